"""Remote ghosts on the GPU through the C ABI against the oracle: the label's
ordered outputs and per-processor counts bit-exact, the packed send buffer
exact (positions minus shifts in the dtype, planes copied), the same-pair rule
fixture, empty and wave-edge sizes."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import remote as R

pytestmark = pytest.mark.gpu

DECOMPS = [
    ((2, 2, 2), 0, (1, 1, 1)),
    ((2, 2, 2), 7, (1, 1, 1)),
    ((3, 3, 3), 13, (1, 1, 1)),
    ((3, 3, 3), 0, (1, 1, 1)),
    ((1, 2, 4), 3, (1, 1, 1)),
    ((2, 2, 2), 0, (0, 0, 0)),
]


def _points(dlo, dhi, n, seed, dt, edges=True):
    rng = np.random.default_rng(seed)
    a = dlo.astype(np.float64)
    b = dhi.astype(np.float64)
    pos = (a[:, None] + rng.random((3, n)) * (b - a)[:, None]).astype(dt)
    if edges and n >= 16:
        # some particles exactly on sub-domain and ghost-layer faces
        pos[0, :4] = dlo[0]
        pos[1, 4:8] = dlo[1]
        pos[2, 8:12] = np.nextafter(dhi[2].astype(dt), dt(-np.inf))
    for k in range(3):
        pos[k] = np.minimum(np.maximum(pos[k], dlo[k].astype(dt)), np.nextafter(dhi[k].astype(dt), dt(-np.inf)))
    return np.ascontiguousarray(pos)


def _run(dims, rank, bc, dt, n, ghost=0.07, seed=11):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    dlo, dhi, blo, bhi, proc, sid, sh = R.regular_decomposition(tdt, (0, 0, 0), (1, 1, 1), bc, dims, rank, ghost)
    pos = _points(dlo, dhi, n, seed, dt)
    cap = n + 7
    pt = torch.zeros((3, cap), dtype=tdt, device="cuda")
    pt[:, :n] = torch.from_numpy(pos).cuda()
    g = R.GhostRemote(tdt, blo, bhi, proc, sid, sh)
    n_out, n_procs = g.label(pt, n)
    po, io = ora.remote_label(pos, n, blo, bhi, proc, sid)
    return g, pt, pos, (blo, bhi, proc, sid, sh), n_out, n_procs, po, io


@pytest.mark.parametrize("dims,rank,bc", DECOMPS)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_remote_label_and_pack_match_oracle(dims, rank, bc, dt):
    n = 30000
    print(f"remote label {dims} rank {rank} bc {bc} {np.dtype(dt).name} n {n}", flush=True)
    g, pt, pos, (blo, bhi, proc, sid, sh), n_out, n_procs, po, io = _run(dims, rank, bc, dt, n)
    assert n_out == len(po)
    gp, gi = g.opart()
    torch.cuda.synchronize()
    assert np.array_equal(gp.cpu().numpy().astype(np.uint32), po)
    assert np.array_equal(gi.cpu().numpy().astype(np.uint64), io)
    procs, counts = g.procs()
    up, uc = np.unique(po, return_counts=True)
    assert n_procs == len(up) and np.array_equal(procs, up) and np.array_equal(counts, uc)
    # pack: positions minus shift, and two property planes (4- and 8-byte)
    tdt = pt.dtype
    prop_f = torch.arange(pt.shape[1], dtype=torch.float32, device="cuda") * 0.5 + 12345.0
    prop_l = torch.arange(pt.shape[1], dtype=torch.int64, device="cuda") * 7 + (1 << 40)
    send = torch.full((3, n_out + 3), -1.0, dtype=tdt, device="cuda")
    out_f = torch.zeros(n_out + 3, dtype=torch.float32, device="cuda")
    out_l = torch.zeros(n_out + 3, dtype=torch.int64, device="cuda")
    g.pack(pt, send, [(prop_f.data_ptr(), 4), (prop_l.data_ptr(), 8)], [(out_f.data_ptr(), 4), (out_l.data_ptr(), 8)])
    torch.cuda.synchronize()
    pid = (io & 0xFFFFFFFF).astype(np.int64)
    s = (io >> 32).astype(np.int64)
    want = (pos[:, pid] - sh[s].T.astype(dt)).astype(dt)
    assert np.array_equal(send[:, :n_out].cpu().numpy(), want)
    assert np.array_equal(out_f[:n_out].cpu().numpy(), prop_f.cpu().numpy()[pid])
    assert np.array_equal(out_l[:n_out].cpu().numpy(), prop_l.cpu().numpy()[pid])
    g.close()


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 4097])
def test_remote_sizes(n):
    print(f"remote sizes n {n}", flush=True)
    g, pt, pos, boxes, n_out, n_procs, po, io = _run((2, 2, 2), 0, (1, 1, 1), np.float32, n, ghost=0.2, seed=n)
    assert n_out == len(po)
    if n_out:
        gp, gi = g.opart()
        torch.cuda.synchronize()
        assert np.array_equal(gp.cpu().numpy().astype(np.uint32), po)
        assert np.array_equal(gi.cpu().numpy().astype(np.uint64), io)
    g.close()


def test_remote_same_pair_rule_on_gpu():
    f = np.float64
    blo = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.4, 0.4, 0.4], [0.9, 0.9, 0.9]], f)
    bhi = np.array([[0.6, 0.6, 0.6], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], f)
    proc = np.array([1, 3, 1, 2], np.uint32)
    sid = np.array([13, 13, 13, 0], np.uint32)
    sh = np.zeros((27, 3), f)
    sh[0] = (-1.0, -1.0, -1.0)
    rng = np.random.default_rng(2)
    n = 5000
    pos = np.ascontiguousarray(rng.random((3, n)))
    pos[:, 0] = 0.45
    pt = torch.from_numpy(pos).cuda().contiguous()
    g = R.GhostRemote(torch.float64, blo, bhi, proc, sid, sh)
    n_out, _ = g.label(pt, n)
    po, io = ora.remote_label(pos, n, blo, bhi, proc, sid)
    gp, gi = g.opart()
    torch.cuda.synchronize()
    assert n_out == len(po)
    assert np.array_equal(gp.cpu().numpy().astype(np.uint32), po)
    assert np.array_equal(gi.cpu().numpy().astype(np.uint64), io)
    # particle 0 is in boxes 0 and 2 with box 1 (another processor) between them: two outputs
    assert int(np.sum((io & 0xFFFFFFFF) == 0)) == 2
    g.close()
