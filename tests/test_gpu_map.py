"""map on one rank on the GPU through the C ABI against the oracle: the wrapped
positions, the staying / lost counts and the per-destination counts bit-exact,
the stay and send buffers (positions and two planes) exactly the oracle's
destination order, the reorder_lbl known answer
(Vector/cuda/vector_dist_cuda_func_test.cu:810-846), empty and wave-edge
sizes."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import remote as R
from test_map import DECOMPS, map_points

pytestmark = pytest.mark.gpu


def _map(dims, rank, bc, dt, n, seed=3):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    slo, shi, sp = R.regular_subdomains(tdt, (0, 0, 0), (1, 1, 1), dims)
    pos = map_points(n, seed, dt, bc)
    dlo, dhi = np.zeros(3, dt), np.ones(3, dt)
    cap = n + 5
    pt = torch.zeros((3, cap), dtype=tdt, device="cuda")
    pt[:, :n] = torch.from_numpy(pos).cuda()
    m = R.Map(tdt, slo, shi, sp, dlo, dhi, bc, rank)
    got = m.label(pt, n)
    want = ora.map_label(pos, n, slo, shi, sp, dlo, dhi, bc, rank)
    return m, pt, got, want


def _check(m, pt, got, want, n):
    w, proc, order, n_stay, stay_begin, n_lost = want
    assert got[0] == n_stay and got[2] == n_lost
    torch.cuda.synchronize()
    assert np.array_equal(pt[:, :n].cpu().numpy(), w)  # wrapped in place
    dest = proc[(proc != m.rank) & (proc != 0xFFFFFFFF)]
    up, uc = np.unique(dest, return_counts=True)
    procs, counts = m.procs()
    assert got[1] == len(up) and np.array_equal(procs, up) and np.array_equal(counts, uc)
    # pack: the staying particles in index order, the others destination-major
    tdt = pt.dtype
    cap = pt.shape[1]
    prop_f = torch.arange(cap, dtype=torch.float32, device="cuda") * 0.25 - 77.0
    prop_l = torch.arange(cap, dtype=torch.int64, device="cuda") * 3 + (1 << 41)
    ns, nsend = m.n_stay, m.n_send
    stay = torch.full((3, ns + 2), -9.0, dtype=tdt, device="cuda")
    send = torch.full((3, nsend + 2), -9.0, dtype=tdt, device="cuda")
    st_f = torch.zeros(ns + 2, dtype=torch.float32, device="cuda")
    st_l = torch.zeros(ns + 2, dtype=torch.int64, device="cuda")
    sd_f = torch.zeros(nsend + 2, dtype=torch.float32, device="cuda")
    sd_l = torch.zeros(nsend + 2, dtype=torch.int64, device="cuda")
    m.pack(pt, stay, send, [(prop_f.data_ptr(), 4), (prop_l.data_ptr(), 8)],
           [(st_f.data_ptr(), 4), (st_l.data_ptr(), 8)], [(sd_f.data_ptr(), 4), (sd_l.data_ptr(), 8)])
    torch.cuda.synchronize()
    o = order.astype(np.int64)
    o_stay = o[stay_begin:stay_begin + n_stay]
    o_send = np.concatenate([o[:stay_begin], o[stay_begin + n_stay:n - n_lost]])
    assert np.array_equal(o_stay, np.flatnonzero(proc == m.rank))
    assert np.array_equal(stay[:, :ns].cpu().numpy(), w[:, o_stay])
    assert np.array_equal(send[:, :nsend].cpu().numpy(), w[:, o_send])
    pf, pl = prop_f.cpu().numpy(), prop_l.cpu().numpy()
    assert np.array_equal(st_f[:ns].cpu().numpy(), pf[o_stay]) and np.array_equal(st_l[:ns].cpu().numpy(), pl[o_stay])
    assert np.array_equal(sd_f[:nsend].cpu().numpy(), pf[o_send])
    assert np.array_equal(sd_l[:nsend].cpu().numpy(), pl[o_send])
    # nothing past the counts is written
    assert np.all(stay[:, ns:].cpu().numpy() == -9.0) and np.all(send[:, nsend:].cpu().numpy() == -9.0)


@pytest.mark.parametrize("dims,rank,bc", DECOMPS)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_map_matches_oracle(dims, rank, bc, dt):
    n = 40000
    print(f"map {dims} rank {rank} bc {bc} {np.dtype(dt).name} n {n}", flush=True)
    m, pt, got, want = _map(dims, rank, bc, dt, n)
    _check(m, pt, got, want, n)
    m.close()


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 4097])
def test_map_sizes(n):
    print(f"map sizes n {n}", flush=True)
    m, pt, got, want = _map((2, 2, 2), 6, (1, 0, 1), np.float32, n, seed=n + 1)
    _check(m, pt, got, want, n)
    m.close()


def test_map_reorder_lbl_known_answer_on_gpu():
    f = np.float64
    slo = np.array([[j / 10, 0, 0] for j in range(10)], f)
    shi = np.array([[(j + 1) / 10, 1, 1] for j in range(10)], f)
    sp = np.arange(10, dtype=np.uint32)
    pos = np.zeros((3, 100), f)
    for i in range(10):
        for j in range(10):
            pos[:, i * 10 + j] = ((j + 0.5) / 10, 0.5, 0.5)
    pt = torch.from_numpy(pos).cuda().contiguous()
    m = R.Map(torch.float64, slo, shi, sp, np.zeros(3), np.ones(3), (1, 1, 1), 4)
    assert m.label(pt, 100) == (10, 9, 0)
    # the particle index as a plane: the packed buffers show the order
    idx = torch.arange(100, dtype=torch.int64, device="cuda")
    st = torch.zeros(10, dtype=torch.int64, device="cuda")
    sd = torch.zeros(90, dtype=torch.int64, device="cuda")
    m.pack(pt, torch.zeros((3, 10), dtype=torch.float64, device="cuda"),
           torch.zeros((3, 90), dtype=torch.float64, device="cuda"), [(idx.data_ptr(), 8)], [(st.data_ptr(), 8)],
           [(sd.data_ptr(), 8)])
    torch.cuda.synchronize()
    full = np.concatenate([sd[:40].cpu().numpy(), st.cpu().numpy(), sd[40:].cpu().numpy()])
    for i in range(10):
        for j in range(10):
            assert int(full[j * 10 + i]) == i * 10 + j
    m.close()


def test_map_rejects_too_many_processors():
    f = np.float64
    nr = 256
    slo = np.array([[j / nr, 0, 0] for j in range(nr)], f)
    shi = np.array([[(j + 1) / nr, 1, 1] for j in range(nr)], f)
    with pytest.raises(RuntimeError):
        R.Map(torch.float64, slo, shi, np.arange(nr, dtype=np.uint32), np.zeros(3), np.ones(3), (1, 1, 1), 0)
