"""Remote ghosts (the send side of ghost_get on one rank), CPU checks: the
library's host-built geometric cell list equals the oracle's, the lists never
lose a box that contains a point, the oracle's label equals brute force over
every box on regular decompositions, the walk's same-pair rule on hand-made
fixtures, and the stable processor order on the reference's
find_buffer_offsets known answer (Vector/cuda/vector_dist_cuda_func_test.cu:
761-808)."""
import numpy as np
import pytest
import torch

from oracle import oracle as ora
from openfpm_pdata_amd import remote as R

DECOMPS = [
    # dims, rank, periodic
    ((2, 2, 2), 0, (1, 1, 1)),
    ((2, 2, 2), 7, (1, 1, 1)),
    ((3, 3, 3), 13, (1, 1, 1)),
    ((3, 3, 3), 0, (1, 1, 1)),
    ((1, 2, 4), 3, (1, 1, 1)),
    ((2, 2, 2), 0, (0, 0, 0)),
    ((4, 1, 1), 1, (1, 0, 1)),
]


def _decomp(dims, rank, bc, dt=torch.float32, ghost=0.07):
    return R.regular_decomposition(dt, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), bc, dims, rank, ghost)


def _points(dlo, dhi, n, seed, dt):
    rng = np.random.default_rng(seed)
    dlo = dlo.astype(np.float64)
    dhi = dhi.astype(np.float64)
    pos = (dlo[:, None] + rng.random((3, n)) * (dhi - dlo)[:, None]).astype(dt)
    # keep them inside the half-open sub-domain after rounding to dt
    for k in range(3):
        pos[k] = np.minimum(np.maximum(pos[k], dlo[k].astype(dt)), np.nextafter(dhi[k].astype(dt), dt(-np.inf)))
    return np.ascontiguousarray(pos)


@pytest.mark.parametrize("dims,rank,bc", DECOMPS)
def test_geo_lists_equal_oracle_and_cover_every_box(dims, rank, bc):
    dlo, dhi, blo, bhi, proc, sid, sh = _decomp(dims, rank, bc)
    g_lib = R.geo(blo.astype(np.float64), bhi.astype(np.float64))
    g_ora = ora.remote_geo(blo, bhi)
    for a, b in zip(g_lib, g_ora):
        assert np.array_equal(a, b)
    glo, cs, div, start, boxes = g_lib
    pos = _points(dlo, dhi, 5000, 3, np.float32).astype(np.float64)
    for i in range(pos.shape[1]):
        x = pos[:, i]
        c = [min(max(int(np.floor((x[k] - glo[k]) / cs[k])), 0), int(div[k]) - 1) for k in range(3)]
        cell = c[0] + int(div[0]) * (c[1] + int(div[1]) * c[2])
        listed = set(boxes[start[cell]:start[cell + 1]].tolist())
        inside = {b for b in range(blo.shape[0]) if np.all(blo[b] <= x) and np.all(x < bhi[b])}
        assert inside <= listed


@pytest.mark.parametrize("dims,rank,bc", DECOMPS)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_label_equals_brute_force_and_ghost_geometry(dims, rank, bc, dt):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    ghost = 0.07
    dlo, dhi, blo, bhi, proc, sid, sh = _decomp(dims, rank, bc, tdt, ghost)
    n = 4000
    pos = _points(dlo, dhi, n, 5, dt)
    po, io = ora.remote_label(pos, n, blo, bhi, proc, sid)
    # one box per (processor, shift): every output is a containing box, each once
    brute = []
    for p in range(n):
        x = pos[:, p]
        for b in range(blo.shape[0]):
            if np.all(blo[b] <= x) and np.all(x < bhi[b]):
                brute.append((int(proc[b]), p, int(sid[b])))
    brute.sort(key=lambda t: t[0])  # stable: particle-major inside a processor
    got = [(int(a), int(b) & 0xFFFFFFFF, int(b) >> 32) for a, b in zip(po, io)]
    assert got == brute
    # x - shift lies in the destination's sub-domain enlarged by the ghost width
    dims_a = np.asarray(dims)
    d = 1.0 / dims_a
    for q, p, s in got[:: max(1, len(got) // 500)]:
        cq = np.array([q % dims[0], (q // dims[0]) % dims[1], q // (dims[0] * dims[1])])
        y = pos[:, p].astype(np.float64) - sh[s].astype(np.float64)
        assert np.all(y >= cq * d - ghost - 1e-6) and np.all(y <= (cq + 1) * d + ghost + 1e-6)


def test_same_pair_rule_fixture():
    """ghost_processorID_general_impl (Decomposition/cuda/ie_ghost_gpu.cuh:49-95):
    after an output, boxes repeating the previous box's (processor, shift) are
    skipped; another pair in between re-arms the output."""
    f = np.float64
    # boxes 0 and 1: processor 1, shift 13, overlapping; box 2: processor 2.
    blo = np.array([[0.0, 0.0, 0.0], [0.4, 0.4, 0.4], [0.9, 0.9, 0.9]], f)
    bhi = np.array([[0.6, 0.6, 0.6], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], f)
    proc = np.array([1, 1, 2], np.uint32)
    sid = np.array([13, 13, 13], np.uint32)
    pos = np.array([[0.5, 0.1, 0.95], [0.5, 0.1, 0.95], [0.5, 0.1, 0.95]], f)
    po, io = ora.remote_label(pos, 3, blo, bhi, proc, sid)
    # particle 0 in boxes 0 and 1 (same pair, consecutive): one output; particle
    # 1 in box 0 only; particle 2 in boxes 1 and 2
    assert [(int(a), int(b)) for a, b in zip(po, io)] == [(1, 0 | 13 << 32), (1, 1 | 13 << 32), (1, 2 | 13 << 32),
                                                          (2, 2 | 13 << 32)]
    # a different pair between two boxes of one pair: the second box emits again.
    # The narrowest box is 0.5 wide, so the grid has 2 cells per dimension and
    # the cell of (0.45, 0.45, 0.45) lists boxes 0, 1, 2; the point is inside 0
    # and 2 (processor 1), not inside 1 (processor 3).
    blo2 = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.4, 0.4, 0.4]], f)
    bhi2 = np.array([[0.6, 0.6, 0.6], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], f)
    proc2 = np.array([1, 3, 1], np.uint32)
    glo, cs, div, start, boxes = ora.remote_geo(blo2, bhi2)
    assert list(div) == [2, 2, 2] and boxes[start[0]:start[1]].tolist() == [0, 1, 2]
    pos2 = np.full((3, 1), 0.45, f)
    po, io = ora.remote_label(pos2, 1, blo2, bhi2, proc2, sid)
    assert [(int(a), int(b)) for a, b in zip(po, io)] == [(1, 0 | 13 << 32), (1, 0 | 13 << 32)]


def test_order_by_proc_find_buffer_offsets_known_answer():
    """vector_dist_gpu_find_buffer_offsets_test: 200,000 outputs, processor
    k / 1000, give 199 boundaries at (i + 1) 1000 -- here 200 processors of 1000
    outputs each, in input order inside a processor (also from shuffled input)."""
    k = np.arange(200000)
    p = (k // 1000).astype(np.uint32)
    po, io = ora.order_by_proc(p, k.astype(np.uint64))
    assert np.array_equal(po, p) and np.array_equal(io, k)
    starts = np.flatnonzero(np.diff(po.astype(np.int64))) + 1
    assert len(starts) == 199 and np.array_equal(starts, (np.arange(199) + 1) * 1000)
    rng = np.random.default_rng(0)
    perm = rng.permutation(k.size)
    po, io = ora.order_by_proc(p[perm], perm.astype(np.uint64))
    assert np.array_equal(po, np.sort(p[perm], kind="stable"))
    for q in (0, 57, 199):
        sel = io[po == q]
        assert np.array_equal(sel, perm[p[perm] == q])  # input order kept


def test_regular_decomposition_box_counts():
    # 2x2x2 periodic: rank 0 meets 3 face neighbours through 2 shifts each, 3 edge
    # neighbours through 4, the corner neighbour through 8
    *_, proc, sid, sh = _decomp((2, 2, 2), 0, (1, 1, 1))
    assert len(proc) == 3 * 2 + 3 * 4 + 8
    # 3x3x3, interior rank: 26 neighbours, one unshifted box each
    *_, proc, sid, sh = _decomp((3, 3, 3), 13, (1, 1, 1))
    assert len(proc) == 26 and np.all(np.all(sh[sid] == 0, axis=1))
    # non-periodic 2x2x2: only unshifted boxes, 7 neighbours
    *_, proc, sid, sh = _decomp((2, 2, 2), 0, (0, 0, 0))
    assert len(proc) == 7 and np.all(np.all(sh[sid] == 0, axis=1))
