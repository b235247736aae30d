"""Python wrapper of the CPU oracle (oracle/ofp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.

Positions are numpy arrays of shape (3, n) (SoA planes, C-contiguous); every
function mirrors the C function of the same name (see ofp_oracle_body.inc for
the reference file:line each one restates).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libofp_oracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def select_native() -> str:
    """Rebuild the oracle with -O3 -march=native for this host (make -B native)
    and load that build from now on (the CPU baseline's timing); returns its path."""
    global _lib, LIB_PATH
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True)
    LIB_PATH = os.path.join(_HERE, "_build", "native", "libofp_oracle.so")
    _lib = None
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.ofo_w3.restype = C.c_double
        _lib.ofo_w3.argtypes = [C.c_double, C.c_double]
        _lib.ofo_dw3.restype = C.c_double
        _lib.ofo_dw3.argtypes = [C.c_double, C.c_double]
        _lib.ofo_ghost_count_f32.restype = C.c_size_t
        _lib.ofo_ghost_count_f64.restype = C.c_size_t
        _lib.ofo_verlet_f32.restype = C.c_size_t
        _lib.ofo_verlet_f64.restype = C.c_size_t
        _lib.ofo_brute_nn_f32.restype = C.c_size_t
        _lib.ofo_brute_nn_f64.restype = C.c_size_t
        _lib.ofo_local_ghost_boxes_f32.restype = C.c_int
        _lib.ofo_local_ghost_boxes_f64.restype = C.c_int
    return _lib


def _sfx(dt) -> str:
    return {np.dtype(np.float32): "_f32", np.dtype(np.float64): "_f64"}[np.dtype(dt)]


def _fn(name, dt):
    return getattr(lib(), name + _sfx(dt))


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _v(vals, dt):
    return np.ascontiguousarray(np.asarray(vals, dtype=dt))


# ------------------------------------------------------------------ geometry
def shift_vectors(lo, hi, dt=np.float64):
    lo, hi = _v(lo, dt), _v(hi, dt)
    out = np.zeros((27, 3), dt)
    _fn("ofo_shift_vectors", dt)(_p(lo), _p(hi), _p(out))
    return out


def local_ghost_boxes(lo, hi, g_lo, g_hi, bc, dt=np.float64):
    lo, hi, g_lo, g_hi = (_v(v, dt) for v in (lo, hi, g_lo, g_hi))
    bc = _v(bc, np.int32)
    blo = np.zeros((26, 3), dt)
    bhi = np.zeros((26, 3), dt)
    sv = np.zeros(26, np.uint32)
    nb = _fn("ofo_local_ghost_boxes", dt)(_p(lo), _p(hi), _p(g_lo), _p(g_hi), _p(bc), _p(blo), _p(bhi), _p(sv))
    return blo[:nb].copy(), bhi[:nb].copy(), sv[:nb].copy()


def magnify(g_lo, g_hi, factor, dt=np.float64):
    g_lo, g_hi = _v(g_lo, dt), _v(g_hi, dt)
    ol = np.zeros(3, dt)
    oh = np.zeros(3, dt)
    f = C.c_float(factor) if np.dtype(dt) == np.float32 else C.c_double(factor)
    _fn("ofo_magnify", dt).argtypes = [C.c_void_p, C.c_void_p, type(f), C.c_void_p, C.c_void_p]
    _fn("ofo_magnify", dt)(_p(g_lo), _p(g_hi), f, _p(ol), _p(oh))
    return ol, oh


def cl_param(lo, hi, r_cut, genl_lo, genl_hi, dt=np.float64):
    lo, hi, gl, gh = (_v(v, dt) for v in (lo, hi, genl_lo, genl_hi))
    ol = np.zeros(3, dt)
    oh = np.zeros(3, dt)
    div = np.zeros(3, np.uint32)
    r = C.c_float(r_cut) if np.dtype(dt) == np.float32 else C.c_double(r_cut)
    fn = _fn("ofo_cl_param", dt)
    fn.argtypes = [C.c_void_p, C.c_void_p, type(r), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    fn(_p(lo), _p(hi), r, _p(gl), _p(gh), _p(ol), _p(oh), _p(div))
    return ol, oh, div


def cl_spacing(lo, hi, div, dt=np.float64):
    lo, hi = _v(lo, dt), _v(hi, dt)
    div = _v(div, np.uint32)
    sp = np.zeros(3, dt)
    _fn("ofo_cl_spacing", dt)(_p(lo), _p(hi), _p(div), _p(sp))
    return sp


class CellGeom:
    """Cell-list geometry of getCellListGPU / getVerlet for a domain and ghost."""

    def __init__(self, dom_lo, dom_hi, ghost, r_cut, factor, dt):
        self.dt = np.dtype(dt)
        g_lo = np.full(3, -ghost, dt)
        g_hi = np.full(3, ghost, dt)
        ml, mh = magnify(g_lo, g_hi, factor, dt)
        self.lo, self.hi, self.div = cl_param(dom_lo, dom_hi, r_cut, ml, mh, dt)
        self.sp = cl_spacing(self.lo, self.hi, self.div, dt)
        self.divc = (self.div + 2).astype(np.uint32)

    @property
    def n_cells(self) -> int:
        return int(np.prod(self.divc.astype(np.int64)))


# --------------------------------------------------------------------- ghosts
def ghost_count(pos, n_dom, blo, bhi, sv):
    dt = pos.dtype
    cnt = np.zeros(max(1, n_dom), np.uint32)
    blo, bhi, sv = _v(blo, dt), _v(bhi, dt), _v(sv, np.uint32)
    tot = _fn("ofo_ghost_count", dt)(_p(pos), C.c_size_t(pos.shape[1]), C.c_size_t(n_dom), _p(blo), _p(bhi),
                                      _p(sv), C.c_int(len(sv)), _p(cnt))
    return cnt[:n_dom], int(tot)


def ghost_emit(pos, n_dom, blo, bhi, sv, shifts):
    dt = pos.dtype
    _, tot = ghost_count(pos, n_dom, blo, bhi, sv)
    blo, bhi, sv, shifts = _v(blo, dt), _v(bhi, dt), _v(sv, np.uint32), _v(shifts, dt)
    g = np.zeros((3, max(1, tot)), dt)
    op = np.zeros(max(1, tot), np.uint32)
    os_ = np.zeros(max(1, tot), np.uint32)
    _fn("ofo_ghost_emit", dt)(_p(pos), C.c_size_t(pos.shape[1]), C.c_size_t(n_dom), _p(blo), _p(bhi), _p(sv),
                              C.c_int(len(sv)), _p(shifts), _p(g), C.c_size_t(g.shape[1]), _p(op), _p(os_))
    return g[:, :tot], op[:tot], os_[:tot]


def apply_bc(pos, lo, hi, bc):
    dt = pos.dtype
    out = np.ascontiguousarray(pos.copy())
    lo, hi, bc = _v(lo, dt), _v(hi, dt), _v(bc, np.int32)
    _fn("ofo_apply_bc", dt)(_p(out), C.c_size_t(out.shape[1]), C.c_size_t(out.shape[1]), _p(lo), _p(hi), _p(bc))
    return out


# ------------------------------------------------------------------ cell list
def cell_ids(pos, geom: CellGeom):
    n = pos.shape[1]
    cid = np.zeros(max(1, n), np.uint32)
    _fn("ofo_cell_ids", pos.dtype)(_p(pos), C.c_size_t(n), C.c_size_t(n), _p(geom.lo), _p(geom.sp),
                                   _p(geom.divc), _p(cid))
    return cid[:n]


def cl_fill(pos, geom: CellGeom):
    n = pos.shape[1]
    cs = np.zeros(geom.n_cells + 1, np.uint32)
    ci = np.zeros(max(1, n), np.uint32)
    _fn("ofo_cl_fill", pos.dtype)(_p(pos), C.c_size_t(n), C.c_size_t(n), _p(geom.lo), _p(geom.sp),
                                  _p(geom.divc), _p(cs), _p(ci))
    return cs, ci[:n]


def verlet(pos, n_dom, geom: CellGeom, r_cut):
    n = pos.shape[1]
    dt = pos.dtype
    off = np.zeros(n_dom + 1, np.uint32)
    r = C.c_float(r_cut) if dt == np.float32 else C.c_double(r_cut)
    fn = _fn("ofo_verlet", dt)
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, type(r),
                   C.c_void_p, C.c_void_p, C.c_size_t]
    E = fn(_p(pos), n, n, n_dom, _p(geom.lo), _p(geom.sp), _p(geom.divc), r, _p(off), None, 0)
    nbr = np.zeros(max(1, E), np.uint32)
    fn(_p(pos), n, n, n_dom, _p(geom.lo), _p(geom.sp), _p(geom.divc), r, _p(off), _p(nbr), E)
    return off, nbr[:E]


def brute_nn(pos, n_dom, r_cut):
    n = pos.shape[1]
    dt = pos.dtype
    off = np.zeros(n_dom + 1, np.uint32)
    r = C.c_float(r_cut) if dt == np.float32 else C.c_double(r_cut)
    fn = _fn("ofo_brute_nn", dt)
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, type(r), C.c_void_p, C.c_void_p, C.c_size_t]
    E = fn(_p(pos), n, n, n_dom, r, _p(off), None, 0)
    nbr = np.zeros(max(1, E), np.uint32)
    fn(_p(pos), n, n, n_dom, r, _p(off), _p(nbr), E)
    return off, nbr[:E]


# ---------------------------------------------------------------------- pairs
def _pair_args(fn, extra):
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p] + extra


def lj(pos, n_dom, geom: CellGeom, r_cut, sigma, eps):
    n = pos.shape[1]
    F = np.zeros((max(1, n_dom), 3), np.float64)
    U = np.zeros(max(1, n_dom), np.float64)
    nn = np.zeros(max(1, n_dom), np.uint32)
    fn = _fn("ofo_lj", pos.dtype)
    _pair_args(fn, [C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p])
    fn(_p(pos), n, n, n_dom, _p(geom.lo), _p(geom.sp), _p(geom.divc), r_cut, sigma, eps, _p(F), _p(U), _p(nn))
    return F[:n_dom], U[:n_dom], nn[:n_dom]


def lj_walk(pos, targets, geom, cs, ci, r_cut, sigma, eps):
    """ofo_lj for the particles `targets` on a prebuilt cell list (cl_fill's cs,
    ci over all of pos): (forces (t, 3), energies, neighbour counts)."""
    t = np.ascontiguousarray(targets, np.uint32)
    m = len(t)
    F = np.zeros((max(1, m), 3), np.float64)
    U = np.zeros(max(1, m), np.float64)
    nn = np.zeros(max(1, m), np.uint32)
    fn = _fn("ofo_lj_walk", pos.dtype)
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    fn(_p(pos), pos.shape[1], _p(t), m, _p(geom.lo), _p(geom.sp), _p(geom.divc), _p(cs), _p(ci), r_cut, sigma, eps,
       _p(F), _p(U), _p(nn))
    return F[:m], U[:m], nn[:m]


def sph_density(pos, n_dom, geom: CellGeom, h, mass):
    n = pos.shape[1]
    rho = np.zeros(max(1, n_dom), np.float64)
    fn = _fn("ofo_sph_density", pos.dtype)
    _pair_args(fn, [C.c_double, C.c_double, C.c_void_p])
    fn(_p(pos), n, n, n_dom, _p(geom.lo), _p(geom.sp), _p(geom.divc), h, mass, _p(rho))
    return rho[:n_dom]


def tait(rho, rho0, c0):
    rho = np.ascontiguousarray(rho, np.float64)
    prs = np.zeros_like(rho)
    lib().ofo_tait.argtypes = [C.c_void_p, C.c_size_t, C.c_double, C.c_double, C.c_void_p]
    lib().ofo_tait(_p(rho), rho.size, rho0, c0, _p(prs))
    return prs


def sph_force(pos, n_dom, geom: CellGeom, h, mass, rho_all, prs_all):
    n = pos.shape[1]
    F = np.zeros((max(1, n_dom), 3), np.float64)
    rho_all = np.ascontiguousarray(rho_all, np.float64)
    prs_all = np.ascontiguousarray(prs_all, np.float64)
    fn = _fn("ofo_sph_force", pos.dtype)
    _pair_args(fn, [C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p])
    fn(_p(pos), n, n, n_dom, _p(geom.lo), _p(geom.sp), _p(geom.divc), h, mass, _p(rho_all), _p(prs_all), _p(F))
    return F[:n_dom]


def w3(r, h) -> float:
    return lib().ofo_w3(r, h)


def dw3(r, h) -> float:
    return lib().ofo_dw3(r, h)


# ------------------------------------------------------------ remote ghosts
def remote_geo(box_lo, box_hi):
    """Geometric cell list of the internal ghost boxes (ofp_oracle_remote.c):
    (grid_lo, cell_size, div, cell_start, cell_boxes)."""
    blo = np.ascontiguousarray(box_lo, np.float64).reshape(-1, 3)
    bhi = np.ascontiguousarray(box_hi, np.float64).reshape(-1, 3)
    f = lib().ofo_remote_geo
    f.restype = C.c_size_t
    f.argtypes = [C.c_int] + [C.c_void_p] * 7
    glo, cs, div = np.zeros(3), np.zeros(3), np.zeros(3, np.uint32)
    m = f(blo.shape[0], _p(blo), _p(bhi), _p(glo), _p(cs), _p(div), None, None)
    start = np.zeros(int(np.prod(div.astype(np.int64))) + 1, np.uint32)
    boxes = np.zeros(max(1, m), np.uint32)
    f(blo.shape[0], _p(blo), _p(bhi), _p(glo), _p(cs), _p(div), _p(start), _p(boxes))
    return glo, cs, div, start, boxes[:m]


def order_by_proc(proc, ids):
    """Stable order of (processor, id) outputs by processor."""
    proc = np.ascontiguousarray(proc, np.uint32)
    ids = np.ascontiguousarray(ids, np.uint64)
    po = np.zeros(max(1, proc.size), np.uint32)
    io = np.zeros(max(1, proc.size), np.uint64)
    f = lib().ofo_order_by_proc
    f.argtypes = [C.c_size_t] + [C.c_void_p] * 4
    f(proc.size, _p(proc), _p(ids), _p(po), _p(io))
    return po[: proc.size], io[: proc.size]


def remote_label(pos, n, box_lo, box_hi, proc, shift_id, geo=None):
    """Outputs of the label walk for particles [0, n) of pos (3, cap), ordered by
    processor: (processor, particle | shift << 32)."""
    pos = np.ascontiguousarray(pos)
    dt = pos.dtype
    blo = np.ascontiguousarray(box_lo, dt).reshape(-1, 3)
    bhi = np.ascontiguousarray(box_hi, dt).reshape(-1, 3)
    proc = np.ascontiguousarray(proc, np.uint32)
    sid = np.ascontiguousarray(shift_id, np.uint32)
    glo, cs, div, start, boxes = geo if geo is not None else remote_geo(blo.astype(np.float64), bhi.astype(np.float64))
    boxes = np.ascontiguousarray(boxes, np.uint32) if boxes.size else np.zeros(1, np.uint32)
    f = _fn("ofo_remote_label", dt)
    f.restype = C.c_size_t
    f.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t] + [C.c_void_p] * 11
    args = [_p(pos), pos.shape[1], n, _p(blo), _p(bhi), _p(proc), _p(sid), _p(glo), _p(cs), _p(div), _p(start),
            _p(boxes)]
    m = f(*args, None, None)
    po = np.zeros(max(1, m), np.uint32)
    io = np.zeros(max(1, m), np.uint64)
    f(*args, _p(po), _p(io))
    return po[:m], io[:m]


def map_label(pos, n, sub_lo, sub_hi, sub_proc, dom_lo, dom_hi, bc, rank):
    """map on one rank for particles [0, n) of pos (3, cap): the wrapped
    positions (ofo_apply_bc), the processor of each particle (ofo_map_proc;
    0xFFFFFFFF for none) and the particle order by destination (stable,
    processor ascending, the lost last).  Returns (wrapped pos, proc, order,
    n_stay, stay_begin, n_lost)."""
    pos = np.ascontiguousarray(pos)
    dt = pos.dtype
    w = apply_bc(pos[:, :n], dom_lo, dom_hi, bc) if n else pos[:, :0].copy()
    blo = np.ascontiguousarray(sub_lo, dt).reshape(-1, 3)
    bhi = np.ascontiguousarray(sub_hi, dt).reshape(-1, 3)
    sp = np.ascontiguousarray(sub_proc, np.uint32)
    glo, cs, div, start, boxes = remote_geo(blo.astype(np.float64), bhi.astype(np.float64))
    boxes = np.ascontiguousarray(boxes, np.uint32) if boxes.size else np.zeros(1, np.uint32)
    proc = np.zeros(max(1, n), np.uint32)
    f = _fn("ofo_map_proc", dt)
    f.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t] + [C.c_void_p] * 11
    f(_p(w), w.shape[1], n, _p(blo), _p(bhi), _p(sp), _p(_v(dom_hi, dt)), _p(_v(bc, np.int32)), _p(glo), _p(cs),
      _p(div), _p(start), _p(boxes), _p(proc))
    proc = proc[:n]
    order = np.argsort(proc, kind="stable").astype(np.uint64)
    sel = proc == np.uint32(rank)
    n_stay = int(sel.sum())
    stay_begin = int(np.sum(proc < np.uint32(rank)))
    n_lost = int(np.sum(proc == np.uint32(0xFFFFFFFF)))
    return w, proc, order, n_stay, stay_begin, n_lost


# ------------------------------------------------------------ symmetric lists
def verlet_sym(off, nbr, row_key, key, ghost_marker, opid, osid, s2n=None):
    """Symmetric half of a full list (getVerletSym's role; the rule of
    ofp_vl_build_sym): row r (particle key row_key[r]) keeps a domain entry j iff
    key[j] > row_key[r], and a ghost entry (ghost k = s2n[j] - ghost_marker of
    source particle opid[k], shift osid[k]) iff key of the source > row_key[r],
    or the source is the row's own particle and osid[k] > 13.  Entries index the
    particle arrays (sorted index when s2n maps them to original indices, else
    original indices); key maps an original index to its place in the total
    order.  Returns (off_sym, nbr_sym)."""
    off = np.asarray(off, np.int64)
    nbr = np.asarray(nbr, np.int64)
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    o = nbr if s2n is None else np.asarray(s2n, np.int64)[nbr]
    ghost = o >= ghost_marker
    k = np.where(ghost, o - ghost_marker, 0)
    src = np.where(ghost, np.asarray(opid, np.int64)[k], o)
    ks = np.asarray(key, np.int64)[src]
    kr = np.asarray(row_key, np.int64)[rows]
    keep = (ks > kr) | (ghost & (ks == kr) & (np.asarray(osid, np.int64)[k] > 13))
    cnt = np.bincount(rows[keep], minlength=len(off) - 1)
    off_sym = np.zeros(len(off), np.int64)
    np.cumsum(cnt, out=off_sym[1:])
    return off_sym, nbr[keep]


def ghost_put_add(vals, ghost_marker, opid, dst=None):
    """ghost_put<add_> of the local ghosts (merge_prp_v with add_, Vector/
    vector_dist_comm.hpp:2158-2180): in ghost order, value of ghost k is added to
    its source opid[k].  vals: original order (dst None) or any order with dst
    mapping original index -> position (the ghost k at dst[ghost_marker + k])."""
    out = np.array(vals, copy=True)
    opid = np.asarray(opid, np.int64)
    g = np.arange(len(opid)) + ghost_marker
    if dst is None:
        np.add.at(out, opid, out[g])
    else:
        dst = np.asarray(dst, np.int64)
        np.add.at(out, dst[opid], out[dst[g]])
    return out


# ---------------------------------------------------------- sparse cell list
def cl_sparse(cell_start, divc, box_nn=1):
    """Sparse form of a dense cell list (CELLLIST_GPU_SPARSE, vector_dist.hpp:49):
    the non-empty cells in ascending id, their starts (+ the total), and for
    each the (2B + 1)^3 cells of its stencil, B = box_nn (x fastest over dx, dy,
    dz in -B..B), as indices among the non-empty cells, -1 where empty or off
    the grid."""
    cs = np.asarray(cell_start, np.int64)
    d0, d1, d2 = (int(x) for x in divc)
    key = np.flatnonzero(cs[1:] > cs[:-1])
    start = np.append(cs[key], cs[-1])
    inv = np.full(len(cs) - 1, -1, np.int64)
    inv[key] = np.arange(len(key))
    cx, cy, cz = key % d0, (key // d0) % d1, key // (d0 * d1)
    rng = range(-box_nn, box_nn + 1)
    nn = np.full((len(key), len(rng) ** 3), -1, np.int64)
    k = 0
    for dz in rng:
        for dy in rng:
            for dx in rng:
                x, y, z = cx + dx, cy + dy, cz + dz
                ok = (x >= 0) & (y >= 0) & (z >= 0) & (x < d0) & (y < d1) & (z < d2)
                c = np.where(ok, x + d0 * (y + d1 * z), 0)
                nn[:, k] = np.where(ok, inv[c], -1)
                k += 1
    return key, start, nn
