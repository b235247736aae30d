"""Seeded synthetic particle sets for the neighbour-search workloads.

Every generator returns positions as a numpy array of shape (3, n) (SoA) plus
the physical parameters of the workload; the same seed always gives the same
particles.  Sizes can be scaled down (``n=...``) for parity tests; the recipe
(density, cut-off relative to spacing, cluster widths) is kept.

Workloads (3-D unit cube unless stated):
  uniform_lj    65,536 FP64 uniform, periodic; r_c = 0.05 = cell size; LJ
                sigma = 0.02, eps = 1, truncated at r_c; ghost width r_c.
  clustered_sph 4,194,304 FP32, periodic; half uniform, half from 64 Gaussian
                clusters (sigma 0.02, wrapped); SPH cubic spline with
                h = 1.3 dx, dx = N^(-1/3), r_c = 2h; Verlet skin 10 %;
                Tait EOS rho0 = 1000, c0 = 20 sqrt(g H), H = 0.5 (the EOS of
                the water column, reused here); mass = rho0 / N.
  water_column  256^3 = 16,777,216 particles on a cubic lattice of spacing
                dx = H / 256 (H = 0.5, a cube of side H), jitter U(-0.01dx,
                0.01dx) per axis; non-periodic box [-2h, H + 2h]^3;
                h = 1.3 dx; Tait rho0 = 1000, c0 = 20 sqrt(g H); mass =
                rho0 dx^3.
  ensemble      256 independent periodic unit boxes, N_b = round(exp(U(ln
                1024, ln 262144))), r_c ~ U(0.03, 0.08), sigma = 0.4 r_c,
                uniform FP32, LJ.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81
H_COL = 0.5
C0 = 20.0 * math.sqrt(G * H_COL)
RHO0 = 1000.0

SEED_BASE = 0x0F9D0000


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


@dataclass
class Workload:
    name: str
    pos: np.ndarray  # (3, n)
    box_lo: tuple
    box_hi: tuple
    bc: tuple
    ghost: float
    r_cut: float
    params: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[1])


def uniform_lj(n: int = 65536, seed: int = SEED_BASE + 1, dtype=np.float64, r_cut: float = 0.05,
               sigma: float = 0.02, eps: float = 1.0) -> Workload:
    rng = _rng(seed)
    pos = rng.random((3, n), dtype=np.float64).astype(dtype)
    return Workload("uniform_lj", np.ascontiguousarray(pos), (0.0,) * 3, (1.0,) * 3, (1, 1, 1), r_cut, r_cut,
                    dict(sigma=sigma, eps=eps))


def clustered_sph(n: int = 4194304, seed: int = SEED_BASE + 2, dtype=np.float32, n_clusters: int = 64,
                  cluster_sigma: float = 0.02, skin: float = 0.10) -> Workload:
    rng = _rng(seed)
    nu = n // 2
    nc = n - nu
    uni = rng.random((3, nu))
    centres = rng.random((3, n_clusters))
    which = rng.integers(0, n_clusters, nc)
    clu = centres[:, which] + rng.normal(0.0, cluster_sigma, (3, nc))
    clu = np.mod(clu, 1.0)
    pos = np.concatenate([uni, clu], axis=1).astype(dtype)
    # values that round up to 1.0 in the target dtype wrap to 0
    pos[pos >= 1.0] = 0.0
    dx = n ** (-1.0 / 3.0)
    h = 1.3 * dx
    r_cut = 2.0 * h
    return Workload("clustered_sph", np.ascontiguousarray(pos), (0.0,) * 3, (1.0,) * 3, (1, 1, 1), r_cut, r_cut,
                    dict(h=h, dx=dx, mass=RHO0 / n, rho0=RHO0, c0=C0, skin=skin, r_verlet=r_cut * (1 + skin)))


def water_column(side: int = 256, seed: int = SEED_BASE + 3, dtype=np.float64, jitter: float = 0.01) -> Workload:
    rng = _rng(seed)
    dx = H_COL / side
    h = 1.3 * dx
    ax = (np.arange(side) + 0.5) * dx
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()])
    pos = pos + rng.uniform(-jitter * dx, jitter * dx, pos.shape)
    pad = 2 * h
    return Workload("water_column", np.ascontiguousarray(pos.astype(dtype)), (-pad,) * 3, (H_COL + pad,) * 3,
                    (0, 0, 0), 2 * h, 2 * h, dict(h=h, dx=dx, mass=RHO0 * dx ** 3, rho0=RHO0, c0=C0))


def ensemble(n_boxes: int = 256, seed: int = SEED_BASE + 4, n_min: int = 1024, n_max: int = 262144,
             dtype=np.float32):
    rng = _rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_boxes))).astype(np.int64)
    rcs = rng.uniform(0.03, 0.08, n_boxes)
    out = []
    for b in range(n_boxes):
        pos = rng.random((3, int(sizes[b]))).astype(dtype)
        pos[pos >= 1.0] = 0.0
        rc = float(rcs[b])
        out.append(Workload(f"box{b}", np.ascontiguousarray(pos), (0.0,) * 3, (1.0,) * 3, (1, 1, 1), rc, rc,
                            dict(sigma=0.4 * rc, eps=1.0)))
    return out


def power_law_clusters(n: int = 1 << 26, seed: int = SEED_BASE + 5, n_clusters: int = 512, width1: float = 0.1,
                       target_neighbours: float = 32.0, dtype=np.float32) -> Workload:
    """Dynamic-run particles: n FP32 particles in the periodic unit cube from
    n_clusters Gaussian clusters of weights w_k ~ k^-1.5, centres uniform, widths
    sigma_k = width1 (w_k / w_1)^(1/3) (equal peak density), wrapped.  r_c sets
    the particle-weighted mean neighbour count to target_neighbours
    (<rho> = n w_1 / (8 pi^1.5 width1^3), clusters treated as isolated);
    LJ sigma = 0.4 r_c, eps = 1, m = 1; Maxwell velocities at T = 1 (unit
    normal per component, mean removed); dt = 1e-4."""
    rng = _rng(seed)
    k = np.arange(1, n_clusters + 1, dtype=np.float64)
    w = k ** -1.5
    w /= w.sum()
    centres = rng.random((n_clusters, 3))
    widths = width1 * (w / w[0]) ** (1.0 / 3.0)
    which = np.searchsorted(np.cumsum(w), rng.random(n), side="right")
    which = np.minimum(which, n_clusters - 1)
    pos = centres[which].T + rng.standard_normal((3, n)) * widths[which]
    pos = np.mod(pos, 1.0).astype(dtype)
    pos[pos >= 1.0] = 0.0
    rho_w = n * w[0] / (8.0 * math.pi ** 1.5 * width1 ** 3)
    rc = (target_neighbours / (4.0 / 3.0 * math.pi * rho_w)) ** (1.0 / 3.0)
    vel = rng.standard_normal((3, n))
    vel -= vel.mean(axis=1, keepdims=True)
    return Workload("power_law_clusters", np.ascontiguousarray(pos), (0.0,) * 3, (1.0,) * 3, (1, 1, 1), rc, rc,
                    dict(sigma=0.4 * rc, eps=1.0, mass=1.0, dt=1e-4, vel=np.ascontiguousarray(vel.astype(dtype)),
                         rho_w=rho_w))


def power_law_clusters_lj(n: int = 1 << 26, seed: int = SEED_BASE + 6, n_clusters: int = 512,
                          mean_density: float = 0.05, peak_density: float = 0.4, d_min: float = 0.85,
                          max_rounds: int = 60):
    """The dynamic run in Lennard-Jones units (sigma = 1, eps = 1, m = 1, so
    dt = 1e-4 is 1e-4 tau and T = 1 is 1 eps): n FP32 particles in a periodic box
    centred on 0 of side L = (n / mean_density)^(1/3), from n_clusters Gaussian
    clusters of weights w_k ~ k^-1.5, centres uniform, widths s_k = s_1 (w_k /
    w_1)^(1/3) with s_1 set so that every cluster centre has the density
    peak_density = n w_k / ((2 pi)^1.5 s_k^3) (8x the mean; a dilute gas, so
    particles can be placed without overlap), r_c = 2.5 sigma.  Where two
    particles are closer than d_min sigma the one of higher index is redrawn from
    its cluster, until no such pair is left (the neighbour search is this
    library's, on the GPU).  Returns (Workload, torch positions [3, n], torch
    velocities [3, n], rounds, pairs still closer than d_min)."""
    import torch

    from . import vector_dist as vdm

    g = torch.Generator(device="cuda").manual_seed(seed)
    # an FP32 side: +-L/2 are then exact in FP32 and rounding keeps particles in [lo, hi]
    L = float(np.float32((n / mean_density) ** (1.0 / 3.0)))
    k = torch.arange(1, n_clusters + 1, dtype=torch.float64, device="cuda")
    w = k ** -1.5
    w /= w.sum()
    centres = (torch.rand((n_clusters, 3), generator=g, device="cuda", dtype=torch.float64) - 0.5) * L
    s1 = (n * float(w[0]) / ((2.0 * math.pi) ** 1.5 * peak_density)) ** (1.0 / 3.0)
    widths = s1 * (w / w[0]) ** (1.0 / 3.0)
    which = torch.searchsorted(torch.cumsum(w, 0), torch.rand(n, generator=g, device="cuda", dtype=torch.float64))
    which = torch.clamp(which, max=n_clusters - 1)
    lo, hi = -0.5 * L, 0.5 * L

    def draw(idx):
        c = which[idx]
        p = centres[c].T + torch.randn((3, idx.numel()), generator=g, device="cuda", dtype=torch.float64) * widths[c]
        p = torch.remainder(p - lo, L) + lo
        p = p.to(torch.float32)
        return torch.where(p >= hi, torch.full_like(p, lo), p)

    pos = draw(torch.arange(n, device="cuda"))
    rounds = 0
    n_close = -1
    for rounds in range(1, max_rounds + 1):
        # cells of 2.5 sigma (cells of d_min over a box of ~1100 sigma would be ~2e9)
        vd = vdm.VectorDistGPU(n, (lo,) * 3, (hi,) * 3, (1, 1, 1), d_min, dtype=torch.float32,
                               capacity=int(n * 1.05) + 4096)
        vd.pos[:, :n] = pos
        vd.ghost_get()
        cl = vd.updateCellListGPU(vd.getCellListGPU(2.5, vdm.CL_NON_SYMMETRIC | vdm.CL_GPU_REORDER_POSITION))
        vl = vdm.VerletListGPU()
        vl.build(cl, vd.pos, d_min)
        off, nbr = vl.arrays()
        a = cl.arrays()
        cnt = off[1:] - off[:-1]
        rows = (cnt > 1).nonzero().flatten()  # a neighbour besides itself
        bad = torch.empty(0, dtype=torch.long, device="cuda")
        n_close = 0
        if rows.numel():
            # particle of a storage index: itself, or the source of a ghost
            gp, _ = vd.ghost_opart()
            src = torch.cat([torch.arange(n, device="cuda"), gp.long()])
            s2n = a["s2n"].long()
            me = src[s2n[a["dom"][rows].long()]]
            c = cnt[rows]
            # the entries of those rows
            first = torch.repeat_interleave(off[:-1][rows] - (torch.cumsum(c, 0) - c), c)
            ent = torch.arange(int(c.sum()), device="cuda") + first
            mine = torch.repeat_interleave(me, c)
            other = src[s2n[nbr[ent].long()]]
            later = other < mine  # the pair's member of higher index is redrawn
            n_close = int(later.sum())
            bad = torch.unique(mine[later])
            del gp, src, s2n, me, c, first, ent, mine, other, later
        del vd, cl, vl, a, off, nbr, cnt, rows
        if bad.numel() == 0:
            break
        pos[:, bad] = draw(bad)
    vel = torch.randn((3, n), generator=g, device="cuda", dtype=torch.float64)
    vel -= vel.mean(dim=1, keepdim=True)
    wl = Workload("power_law_clusters_lj", np.zeros((3, 0), np.float32), (lo,) * 3, (hi,) * 3, (1, 1, 1), 2.5, 2.5,
                  dict(sigma=1.0, eps=1.0, mass=1.0, dt=1e-4, n=n, box=L, d_min=d_min, peak_density=peak_density))
    return wl, pos, vel.to(torch.float32), rounds, n_close


def lattice(ng: int, dtype=np.float32) -> np.ndarray:
    """The reference lattice test's particles (vector_dist_unit_test.cpp:1232-1247):
    key * spacing with spacing = 1/Ng computed in the dtype."""
    sp = dtype(1.0) / dtype(ng)
    k = np.arange(ng)
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()]).astype(dtype) * sp
    return np.ascontiguousarray(pos.astype(dtype))
