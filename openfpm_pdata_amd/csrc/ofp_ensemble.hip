// ofp_ensemble.hip -- many independent periodic systems driven by one set of
// launches: per-system ghosts from one count/scan/emit, one radix-sorted cell
// list over the grids of all systems (global cell ids), one Verlet build and one
// LJ pass over device-side work queues, velocity Verlet, and hipGraph capture of
// the rebuild and of the step.
//
// Device-resident counts: the ghost total, the Verlet entry total and the
// rebuild decision never leave the device inside a captured graph.  Capacities
// are fixed by a sizing rebuild (which reads the counts back); a captured rebuild
// that would overflow writes nothing past them and raises a flag the host reads
// after it, then redoes the rebuild with sizing.
//
// Storage order: at every rebuild after the first the stored particles are
// permuted into the previous build's domain-row order (cell order), so the next
// sort reads nearly sorted input and the per-step reorder/restore gathers are
// nearly sequential.  gid keeps the caller's order.
#include "ofp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ofp {

enum : uint32_t { kOvfGhost = 1u, kOvfVerlet = 2u };

// periodic_l: one wrap, as the reference's map.  A particle that moved more than
// a box length since the last wrap stays outside the box (its cell is clamped
// into the padding layer); reducing it by floor((x - lo) / L) would, once the
// ulp of x exceeds the box, collapse such particles onto the same points.
template <class T>
__device__ __forceinline__ T wrap1(T x, T lo, T hi, T L) {
#pragma clang fp contract(off)
  if (x >= hi) return x - L;
  if (x < lo) return x + L;
  return x;
}

// map(): wrap into the box; xb = reference positions of the displacement test.
// With s2n != null the stored particles are also permuted into the previous
// build's domain-row order, written to the *2 arrays, which then become the
// stored arrays (the host swaps the two sets).
template <class T>
__global__ __launch_bounds__(256) void k_ens_map(uint32_t N, size_t cap, const uint32_t* __restrict__ dom,
                                                 const uint32_t* __restrict__ s2n, T* __restrict__ pos,
                                                 const T* __restrict__ vel, const T* __restrict__ frc,
                                                 const uint32_t* __restrict__ gid, T* __restrict__ pos2,
                                                 T* __restrict__ vel2, T* __restrict__ frc2,
                                                 uint32_t* __restrict__ gid2, T* __restrict__ xb,
                                                 const uint32_t* __restrict__ sid, const SysDyn<T>* __restrict__ dyn) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const SysDyn<T>& D = dyn[sid[k]];  // rows and stored particles share the system ranges
  if (s2n) {
    const uint32_t i = s2n[dom[k]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T x = wrap1(pos[d * cap + i], D.lo[d], D.hi[d], D.L[d]);
      pos2[d * cap + k] = x;
      xb[d * cap + k] = x;
      vel2[d * cap + k] = vel[d * cap + i];
      if (frc) frc2[d * cap + k] = frc[d * cap + i];
    }
    gid2[k] = gid[i];
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T x = wrap1(pos[d * cap + k], D.lo[d], D.hi[d], D.L[d]);
      pos[d * cap + k] = x;
      xb[d * cap + k] = x;
    }
  }
}

// Ghost count / emit of the one-system path with the boxes of each particle's
// system; ghosts land at N + (scan of the counts), at most gcap of them.
template <class T>
__global__ __launch_bounds__(256) void k_ens_ghost_count(const GhostBoxes<T>* __restrict__ gb,
                                                         const T* __restrict__ pos, size_t cap, uint32_t N,
                                                         const uint32_t* __restrict__ sid,
                                                         uint32_t* __restrict__ cnt) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const T x = pos[p], y = pos[cap + p], z = pos[2 * cap + p];
  auto count = [&](const GhostBoxes<T>& B) {
    uint32_t old = 0xFFFFFFFFu, n = 0;
    if (!no_box(B, x, y, z)) {
      for (int b = 0; b < B.nb; ++b) {
        if (inside_np(B, b, x, y, z) && B.sv[b] != old) {
          old = B.sv[b];
          ++n;
        }
      }
    }
    return n;
  };
  // one system across the wave (the usual case): its boxes by scalar loads
  const uint32_t sy = sid[p];
  const uint32_t s0 = __builtin_amdgcn_readfirstlane(sy);
  cnt[p] = __all(sy == s0) ? count(gb[s0]) : count(gb[sy]);
}

template <class T>
__global__ __launch_bounds__(256) void k_ens_ghost_emit(const GhostBoxes<T>* __restrict__ gb, T* __restrict__ pos,
                                                        size_t cap, uint32_t N, uint32_t* __restrict__ sid,
                                                        const uint32_t* __restrict__ gstart, uint32_t gcap,
                                                        uint32_t* __restrict__ opid, uint32_t* __restrict__ osid) {
#pragma clang fp contract(off)
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t sy = sid[p];
  const T x = pos[p], y = pos[cap + p], z = pos[2 * cap + p];
  auto emit = [&](const GhostBoxes<T>& B) {
    if (no_box(B, x, y, z)) return;
    const uint32_t base = gstart[p];
    uint32_t old = 0xFFFFFFFFu, n = 0;
    for (int b = 0; b < B.nb; ++b) {
      if (inside_np(B, b, x, y, z) && B.sv[b] != old) {
        const uint32_t sh = B.sv[b];
        const uint32_t g = base + n;
        if (g < gcap) {
          pos[N + g] = x - B.shift[sh][0];
          pos[cap + N + g] = y - B.shift[sh][1];
          pos[2 * cap + N + g] = z - B.shift[sh][2];
          opid[g] = p;
          osid[g] = sh;
          sid[N + g] = sy;
        }
        old = sh;
        ++n;
      }
    }
  };
  // one system across the wave (the usual case): its boxes by scalar loads
  const uint32_t s0 = __builtin_amdgcn_readfirstlane(sy);
  if (__all(sy == s0))
    emit(gb[s0]);
  else
    emit(gb[sy]);
}

// counts[0] = particles binned (N + ghosts kept), counts[1] = ghosts labelled.
__global__ void k_ens_counts(uint32_t N, const uint32_t* __restrict__ gstart, uint32_t gcap,
                             uint32_t* __restrict__ counts, uint32_t* __restrict__ flags) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t G = gstart[N];
  counts[1] = G;
  counts[0] = N + (G < gcap ? G : gcap);
  if (G > gcap) atomicOr(&flags[1], kOvfGhost);
}

__global__ void k_ens_check_entries(const unsigned long long* __restrict__ e_total, unsigned long long e_cap,
                                    const unsigned long long* __restrict__ mstate, uint32_t* __restrict__ flags,
                                    unsigned long long* __restrict__ e_out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const unsigned long long E = *e_total;
  *e_out = E;
  if (E > e_cap || mstate[1] != 0ull) atomicOr(&flags[1], kOvfVerlet);
}

// Per step: ghost positions from their sources (SKIP_LABELLING refresh).
template <class T>
__global__ __launch_bounds__(256) void k_ens_ghost_refresh(const GhostBoxes<T>* __restrict__ gb, T* __restrict__ pos,
                                                           size_t cap, uint32_t N, const uint32_t* __restrict__ counts,
                                                           const uint32_t* __restrict__ sid,
                                                           const uint32_t* __restrict__ opid,
                                                           const uint32_t* __restrict__ osid) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (N + i >= counts[0]) return;
  const uint32_t p = opid[i];
  const GhostBoxes<T>& B = gb[sid[p]];
  const uint32_t sh = osid[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) pos[d * cap + N + i] = pos[d * cap + p] - B.shift[sh][d];
}

// Per step of a run whose pair kernel kicks (records of the domain slots
// written by the previous step's pair kernel): ghost positions from their
// sources, also into the sorted positions spos and, with quant (FP32C), into
// the quantised records of the ghost slots, against the cell of the last
// build (w of the cell list's own record, cell).
template <class T>
__global__ __launch_bounds__(256) void k_ens_ghost_rec(const GhostBoxes<T>* __restrict__ gb, T* __restrict__ pos,
                                                       size_t cap, uint32_t N, const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ sid,
                                                       const uint32_t* __restrict__ opid,
                                                       const uint32_t* __restrict__ osid,
                                                       const uint32_t* __restrict__ n2s, T* __restrict__ spos,
                                                       int4* __restrict__ quant, const int4* __restrict__ cell,
                                                       const SysGrid<T>* __restrict__ sys) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (N + i >= counts[0]) return;
  const uint32_t p = opid[i];
  const uint32_t sy = sid[p];
  const GhostBoxes<T>& B = gb[sy];
  const uint32_t sh = osid[i];
  const uint32_t f = n2s[N + i];
  T x[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x[d] = pos[d * cap + p] - B.shift[sh][d];
    pos[d * cap + N + i] = x[d];
    spos[d * cap + f] = x[d];
  }
  if (quant) {
    const int w = cell[f].w;
    const uint32_t cc[3] = {uint32_t(w) & 2047u, (uint32_t(w) >> 11) & 2047u, (uint32_t(w) >> 22) & 1023u};
    const SysGrid<T>& G = sys[sy];
    const double xs[3] = {double(x[0]), double(x[1]), double(x[2])};
    quant[f] = quantise(xs, cc, G.qlo, G.qsp, G.qinv);
  }
}

// Per step: positions into sorted order, and the quantised stream relative to
// the cells of the last build.
template <class T>
__global__ __launch_bounds__(256) void k_ens_reorder(const uint32_t* __restrict__ counts, size_t cap,
                                                     const uint32_t* __restrict__ s2n, const T* __restrict__ pos,
                                                     T* __restrict__ spos, int4* __restrict__ quant,
                                                     const uint32_t* __restrict__ sid_sorted,
                                                     const SysGrid<T>* __restrict__ sys) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= counts[0]) return;
  const uint32_t o = s2n[f];
  const T x = pos[o], y = pos[cap + o], z = pos[2 * cap + o];
  spos[f] = x;
  spos[cap + f] = y;
  spos[2 * cap + f] = z;
  if (quant) {
    const int w = quant[f].w;
    const uint32_t cc[3] = {uint32_t(w) & 2047u, (uint32_t(w) >> 11) & 2047u, (uint32_t(w) >> 22) & 1023u};
    const SysGrid<T>& G = sys[sid_sorted[f]];
    const double xs[3] = {double(x), double(y), double(z)};
    quant[f] = quantise(xs, cc, G.qlo, G.qsp, G.qinv);
  }
}

// First half of velocity Verlet: v += dt/(2m) F, x += dt v; a rebuild is
// flagged (flags[0]) if the drift carried a particle beyond skin/2 of its
// position at the last build.  sid null: one system, constants dyn[0].
template <class T>
__global__ __launch_bounds__(256) void k_ens_kick_drift(uint32_t N, size_t cap, T* __restrict__ pos,
                                                        T* __restrict__ vel, const T* __restrict__ frc,
                                                        const T* __restrict__ xb, const uint32_t* __restrict__ sid,
                                                        const SysDyn<T>* __restrict__ dyn, T dt,
                                                        uint32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const SysDyn<T>& D = dyn[sid ? sid[i] : 0u];
  const T h = (T(0.5) * dt) * D.inv_m;
  T dd = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T v = vel[d * cap + i] + h * frc[d * cap + i];
    vel[d * cap + i] = v;
    const T x = pos[d * cap + i] + dt * v;
    pos[d * cap + i] = x;
    const T r = x - xb[d * cap + i];
    dd = dd + r * r;
  }
  if (dd > D.lim2) atomicOr(&flags[0], 1u);
}

// Forces (sorted slots, FT) back to storage order.  kick 1: second half kick
// v += dt/(2m) F.  kick 2: that kick and, fused with it, the next step's first
// half kick and drift (k_ens_kick_drift's arithmetic, with its displacement
// check), so a run of steps reads and writes the particle arrays once per step.
// sid null: one system, constants dyn[0].
template <class T, class FT>
__global__ __launch_bounds__(256) void k_ens_kick_end(uint32_t N, size_t cap, const uint32_t* __restrict__ n2s,
                                                      const FT* __restrict__ fs, const FT* __restrict__ es,
                                                      T* __restrict__ en, T* __restrict__ pos,
                                                      T* __restrict__ vel, T* __restrict__ frc,
                                                      const T* __restrict__ xb, const uint32_t* __restrict__ sid,
                                                      const SysDyn<T>* __restrict__ dyn, T dt, int kick,
                                                      uint32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t f = n2s[i];
  T a[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) a[d] = T(fs[d * cap + f]);
  // the stored forces are read by the next run's first kick and by
  // ofp_ens_get: a step fused with the next one (kick 2) leaves them to the
  // run's last step
  if (kick != 2) {
#pragma unroll
    for (int d = 0; d < 3; ++d) frc[d * cap + i] = a[d];
  }
  if (en) en[i] = T(es[f]);
  if (!kick) return;
  const SysDyn<T>& D = dyn[sid ? sid[i] : 0u];
  if (kick == 1) {
    const T h = (T(0.5) * dt) * D.inv_m;
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d * cap + i] = vel[d * cap + i] + h * a[d];
    return;
  }
  T x[3];
  kick_kick_drift<T>(i, cap, pos, vel, xb, D, dt, a, flags, x);
}

// Caller order <-> storage order.
template <class T>
__global__ __launch_bounds__(256) void k_ens_set(uint32_t N, size_t cap, const T* __restrict__ pos,
                                                 const T* __restrict__ vel, size_t stride, T* __restrict__ spos,
                                                 T* __restrict__ svel, T* __restrict__ sfrc,
                                                 uint32_t* __restrict__ gid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    spos[d * cap + i] = pos[d * stride + i];
    svel[d * cap + i] = vel ? vel[d * stride + i] : T(0);
    sfrc[d * cap + i] = T(0);
  }
  gid[i] = i;
}

template <class T>
__global__ __launch_bounds__(256) void k_ens_get(uint32_t N, size_t cap, const uint32_t* __restrict__ gid,
                                                 const T* __restrict__ spos, const T* __restrict__ svel,
                                                 const T* __restrict__ sfrc, T* __restrict__ pos, T* __restrict__ vel,
                                                 T* __restrict__ frc, size_t stride) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const uint32_t i = gid[k];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (pos) pos[d * stride + i] = spos[d * cap + k];
    if (vel) vel[d * stride + i] = svel[d * cap + k];
    if (frc) frc[d * stride + i] = sfrc[d * cap + k];
  }
}

}  // namespace ofp

using namespace ofp;

struct ofp_ens {
  int dtype = OFP_F32;
  uint32_t S = 0;
  uint64_t N = 0;
  ofp_ens_opts opt{};
  std::vector<ofp_sys_desc> desc;
  std::vector<uint64_t> first;      // S + 1
  std::vector<uint32_t> cbase;      // S + 1
  std::vector<uint32_t> divc;       // 3 S
  std::vector<double> clo, csp;     // 3 S
  uint64_t n_cells = 0;
  int n_cu = 256;
  size_t cap = 0, gcap = 0, ecap = 0;
  hipStream_t cs = nullptr;  // capture stream
  // device tables
  DevBuf sysgrid, rbase, gboxes, dyn, r2v, ljtab[2], ljq[2];
  int lj_cubic[2] = {0, 0};  // every system's cells cubic (set with the tables)
  // particle arrays (3 planes of cap unless noted)
  DevBuf pos, vel, frc, xb, pos2, vel2, frc2, gid, gid2, sid, spos, fsorted, esorted, energy_tmp, nn;
  // the second set of sorted positions and FP32C records a run's steps
  // alternate with (set 0: spos and the cell list's records)
  DevBuf spos2, quant2;
  int rec_set = 0;  // the set holding the current sorted positions and records
  DevBuf gcnt, gstart, gtmp, opid, osid;
  DevBuf counts, flags, e_dev;  // counts[2], flags[2], e_dev[1] (u64)
  uint32_t* hflags = nullptr;   // pinned: flags[2], counts[2], E (2 words)
  ofp_cl* cl = nullptr;
  ofp_vl* vl = nullptr;
  bool have_pos = false, built = false, sized = false, rebuild_pending = true;
  int force_prec = -1;  // precision of the forces in frc (-1: none)
  uint64_t n_ghost = 0, n_entries = 0, rebuilds = 0, steps = 0, recaptures = 0;
  // graphs: [step variant][precision] step (variants: see step_variant),
  // rebuild
  // A rebuild permutes the stored particles into the *2 arrays and swaps the
  // two sets; spar counts the swaps mod 2, and the graphs are kept per spar
  // (they name the arrays of the set that was current when captured).
  static constexpr int kVariants = 6;
  int spar = 0;
  hipGraphExec_t g_step[2][kVariants][2] = {};
  hipGraphExec_t g_rebuild[2] = {};
  // realloc_epoch() at each capture: a graph is replayed only while it is unchanged
  uint64_t g_step_epoch[2][kVariants][2] = {};
  uint64_t g_rebuild_epoch[2] = {};
  double g_dt[2][kVariants][2] = {};
  void* g_energy = nullptr;
};

namespace ofp {

static int ens_destroy_graphs(ofp_ens* e) {
  for (int q = 0; q < 2; ++q) {
    for (int l = 0; l < ofp_ens::kVariants; ++l)
      for (int p = 0; p < 2; ++p)
        if (e->g_step[q][l][p]) {
          OFP_HIP(hipGraphExecDestroy(e->g_step[q][l][p]));
          e->g_step[q][l][p] = nullptr;
        }
    if (e->g_rebuild[q]) {
      OFP_HIP(hipGraphExecDestroy(e->g_rebuild[q]));
      e->g_rebuild[q] = nullptr;
    }
  }
  return OFP_OK;
}

// The permuted copies become the stored particles (the forces too when they
// were permuted).
static void ens_swap_storage(ofp_ens* e, bool with_frc) {
  std::swap(e->pos, e->pos2);
  std::swap(e->vel, e->vel2);
  if (with_frc) std::swap(e->frc, e->frc2);
  std::swap(e->gid, e->gid2);
  e->spar ^= 1;
}

// Drop the graphs captured before a device buffer was allocated, grown or
// released anywhere in the library: their kernel arguments may name freed
// memory or old capacities (a sizing rebuild can grow the Verlet mask buffer
// without changing the particle or entry capacities).
static int ens_drop_stale_graphs(ofp_ens* e) {
  const uint64_t now = realloc_epoch();
  for (int q = 0; q < 2; ++q) {
    for (int l = 0; l < ofp_ens::kVariants; ++l)
      for (int p = 0; p < 2; ++p)
        if (e->g_step[q][l][p] && e->g_step_epoch[q][l][p] != now) {
          OFP_HIP(hipGraphExecDestroy(e->g_step[q][l][p]));
          e->g_step[q][l][p] = nullptr;
        }
    if (e->g_rebuild[q] && e->g_rebuild_epoch[q] != now) {
      OFP_HIP(hipGraphExecDestroy(e->g_rebuild[q]));
      e->g_rebuild[q] = nullptr;
    }
  }
  return OFP_OK;
}

template <class T>
static int ens_tables(ofp_ens* e) {
  const uint32_t S = e->S;
  std::vector<SysGrid<T>> sg(S);
  std::vector<GhostBoxes<T>> gb(S);
  std::vector<SysDyn<T>> dy(S);
  std::vector<T> r2(S);
  std::vector<uint32_t> rb(S + 1);
  e->cbase.assign(S + 1, 0);
  e->divc.assign(3 * S, 0);
  e->clo.assign(3 * S, 0);
  e->csp.assign(3 * S, 0);
  uint64_t cells = 0, rows = 0;
  double gmax = 0;
  for (uint32_t s = 0; s < S; ++s) {
    const ofp_sys_desc& d = e->desc[s];
    const double rv = d.r_cut * (1.0 + e->opt.skin);
    T lo[3], hi[3], glo[3], ghi[3], blo[3], bhi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = T(d.box_lo[k]);
      hi[k] = T(d.box_hi[k]);
      glo[k] = T(-rv);
      ghi[k] = T(rv);
    }
    uint32_t div[3];
    OFP_TRY(ofp_cl_param(e->dtype, lo, hi, rv / e->opt.sub, glo, ghi, 1.013, blo, bhi, div));
    SysGrid<T>& G = sg[s];
    for (int k = 0; k < 3; ++k) {
      OFP_CHECK(div[k] + 2 <= (k == 2 ? 1024u : 2048u), OFP_E_LIMIT, "system %u: %u cells along dim %d", s,
                div[k] + 2, k);
      G.lo[k] = blo[k];
      G.sp[k] = T((bhi[k] - blo[k]) / T(div[k]));
      G.divc[k] = div[k] + 2;
      G.qlo[k] = double(G.lo[k]);
      G.qsp[k] = double(G.sp[k]);
      G.qinv[k] = double(kQuantOne) / double(G.sp[k]);
      e->divc[3 * s + k] = G.divc[k];
      e->clo[3 * s + k] = double(G.lo[k]);
      e->csp[3 * s + k] = double(G.sp[k]);
      OFP_CHECK(double(G.sp[k]) * e->opt.sub >= rv * (1.0 - 1e-6), OFP_E_LIMIT,
                "system %u: stencil smaller than the Verlet radius", s);
    }
    G.cbase = uint32_t(cells);
    e->cbase[s] = uint32_t(cells);
    rb[s] = uint32_t(rows);
    cells += uint64_t(G.divc[0]) * G.divc[1] * G.divc[2];
    rows += uint64_t(G.divc[1]) * G.divc[2];
    OFP_CHECK(cells < (1ull << 31), OFP_E_LIMIT, "too many cells in the ensemble");
    // ghost boxes of the single-domain periodic system
    const int bc[3] = {1, 1, 1};
    T gblo[3 * kMaxBoxes], gbhi[3 * kMaxBoxes], sh[81];
    uint32_t sv[kMaxBoxes];
    int nb = 0;
    OFP_TRY(ofp_local_ghost_boxes(e->dtype, lo, hi, glo, ghi, bc, gblo, gbhi, sv, &nb));
    OFP_TRY(ofp_shift_vectors(e->dtype, lo, hi, sh));
    fill_boxes(gb[s], nb, gblo, gbhi, sv, sh);
    SysDyn<T>& D = dy[s];
    for (int k = 0; k < 3; ++k) {
      D.lo[k] = lo[k];
      D.hi[k] = hi[k];
      D.L[k] = T(hi[k] - lo[k]);
    }
    const double half = 0.5 * (rv - d.r_cut);
    D.lim2 = T(half * half);
    D.inv_m = T(1.0 / d.mass);
    const T rvt = T(rv);
    r2[s] = rvt * rvt;
    const double L = d.box_hi[0] - d.box_lo[0];
    gmax = std::max(gmax, rv / L);
  }
  e->cbase[S] = uint32_t(cells);
  rb[S] = uint32_t(rows);
  e->n_cells = cells;
  OFP_TRY(e->sysgrid.ensure(S * sizeof(SysGrid<T>)));
  OFP_TRY(e->gboxes.ensure(S * sizeof(GhostBoxes<T>)));
  OFP_TRY(e->dyn.ensure(S * sizeof(SysDyn<T>)));
  OFP_TRY(e->r2v.ensure(S * sizeof(T)));
  OFP_TRY(e->rbase.ensure((S + 1) * sizeof(uint32_t)));
  OFP_HIP(hipMemcpy(e->sysgrid.p, sg.data(), S * sizeof(SysGrid<T>), hipMemcpyHostToDevice));
  OFP_HIP(hipMemcpy(e->gboxes.p, gb.data(), S * sizeof(GhostBoxes<T>), hipMemcpyHostToDevice));
  OFP_HIP(hipMemcpy(e->dyn.p, dy.data(), S * sizeof(SysDyn<T>), hipMemcpyHostToDevice));
  OFP_HIP(hipMemcpy(e->r2v.p, r2.data(), S * sizeof(T), hipMemcpyHostToDevice));
  OFP_HIP(hipMemcpy(e->rbase.p, rb.data(), (S + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  OFP_TRY(cl_create_many(&e->cl, e->dtype, S, e->sysgrid.p, e->rbase.as<uint32_t>(), uint32_t(rows), cells,
                         OFP_CL_NON_SYMMETRIC | OFP_CL_QUANT));
  OFP_TRY(ofp_cl_set_box_nn(e->cl, e->opt.sub));
  // initial ghost capacity: uniform-density estimate of every system's ghost layer
  double gest = 0;
  for (uint32_t s = 0; s < S; ++s) {
    const ofp_sys_desc& d = e->desc[s];
    const double rv = d.r_cut * (1.0 + e->opt.skin);
    double f = 1;
    for (int k = 0; k < 3; ++k) f *= 1.0 + 2.0 * rv / (d.box_hi[k] - d.box_lo[k]);
    gest += double(d.n) * (f - 1.0);
  }
  e->gcap = size_t(gest * e->opt.headroom) + 4096;
  return OFP_OK;
}

template <class T>
static size_t tsz() {
  return sizeof(T);
}

// (Re)allocate the cap-strided particle arrays for the current capacities,
// keeping the stored domain data.
template <class T>
static int ens_alloc(ofp_ens* e, hipStream_t s) {
  const size_t cap_new = size_t(e->N) + e->gcap + 64;
  if (cap_new <= e->cap && e->pos.p) return OFP_OK;
  const size_t old = e->cap;
  auto grow3 = [&](DevBuf& b, size_t esz, bool keep) -> int {
    DevBuf nb;
    OFP_TRY(nb.ensure(3 * cap_new * esz));
    if (keep && b.p)
      for (int d = 0; d < 3; ++d)
        OFP_HIP(hipMemcpyAsync(static_cast<char*>(nb.p) + d * cap_new * esz,
                               static_cast<char*>(b.p) + d * old * esz, e->N * esz, hipMemcpyDeviceToDevice, s));
    OFP_HIP(hipStreamSynchronize(s));
    b.release();
    b = nb;
    return OFP_OK;
  };
  OFP_TRY(grow3(e->pos, sizeof(T), true));
  OFP_TRY(grow3(e->vel, sizeof(T), true));
  OFP_TRY(grow3(e->frc, sizeof(T), true));
  OFP_TRY(grow3(e->xb, sizeof(T), true));
  OFP_TRY(grow3(e->pos2, sizeof(T), false));
  OFP_TRY(grow3(e->vel2, sizeof(T), false));
  OFP_TRY(grow3(e->frc2, sizeof(T), false));
  OFP_TRY(grow3(e->spos, sizeof(T), false));
  OFP_TRY(grow3(e->spos2, sizeof(T), false));
  OFP_TRY(e->quant2.ensure(cap_new * sizeof(int4)));
  OFP_TRY(grow3(e->fsorted, sizeof(double), false));
  OFP_TRY(e->esorted.ensure(cap_new * sizeof(double)));
  OFP_TRY(e->energy_tmp.ensure(cap_new * sizeof(T)));
  OFP_TRY(e->nn.ensure(cap_new * sizeof(uint32_t)));
  if (!e->sid.p || cap_new > old) {
    DevBuf nb;
    OFP_TRY(nb.ensure(cap_new * 4));
    if (e->sid.p) OFP_HIP(hipMemcpyAsync(nb.p, e->sid.p, e->N * 4, hipMemcpyDeviceToDevice, s));
    OFP_HIP(hipStreamSynchronize(s));
    e->sid.release();
    e->sid = nb;
  }
  OFP_TRY(e->gid.ensure(e->N * 4 + 4));
  OFP_TRY(e->gid2.ensure(e->N * 4 + 4));
  OFP_TRY(e->opid.ensure(e->gcap * 4 + 4));
  OFP_TRY(e->osid.ensure(e->gcap * 4 + 4));
  e->cap = cap_new;
  return OFP_OK;
}

// The rebuild: map (+ storage permutation), ghosts, cell list, Verlet list.
// sizing: counts read back and capacities grown (synchronous, not capturable).
// keep_frc = false: the stored forces are not read before they are written
// again (a rebuild between a run's steps: the run's last step stores the
// forces), so they are not permuted.
template <class T>
static int ens_rebuild_t(ofp_ens* e, int sizing, hipStream_t s, bool keep_frc) {
  const uint32_t N = uint32_t(e->N);
  const size_t cap = e->cap;
  uint32_t* flags = e->flags.as<uint32_t>();
  uint32_t* counts = e->counts.as<uint32_t>();
  const uint32_t* sid = e->sid.as<uint32_t>();
  const SysDyn<T>* dyn = e->dyn.as<SysDyn<T>>();
  OFP_TRY(fill_u32(flags + 1, 0, 1, s));
  const unsigned gN = grid_for(N, 256);
  if (e->built) {
    ofp_cl_view v;
    OFP_TRY(ofp_cl_get_view(e->cl, &v));
    k_ens_map<T><<<gN, 256, 0, s>>>(N, cap, v.dom_sort_ids, v.sort_to_nonsort, e->pos.as<T>(), e->vel.as<T>(),
                                    keep_frc ? e->frc.as<T>() : nullptr, e->gid.as<uint32_t>(), e->pos2.as<T>(),
                                    e->vel2.as<T>(), keep_frc ? e->frc2.as<T>() : nullptr,
                                    e->gid2.as<uint32_t>(), e->xb.as<T>(), sid, dyn);
    OFP_LAUNCHED();
    ens_swap_storage(e, keep_frc);
  } else {
    k_ens_map<T><<<gN, 256, 0, s>>>(N, cap, nullptr, nullptr, e->pos.as<T>(), nullptr, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, e->xb.as<T>(), sid, dyn);
    OFP_LAUNCHED();
  }
  // ghosts
  OFP_TRY(e->gcnt.ensure((size_t(N) + 1) * 4));
  OFP_TRY(e->gstart.ensure((size_t(N) + 1) * 4));
  OFP_TRY(e->gtmp.ensure(scan_tmp_bytes(N)));
  const GhostBoxes<T>* gb = e->gboxes.as<GhostBoxes<T>>();
  k_ens_ghost_count<T><<<gN, 256, 0, s>>>(gb, e->pos.as<T>(), cap, N, sid, e->gcnt.as<uint32_t>());
  OFP_LAUNCHED();
  OFP_TRY(scan_excl_u32(e->gcnt.as<uint32_t>(), e->gstart.as<uint32_t>(), N, e->gtmp.p, s));
  if (sizing) {
    uint32_t G = 0;
    OFP_HIP(hipMemcpyAsync(&G, e->gstart.as<uint32_t>() + N, 4, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    if (G > e->gcap) {
      e->gcap = size_t(double(G) * std::max(1.0, e->opt.headroom)) + 4096;
      OFP_TRY(ens_alloc<T>(e, s));
    }
  }
  k_ens_ghost_emit<T><<<gN, 256, 0, s>>>(gb, e->pos.as<T>(), e->cap, N, e->sid.as<uint32_t>(),
                                         e->gstart.as<uint32_t>(), uint32_t(e->gcap), e->opid.as<uint32_t>(),
                                         e->osid.as<uint32_t>());
  OFP_LAUNCHED();
  k_ens_counts<<<1, 64, 0, s>>>(N, e->gstart.as<uint32_t>(), uint32_t(e->gcap), counts, flags);
  OFP_LAUNCHED();
  // cell list over all grids
  OFP_TRY(cl_construct_many(e->cl, e->pos.p, e->cap, e->sid.as<uint32_t>(), e->cap, counts, N, e->spos.p, e->cap, s));
  if (e->opt.pair_walk == 1) {
    // cell tiles: the item queue of the new cells is the pair pass's work (built
    // here, so a captured rebuild rebuilds it); no Verlet list
    const TargetItem* items = nullptr;
    const uint32_t* n_items = nullptr;
    uint32_t* counter = nullptr;
    size_t icap = 0;
    OFP_TRY(ofp_cl_items_(e->cl, &items, &n_items, &counter, &icap, s));
    OFP_TRY(fill_u64(e->e_dev.as<uint64_t>(), 0, 1, s));
    if (sizing) {
      uint32_t c[2];
      OFP_HIP(hipMemcpyAsync(c, counts, 8, hipMemcpyDeviceToHost, s));
      OFP_HIP(hipStreamSynchronize(s));
      e->n_ghost = c[1];
      e->n_entries = 0;
      e->sized = true;
    }
    e->built = true;
    return OFP_OK;
  }
  // Verlet list
  const unsigned g = unsigned(e->n_cu) * std::max(1u, e->opt.waves_per_cu / 4);  // workgroups of 4 waves
  uint64_t E = 0;
  OFP_TRY(vl_build_many(e->vl, e->cl, e->dtype, e->spos.p, e->cap, e->sysgrid.p, e->r2v.p, &e->ecap,
                        e->opt.headroom, sizing, g, &E, s, int(e->opt.verlet_rows)));
  const unsigned long long* off = nullptr;
  OFP_TRY(vl_offsets(e->vl, &off));
  const unsigned long long* mstate = nullptr;
  OFP_TRY(vl_mask_state(e->vl, &mstate));
  k_ens_check_entries<<<1, 64, 0, s>>>(off + N, (unsigned long long)e->ecap, mstate, flags,
                                       e->e_dev.as<unsigned long long>());
  OFP_LAUNCHED();
  if (sizing) {
    uint32_t c[2];
    OFP_HIP(hipMemcpyAsync(c, counts, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    e->n_ghost = c[1];
    e->n_entries = E;
    e->sized = true;
  }
  e->built = true;
  return OFP_OK;
}

static int ens_tables_lj(ofp_ens* e, int prec) {
  if (e->ljtab[prec].p) return OFP_OK;
  std::vector<ofp_lj_params> prm(e->S);
  for (uint32_t s = 0; s < e->S; ++s) prm[s] = ofp_lj_params{e->desc[s].r_cut, e->desc[s].sigma, e->desc[s].eps};
  double* fs = e->fsorted.as<double>();
  const size_t cap = e->cap;
  uint32_t* nn = e->nn.as<uint32_t>();
  const int f64 = e->dtype == OFP_F64 ? 1 : 0;
  if (prec == OFP_PREC_F64)
    return lj_tables_many(prec, e->S, prm.data(), e->csp.data(), fs, fs + cap, fs + 2 * cap, e->esorted.p, nn,
                          e->spos.p, cap, f64, &e->ljtab[prec], &e->ljq[prec], &e->lj_cubic[prec]);
  float* f32 = e->fsorted.as<float>();
  return lj_tables_many(prec, e->S, prm.data(), e->csp.data(), f32, f32 + cap, f32 + 2 * cap, e->esorted.p, nn,
                        e->spos.p, cap, f64, &e->ljtab[prec], &e->ljq[prec], &e->lj_cubic[prec]);
}

// The sorted positions and FP32C records of set k (0: spos and the cell
// list's own records, 1: spos2, quant2).
struct RecSet {
  void* spos;
  int4* quant;
};
static RecSet rec_set(ofp_ens* e, int k, const ofp_cl_view& v) {
  return k == 0 ? RecSet{e->spos.p, (int4*)v.quant} : RecSet{e->spos2.p, e->quant2.as<int4>()};
}
// Steps over Verlet rows kick inside the pair kernel (unless told not to).
static bool kick_in_pairs(const ofp_ens* e) { return !e->opt.pair_walk && !e->opt.split_kick; }

// The force pass: ghosts, sorted positions and records, LJ, back to storage
// order.  kick 0: forces stored (the explicit pass); 1: the last step of a
// run (forces stored, second half kick); 2: a step followed by another (the
// second half kick fused with the next step's first half kick and drift).
// set < 0: ghost refresh, then the sorted positions and records of every slot
// made from the positions (reorder into set 0).  set k >= 0 (kick_in_pairs
// only): the domain slots of set k hold the current positions (written by the
// previous step's pair kernel), the ghost refresh writes the ghost slots.
// With kick 2 and kick_in_pairs, the pair kernel kicks and drifts each row's
// particle and writes its new sorted position and record to the other set.
template <class T>
static int ens_forces_t(ofp_ens* e, int prec, int kick, double dt, void* energy, hipStream_t s, int set) {
  const uint32_t N = uint32_t(e->N);
  const size_t cap = e->cap;
  const uint32_t* counts = e->counts.as<uint32_t>();
  const GhostBoxes<T>* gb = e->gboxes.as<GhostBoxes<T>>();
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(e->cl, &v));
  const int cur = set < 0 ? 0 : set;
  const RecSet R = rec_set(e, cur, v);
  if (set < 0) {
    if (e->gcap > 0) {
      k_ens_ghost_refresh<T><<<grid_for(e->gcap, 256), 256, 0, s>>>(gb, e->pos.as<T>(), cap, N, counts,
                                                                    e->sid.as<uint32_t>(), e->opid.as<uint32_t>(),
                                                                    e->osid.as<uint32_t>());
      OFP_LAUNCHED();
    }
    const uint32_t* sid_sorted = nullptr;
    OFP_TRY(cl_sid_sorted(e->cl, &sid_sorted));
    k_ens_reorder<T><<<grid_for(cap, 256), 256, 0, s>>>(counts, cap, v.sort_to_nonsort, e->pos.as<T>(),
                                                        e->spos.as<T>(),
                                                        prec == OFP_PREC_F32C ? (int4*)v.quant : nullptr,
                                                        sid_sorted, e->sysgrid.as<SysGrid<T>>());
    OFP_LAUNCHED();
  } else {
    OFP_CHECK(kick_in_pairs(e), OFP_E_STATE, "record sets alternate only when the pair kernel kicks");
    if (e->gcap > 0) {
      k_ens_ghost_rec<T><<<grid_for(e->gcap, 256), 256, 0, s>>>(
          gb, e->pos.as<T>(), cap, N, counts, e->sid.as<uint32_t>(), e->opid.as<uint32_t>(), e->osid.as<uint32_t>(),
          v.nonsort_to_sort, static_cast<T*>(R.spos), prec == OFP_PREC_F32C ? R.quant : nullptr,
          (const int4*)v.quant, e->sysgrid.as<SysGrid<T>>());
      OFP_LAUNCHED();
    }
  }
  e->rec_set = cur;
  OFP_TRY(ens_tables_lj(e, prec));
  const uint32_t* sidp = e->S > 1 ? e->sid.as<uint32_t>() : nullptr;  // one system: its constants for all
  if (kick == 2 && kick_in_pairs(e)) {
    const RecSet X = rec_set(e, 1 - cur, v);
    const StepKick<T> K{v.sort_to_nonsort, e->pos.as<T>(), e->vel.as<T>(), e->xb.as<T>(), cap,
                        e->dyn.as<SysDyn<T>>(), sidp ? 1 : 0, T(dt), e->flags.as<uint32_t>(),
                        static_cast<T*>(X.spos), prec == OFP_PREC_F32C ? X.quant : nullptr,
                        e->sysgrid.as<SysGrid<T>>()};
    return pair_lj_many(e->cl, e->vl, e->dtype, R.spos, cap, prec, e->ljtab[prec].p, e->ljq[prec].p,
                        int(e->opt.lanes), e->ecap, 0, e->lj_cubic[prec], s, e->sysgrid.p, &K, R.quant);
  }
  // a step (kick) needs the forces only; energy and neighbour counts are for
  // the explicit force pass
  OFP_TRY(pair_lj_many(e->cl, e->opt.pair_walk ? nullptr : e->vl, e->dtype, R.spos, cap, prec, e->ljtab[prec].p,
                       e->ljq[prec].p, int(e->opt.lanes), e->ecap, kick ? 0 : 1, e->lj_cubic[prec], s,
                       e->sysgrid.p, nullptr, R.quant));
  const unsigned gN = grid_for(N, 256);
  T* en = static_cast<T*>(energy);
  if (prec == OFP_PREC_F64)
    k_ens_kick_end<T, double><<<gN, 256, 0, s>>>(N, cap, v.nonsort_to_sort, e->fsorted.as<double>(),
                                                 e->esorted.as<double>(), en, e->pos.as<T>(), e->vel.as<T>(),
                                                 e->frc.as<T>(), e->xb.as<T>(), sidp, e->dyn.as<SysDyn<T>>(), T(dt),
                                                 kick, e->flags.as<uint32_t>());
  else
    k_ens_kick_end<T, float><<<gN, 256, 0, s>>>(N, cap, v.nonsort_to_sort, e->fsorted.as<float>(),
                                                e->esorted.as<float>(), en, e->pos.as<T>(), e->vel.as<T>(),
                                                e->frc.as<T>(), e->xb.as<T>(), sidp, e->dyn.as<SysDyn<T>>(), T(dt),
                                                kick, e->flags.as<uint32_t>());
  OFP_LAUNCHED();
  return OFP_OK;
}

// First half kick and drift of a run (flags[0]: rebuild due).
template <class T>
static int ens_drift_t(ofp_ens* e, double dt, hipStream_t s) {
  OFP_TRY(fill_u32(e->flags.as<uint32_t>(), 0, 1, s));
  const uint32_t* sidp = e->S > 1 ? e->sid.as<uint32_t>() : nullptr;  // one system: its constants for all
  k_ens_kick_drift<T><<<grid_for(e->N, 256), 256, 0, s>>>(uint32_t(e->N), e->cap, e->pos.as<T>(), e->vel.as<T>(),
                                                          e->frc.as<T>(), e->xb.as<T>(), sidp, e->dyn.as<SysDyn<T>>(),
                                                          T(dt), e->flags.as<uint32_t>());
  OFP_LAUNCHED();
  return OFP_OK;
}

// Forces at the drifted positions, then (last) the second half kick, or (not
// last) that kick fused with the next step's first half kick and drift.  set:
// as for ens_forces_t.
template <class T>
static int ens_step_t(ofp_ens* e, int prec, double dt, int last, int set, hipStream_t s) {
  if (!last) OFP_TRY(fill_u32(e->flags.as<uint32_t>(), 0, 1, s));
  return ens_forces_t<T>(e, prec, last ? 1 : 2, dt, nullptr, s, set);
}

static int ens_rebuild(ofp_ens* e, int sizing, hipStream_t s, bool keep_frc = true) {
  return e->dtype == OFP_F32 ? ens_rebuild_t<float>(e, sizing, s, keep_frc)
                             : ens_rebuild_t<double>(e, sizing, s, keep_frc);
}
static int ens_drift(ofp_ens* e, double dt, hipStream_t s) {
  return e->dtype == OFP_F32 ? ens_drift_t<float>(e, dt, s) : ens_drift_t<double>(e, dt, s);
}
static int ens_step(ofp_ens* e, int prec, double dt, int last, int set, hipStream_t s) {
  return e->dtype == OFP_F32 ? ens_step_t<float>(e, prec, dt, last, set, s)
                             : ens_step_t<double>(e, prec, dt, last, set, s);
}
static int ens_forces(ofp_ens* e, int prec, void* energy, hipStream_t s) {
  return e->dtype == OFP_F32 ? ens_forces_t<float>(e, prec, 0, 0.0, energy, s, -1)
                             : ens_forces_t<double>(e, prec, 0, 0.0, energy, s, -1);
}
// Graph of a step: set -1, 0, 1 (see ens_forces_t) x last step or not.
static int step_variant(int set, int last) { return (set + 1) * 2 + last; }

// Capture fn(capture stream) as a graph; *epoch = realloc_epoch() after it.
template <class F>
static int capture(ofp_ens* e, hipGraphExec_t* out, uint64_t* epoch, F&& fn) {
  hipGraph_t g = nullptr;
  OFP_HIP(hipStreamBeginCapture(e->cs, hipStreamCaptureModeThreadLocal));
  const int rc = fn(e->cs);
  hipError_t ec = hipStreamEndCapture(e->cs, &g);
  if (rc != OFP_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return set_error(OFP_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
  hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return set_error(OFP_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  *epoch = realloc_epoch();
  e->recaptures++;
  return OFP_OK;
}

// Sizing rebuild (host-synchronous) and, if the capacities moved, new graphs.
static int ens_sizing(ofp_ens* e, hipStream_t s) {
  const size_t cap0 = e->cap, e0 = e->ecap;
  OFP_TRY(ens_rebuild(e, 1, s));
  OFP_HIP(hipStreamSynchronize(s));
  if (e->cap != cap0 || e->ecap != e0) {
    OFP_TRY(ens_destroy_graphs(e));
    for (int p = 0; p < 2; ++p) {
      e->ljtab[p].release();
      e->ljq[p].release();
    }
  }
  e->rebuilds++;
  e->rebuild_pending = false;
  return OFP_OK;
}

}  // namespace ofp

extern "C" {

int ofp_ens_create(ofp_ens** out, int dtype, uint32_t n_sys, const ofp_sys_desc* sys, const ofp_ens_opts* opts) {
  OFP_CHECK(out && sys && n_sys > 0, OFP_E_ARG, "ofp_ens_create: null argument or no systems");
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  ofp_ens_opts o{};
  if (opts) o = *opts;
  if (o.skin <= 0) o.skin = 0.1;
  if (o.sub == 0) o.sub = 2;
  if (o.lanes == 0) o.lanes = 16;
  if (o.waves_per_cu == 0) o.waves_per_cu = 32;
  if (o.headroom < 1.0) o.headroom = 1.25;
  OFP_CHECK(o.sub >= 1 && o.sub <= uint32_t(kMaxBoxNN), OFP_E_LIMIT, "sub must be in [1, %d]", kMaxBoxNN);
  OFP_CHECK(o.verlet_rows <= 2, OFP_E_ARG, "verlet_rows must be 0, 1 or 2");
  OFP_CHECK(o.pair_walk <= 1, OFP_E_ARG, "pair_walk must be 0 (Verlet rows) or 1 (cell tiles)");
  OFP_CHECK(o.split_kick <= 1, OFP_E_ARG, "split_kick must be 0 or 1");
  OFP_CHECK(o.lanes == 1 || o.lanes == 2 || o.lanes == 4 || o.lanes == 8 || o.lanes == 16 || o.lanes == 32,
            OFP_E_ARG, "lanes must be 1, 2, 4, 8, 16 or 32");
  uint64_t N = 0;
  for (uint32_t s = 0; s < n_sys; ++s) {
    const ofp_sys_desc& d = sys[s];
    OFP_CHECK(d.r_cut > 0 && d.mass > 0 && d.sigma > 0, OFP_E_ARG, "system %u: r_cut, sigma, mass must be > 0", s);
    for (int k = 0; k < 3; ++k) {
      const double L = d.box_hi[k] - d.box_lo[k];
      OFP_CHECK(L > 2 * d.r_cut * (1 + o.skin), OFP_E_LIMIT,
                "system %u: box side %g must exceed twice the Verlet radius", s, L);
    }
    N += d.n;
  }
  OFP_CHECK(N > 0 && N < (1ull << 31), OFP_E_LIMIT, "total particles must be in [1, 2^31)");
  ofp_ens* e = new ofp_ens();
  e->dtype = dtype;
  e->S = n_sys;
  e->N = N;
  e->opt = o;
  e->desc.assign(sys, sys + n_sys);
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      e->n_cu = prop.multiProcessorCount;
  }
  e->first.assign(n_sys + 1, 0);
  for (uint32_t s = 0; s < n_sys; ++s) e->first[s + 1] = e->first[s] + sys[s].n;
  int rc = dtype == OFP_F32 ? ens_tables<float>(e) : ens_tables<double>(e);
  if (rc == OFP_OK) rc = ofp_vl_create(&e->vl);
  if (rc == OFP_OK) rc = e->counts.ensure(16);
  if (rc == OFP_OK) rc = e->flags.ensure(16);
  if (rc == OFP_OK) rc = e->e_dev.ensure(16);
  if (rc == OFP_OK) rc = dtype == OFP_F32 ? ens_alloc<float>(e, nullptr) : ens_alloc<double>(e, nullptr);
  if (rc == OFP_OK && hipHostMalloc(reinterpret_cast<void**>(&e->hflags), 64) != hipSuccess)
    rc = set_error(OFP_E_HIP, "hipHostMalloc failed");
  if (rc == OFP_OK && hipStreamCreateWithFlags(&e->cs, hipStreamNonBlocking) != hipSuccess)
    rc = set_error(OFP_E_HIP, "hipStreamCreateWithFlags failed");
  if (rc == OFP_OK) {
    // system of each stored domain particle (fixed: systems keep their ranges)
    std::vector<uint32_t> sid(N);
    for (uint32_t s = 0; s < n_sys; ++s)
      std::fill(sid.begin() + e->first[s], sid.begin() + e->first[s + 1], s);
    if (hipMemcpy(e->sid.p, sid.data(), N * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = set_error(OFP_E_HIP, "hipMemcpy(sid) failed");
  }
  if (rc != OFP_OK) {
    ofp_ens_destroy(e);
    return rc;
  }
  *out = e;
  return OFP_OK;
}

int ofp_ens_destroy(ofp_ens* e) {
  if (!e) return OFP_OK;
  (void)ens_destroy_graphs(e);
  if (e->cl) ofp_cl_destroy(e->cl);
  if (e->vl) ofp_vl_destroy(e->vl);
  if (e->cs) (void)hipStreamDestroy(e->cs);
  if (e->hflags) (void)hipHostFree(e->hflags);
  DevBuf* bufs[] = {&e->sysgrid, &e->rbase,  &e->gboxes, &e->dyn,     &e->r2v,    &e->ljtab[0], &e->ljtab[1],
                    &e->ljq[0],  &e->ljq[1], &e->pos,    &e->vel,     &e->frc,    &e->xb,       &e->pos2,
                    &e->vel2,    &e->frc2,   &e->gid,    &e->gid2,    &e->sid,    &e->spos,     &e->fsorted,
                    &e->esorted, &e->energy_tmp, &e->gcnt, &e->gstart, &e->gtmp,  &e->opid,     &e->osid,
                    &e->counts,  &e->flags,  &e->e_dev,  &e->nn};
  for (DevBuf* b : bufs) b->release();
  delete e;
  return OFP_OK;
}

int ofp_ens_first(const ofp_ens* e, uint64_t* first) {
  OFP_CHECK(e && first, OFP_E_ARG, "ofp_ens_first: null argument");
  std::copy(e->first.begin(), e->first.end(), first);
  return OFP_OK;
}

int ofp_ens_arrays_get(const ofp_ens* e, ofp_ens_arrays* a) {
  OFP_CHECK(e && a, OFP_E_ARG, "ofp_ens_arrays_get: null argument");
  a->pos = e->pos.p;
  a->vel = e->vel.p;
  a->force = e->frc.p;
  a->gid = e->gid.as<uint32_t>();
  a->sid = e->sid.as<uint32_t>();
  a->force_sorted = e->fsorted.p;
  a->n_neigh = e->nn.as<uint32_t>();
  a->cap = e->cap;
  a->n_dom = e->N;
  return OFP_OK;
}

int ofp_ens_lists(const ofp_ens* e, const ofp_cl** cl, const ofp_vl** vl) {
  OFP_CHECK(e && cl && vl, OFP_E_ARG, "ofp_ens_lists: null argument");
  OFP_CHECK(e->built, OFP_E_STATE, "ensemble not built");
  *cl = e->cl;
  *vl = e->opt.pair_walk ? nullptr : e->vl;  // cell tiles keep no Verlet list
  return OFP_OK;
}

int ofp_ens_cells(const ofp_ens* e, uint32_t* base, uint32_t* divc, double* lo, double* spacing) {
  OFP_CHECK(e, OFP_E_ARG, "ofp_ens_cells: null handle");
  if (base) std::copy(e->cbase.begin(), e->cbase.end(), base);
  if (divc) std::copy(e->divc.begin(), e->divc.end(), divc);
  if (lo) std::copy(e->clo.begin(), e->clo.end(), lo);
  if (spacing) std::copy(e->csp.begin(), e->csp.end(), spacing);
  return OFP_OK;
}

int ofp_ens_set(ofp_ens* e, const void* pos, const void* vel, size_t stride, void* stream) {
  OFP_CHECK(e && pos, OFP_E_ARG, "ofp_ens_set: null argument");
  OFP_CHECK(stride >= e->N, OFP_E_ARG, "stride %zu < particles %llu", stride, (unsigned long long)e->N);
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_for(e->N, 256);
  if (e->dtype == OFP_F32)
    k_ens_set<float><<<g, 256, 0, s>>>(uint32_t(e->N), e->cap, (const float*)pos, (const float*)vel, stride,
                                       e->pos.as<float>(), e->vel.as<float>(), e->frc.as<float>(),
                                       e->gid.as<uint32_t>());
  else
    k_ens_set<double><<<g, 256, 0, s>>>(uint32_t(e->N), e->cap, (const double*)pos, (const double*)vel, stride,
                                        e->pos.as<double>(), e->vel.as<double>(), e->frc.as<double>(),
                                        e->gid.as<uint32_t>());
  OFP_LAUNCHED();
  e->have_pos = true;
  e->built = false;
  e->sized = false;
  e->rebuild_pending = true;
  e->force_prec = -1;
  return OFP_OK;
}

int ofp_ens_get(const ofp_ens* e, void* pos, void* vel, void* force, size_t stride, void* stream) {
  OFP_CHECK(e && e->have_pos, OFP_E_STATE, "ofp_ens_get: no particles set");
  OFP_CHECK(stride >= e->N, OFP_E_ARG, "stride < particles");
  hipStream_t s = as_stream(stream);
  const unsigned g = grid_for(e->N, 256);
  if (e->dtype == OFP_F32)
    k_ens_get<float><<<g, 256, 0, s>>>(uint32_t(e->N), e->cap, e->gid.as<uint32_t>(), e->pos.as<float>(),
                                       e->vel.as<float>(), e->frc.as<float>(), (float*)pos, (float*)vel,
                                       (float*)force, stride);
  else
    k_ens_get<double><<<g, 256, 0, s>>>(uint32_t(e->N), e->cap, e->gid.as<uint32_t>(), e->pos.as<double>(),
                                        e->vel.as<double>(), e->frc.as<double>(), (double*)pos, (double*)vel,
                                        (double*)force, stride);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_ens_rebuild(ofp_ens* e, int sizing, void* stream) {
  OFP_CHECK(e && e->have_pos, OFP_E_STATE, "ofp_ens_rebuild: no particles set");
  hipStream_t s = as_stream(stream);
  if (sizing || !e->sized) return ens_sizing(e, s);
  OFP_TRY(ens_rebuild(e, 0, s));
  // a rebuild that outgrew a capacity wrote nothing beyond it and flagged the
  // overflow: redo it as a sizing rebuild (as the step loop does)
  OFP_HIP(hipMemcpyAsync(e->hflags, e->flags.p, 8, hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  if (e->hflags[1] != 0) return ens_sizing(e, s);
  e->rebuilds++;
  e->rebuild_pending = false;
  return OFP_OK;
}

int ofp_ens_forces(ofp_ens* e, int precision, void* energy, void* stream) {
  OFP_CHECK(e && e->have_pos, OFP_E_STATE, "ofp_ens_forces: no particles set");
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG, "bad precision %d", precision);
  hipStream_t s = as_stream(stream);
  if (!e->built || !e->sized) OFP_TRY(ens_sizing(e, s));
  OFP_TRY(ens_forces(e, precision, e->energy_tmp.p, s));
  if (energy) {
    const size_t esz = e->dtype == OFP_F32 ? 4 : 8;
    OFP_HIP(hipMemcpyAsync(energy, e->energy_tmp.p, e->N * esz, hipMemcpyDeviceToDevice, s));
  }
  e->force_prec = precision;
  return OFP_OK;
}

int ofp_ens_pair_pass(ofp_ens* e, int precision, int with_energy, void* stream) {
  OFP_CHECK(e && e->built && e->force_prec >= 0, OFP_E_STATE, "ofp_ens_pair_pass: run ofp_ens_forces first");
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG, "bad precision %d", precision);
  OFP_TRY(ens_tables_lj(e, precision));
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(e->cl, &v));
  const RecSet R = rec_set(e, e->rec_set, v);  // the positions of the last pass
  return pair_lj_many(e->cl, e->opt.pair_walk ? nullptr : e->vl, e->dtype, R.spos, e->cap, precision,
                      e->ljtab[precision].p, e->ljq[precision].p, int(e->opt.lanes), e->ecap, with_energy ? 1 : 0,
                      e->lj_cubic[precision], as_stream(stream), e->sysgrid.p, nullptr, R.quant);
}

int ofp_ens_run(ofp_ens* e, int precision, double dt, int n_steps, int use_graph, void* stream) {
  OFP_CHECK(e && e->have_pos, OFP_E_STATE, "ofp_ens_run: no particles set");
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG, "bad precision %d", precision);
  OFP_CHECK(dt >= 0 && n_steps >= 0, OFP_E_ARG, "dt and n_steps must be >= 0");
  hipStream_t s = as_stream(stream);
  if (!e->built || !e->sized) OFP_TRY(ens_sizing(e, s));
  if (e->force_prec != precision) {
    OFP_TRY(ens_forces(e, precision, nullptr, s));
    e->force_prec = precision;
  }
  if (n_steps == 0) return OFP_OK;
  OFP_TRY(ens_tables_lj(e, precision));
  for (int q = 0; q < 2; ++q)
    for (int l = 0; l < ofp_ens::kVariants; ++l)
      if (use_graph && e->g_step[q][l][precision] && e->g_dt[q][l][precision] != dt) {
        OFP_HIP(hipGraphExecDestroy(e->g_step[q][l][precision]));
        e->g_step[q][l][precision] = nullptr;
      }
  uint32_t* hf = e->hflags;
  auto read_flag = [&]() -> int {
    OFP_HIP(hipMemcpyAsync(hf, e->flags.p, 4, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    if (hf[0] != 0) e->rebuild_pending = true;
    return OFP_OK;
  };
  // The run: first half kick and drift; then per step a rebuild if one is due
  // (a particle beyond skin/2 of its position at the last build, or set up by
  // the caller), the forces at the drifted positions, and the second half kick
  // -- fused, except after the last step, with the next step's first half kick
  // and drift, whose displacement check decides the next rebuild.  set: the
  // sorted positions and records the next step reads (-1: made from the
  // positions, as after the drift and after a rebuild; 0 or 1: written by the
  // previous step's pair kernel).
  OFP_TRY(ens_drift(e, dt, s));
  OFP_TRY(read_flag());
  int set = -1;
  for (int it = 0; it < n_steps; ++it) {
    if (e->rebuild_pending) {
      if (use_graph) {
        OFP_TRY(ens_drop_stale_graphs(e));
        // a capture swaps the stored sets as the rebuild does; a replay of an
        // earlier capture swaps them here
        const int q = e->spar;
        if (!e->g_rebuild[q])
          OFP_TRY(capture(e, &e->g_rebuild[q], &e->g_rebuild_epoch[q],
                          [&](hipStream_t cs) { return ens_rebuild(e, 0, cs, false); }));
        else
          ens_swap_storage(e, false);
        OFP_HIP(hipGraphLaunch(e->g_rebuild[q], s));
      } else {
        OFP_TRY(ens_rebuild(e, 0, s, false));
      }
      OFP_HIP(hipMemcpyAsync(hf, e->flags.p, 8, hipMemcpyDeviceToHost, s));
      OFP_HIP(hipStreamSynchronize(s));
      if (hf[1] != 0) OFP_TRY(ens_sizing(e, s));  // capacities grown, graphs dropped
      e->rebuilds += hf[1] != 0 ? 0 : 1;
      e->rebuild_pending = false;
      OFP_TRY(ens_tables_lj(e, precision));
      set = -1;
    }
    const int last = it == n_steps - 1 ? 1 : 0;
    const int var = step_variant(set, last);
    if (use_graph) {
      OFP_TRY(ens_drop_stale_graphs(e));
      const int q = e->spar;
      if (!e->g_step[q][var][precision]) {
        OFP_TRY(capture(e, &e->g_step[q][var][precision], &e->g_step_epoch[q][var][precision],
                        [&](hipStream_t cs) { return ens_step(e, precision, dt, last, set, cs); }));
        e->g_dt[q][var][precision] = dt;
      }
      OFP_HIP(hipGraphLaunch(e->g_step[q][var][precision], s));
      e->rec_set = set < 0 ? 0 : set;  // as the captured pass sets it
    } else {
      OFP_TRY(ens_step(e, precision, dt, last, set, s));
    }
    // the next step reads what this one's pair kernel wrote
    if (!last && kick_in_pairs(e)) set = 1 - (set < 0 ? 0 : set);
    if (!last) OFP_TRY(read_flag());
    e->steps++;
  }
  return OFP_OK;
}

// Test hook (not in the public header): give the Verlet list a mask buffer of
// exactly mask_cap masks (0: keep it) and, if pending != 0, make the next step
// start with a rebuild.  A later rebuild that needs more masks flags the
// overflow and is redone as a sizing rebuild, as after any configuration change.
int ofp_ens_test_masks_(ofp_ens* e, size_t mask_cap, int pending, size_t* old_cap) {
  OFP_CHECK(e && e->built, OFP_E_STATE, "ofp_ens_test_masks_: ensemble not built");
  if (mask_cap) OFP_TRY(vl_set_mask_cap(e->vl, mask_cap, old_cap));
  if (pending) e->rebuild_pending = true;
  return OFP_OK;
}

int ofp_ens_status(ofp_ens* e, ofp_ens_state* st, void* stream) {
  OFP_CHECK(e && st, OFP_E_ARG, "ofp_ens_status: null argument");
  std::memset(st, 0, sizeof(*st));
  hipStream_t s = as_stream(stream);
  if (e->built) {
    uint32_t c[2] = {0, 0}, f[2] = {0, 0};
    unsigned long long E = 0;
    OFP_HIP(hipMemcpyAsync(c, e->counts.p, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipMemcpyAsync(f, e->flags.p, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipMemcpyAsync(&E, e->e_dev.p, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    st->n_ghost = c[1];
    st->n_entries = E;
    st->overflow = f[1];
  }
  st->n_dom = e->N;
  st->ghost_cap = e->gcap;
  st->entry_cap = e->ecap;
  st->rebuilds = e->rebuilds;
  st->steps = e->steps;
  st->recaptures = e->recaptures;
  st->rebuild_pending = e->rebuild_pending ? 1u : 0u;
  return OFP_OK;
}

}  // extern "C"
