// ofp_internal.hpp -- shared helpers of libofp_nn.so (gfx950 only).
//
// Error reporting, grow-only device buffers, the plane-copy argument block, the
// cell-key formula shared by every kernel, and the wave/block scan primitives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>

#include "ofp_nn.h"

namespace ofp {

// ------------------------------------------------------------------ errors
int set_error(int code, const char* fmt, ...);
void clear_error();

#define OFP_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::ofp::set_error(OFP_E_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #call,    \
                              hipGetErrorString(e_));                                    \
  } while (0)

#define OFP_CHECK(cond, code, ...)                                                       \
  do {                                                                                   \
    if (!(cond)) return ::ofp::set_error((code), __VA_ARGS__);                           \
  } while (0)

#define OFP_TRY(expr)                                                                    \
  do {                                                                                   \
    int r_ = (expr);                                                                     \
    if (r_ != OFP_OK) return r_;                                                         \
  } while (0)

// Last launch error (a bad launch configuration is reported here).
#define OFP_LAUNCHED() OFP_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------- device buffers
// Process-wide count of device-buffer (re)allocations and releases.  A captured
// hipGraph holds buffer addresses and capacities as kernel arguments, so it is
// valid only while this count is unchanged since its capture.
uint64_t realloc_epoch();
void bump_realloc_epoch();

// Grow-only device allocation owned by a handle.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap && p) return OFP_OK;
    bump_realloc_epoch();
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return set_error(OFP_E_HIP, "hipFree: %s", hipGetErrorString(e));
    }
    size_t b = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return set_error(OFP_E_HIP, "hipMalloc(%zu): %s", b, hipGetErrorString(e));
    }
    cap = b;
    return OFP_OK;
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      bump_realloc_epoch();
    }
    p = nullptr;
    cap = 0;
  }
  template <class U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

// --------------------------------------------------------- plane copies
constexpr int kMaxPlanes = 24;

// Up to kMaxPlanes (src, dst) plane pairs, passed to kernels by value.
struct PlaneArgs {
  uint64_t src[kMaxPlanes];
  uint64_t dst[kMaxPlanes];
  uint32_t esz[kMaxPlanes];
  int n;
};

int make_plane_args(const ofp_plane* in, const ofp_plane* out, int n, PlaneArgs* a);

__device__ __forceinline__ void copy_planes(const PlaneArgs& a, size_t from, size_t to) {
  for (int k = 0; k < a.n; ++k) {
    if (a.esz[k] == 8) {
      reinterpret_cast<uint64_t*>(a.dst[k])[to] = reinterpret_cast<const uint64_t*>(a.src[k])[from];
    } else {
      reinterpret_cast<uint32_t*>(a.dst[k])[to] = reinterpret_cast<const uint32_t*>(a.src[k])[from];
    }
  }
}

// ------------------------------------------------------------ cell key
// Cell grid of CellList_gpu: one padding layer, key_d = floor((x_d - lo_d) /
// spacing_d) + 1 in the position dtype, linearised x-fastest over divc = div + 2.
// Coordinates outside the padded grid are clamped to it.
template <class T>
struct CellGeom {
  T lo[3];
  T sp[3];
  uint32_t divc[3];
};

template <class T>
__host__ __device__ __forceinline__ uint32_t cell_coord(T x, T lo, T sp, uint32_t divc) {
#pragma clang fp contract(off)
  T f = floor((x - lo) / sp);
  if (!(f == f)) return 0u;  // NaN -> padding cell
  if (f < T(-1)) return 0u;
  if (f > T(divc) - T(2)) return divc - 1u;
  return uint32_t(int(f) + 1);
}

template <class T>
__host__ __device__ __forceinline__ uint32_t cell_of(const CellGeom<T>& g, T x, T y, T z,
                                                     uint32_t* cx, uint32_t* cy,
                                                     uint32_t* cz) {
  *cx = cell_coord(x, g.lo[0], g.sp[0], g.divc[0]);
  *cy = cell_coord(y, g.lo[1], g.sp[1], g.divc[1]);
  *cz = cell_coord(z, g.lo[2], g.sp[2], g.divc[2]);
  return *cx + g.divc[0] * (*cy + g.divc[1] * *cz);
}

// One system of an ensemble: its cell grid (same formula as CellGeom) placed at
// global cell ids [cbase, cbase + divc0 divc1 divc2), and the FP64 copies the
// quantised stream is computed from.
template <class T>
struct SysGrid {
  T lo[3];
  T sp[3];
  uint32_t divc[3];
  uint32_t cbase;
  double qlo[3], qsp[3], qinv[3];
  __host__ __device__ CellGeom<T> geom() const {
    CellGeom<T> g;
    for (int k = 0; k < 3; ++k) {
      g.lo[k] = lo[k];
      g.sp[k] = sp[k];
      g.divc[k] = divc[k];
    }
    return g;
  }
};

// Cell key of particle i: one grid, or the grid of the particle's system.
template <class T>
struct KeyOne {
  CellGeom<T> g;
  const T* pos;
  size_t stride;
  __device__ __forceinline__ uint32_t operator()(uint32_t i) const {
    uint32_t cx, cy, cz;
    return cell_of(g, pos[i], pos[stride + i], pos[2 * stride + i], &cx, &cy, &cz);
  }
};

template <class T>
struct KeyMany {
  const SysGrid<T>* sys;
  const uint32_t* sid;  // system of every particle (domain and ghost)
  const T* pos;
  size_t stride;
  __device__ __forceinline__ uint32_t operator()(uint32_t i) const {
    const SysGrid<T>& G = sys[sid[i]];
    uint32_t cx, cy, cz;
    return G.cbase + cell_of(G.geom(), pos[i], pos[stride + i], pos[2 * stride + i], &cx, &cy, &cz);
  }
};

// ------------------------------------------------------------ ghost boxes
// Local ghost boxes of one single-domain periodic system (sorted by shift id)
// and its 27 shift vectors.
constexpr int kMaxBoxes = 26;

// in_lo / in_hi: a region that no box reaches (filt != 0), so a particle
// inside it has no ghost and skips the box tests.  Every box is a slab of the
// boxes' bounding box in some dimension d: it starts at the lower face (then
// in_lo[d] >= its upper bound) or ends at the upper face (then in_hi[d] <= its
// lower bound); a particle with in_lo <= x < in_hi in every dimension is thus
// outside every box.  A box that is a slab in no dimension turns the filter off.
template <class T>
struct GhostBoxes {
  T lo[kMaxBoxes][3];
  T hi[kMaxBoxes][3];
  T shift[27][3];
  uint32_t sv[kMaxBoxes];
  int nb;
  int filt;
  T in_lo[3], in_hi[3];
};

// Box::isInsideNP: half-open on every side.
template <class T>
__device__ __forceinline__ bool inside_np(const GhostBoxes<T>& B, int b, T x, T y, T z) {
  return x >= B.lo[b][0] && x < B.hi[b][0] && y >= B.lo[b][1] && y < B.hi[b][1] &&
         z >= B.lo[b][2] && z < B.hi[b][2];
}

// True when (x, y, z) lies in the region no box reaches (no ghost).
template <class T>
__device__ __forceinline__ bool no_box(const GhostBoxes<T>& B, T x, T y, T z) {
  return B.filt && x >= B.in_lo[0] && x < B.in_hi[0] && y >= B.in_lo[1] && y < B.in_hi[1] && z >= B.in_lo[2] &&
         z < B.in_hi[2];
}

template <class T>
inline void fill_boxes(GhostBoxes<T>& B, int nb, const void* blo, const void* bhi, const uint32_t* sv,
                       const void* shifts27) {
  B = GhostBoxes<T>{};
  const T* lo = static_cast<const T*>(blo);
  const T* hi = static_cast<const T*>(bhi);
  const T* sh = static_cast<const T*>(shifts27);
  B.nb = nb;
  for (int b = 0; b < nb; ++b) {
    for (int k = 0; k < 3; ++k) {
      B.lo[b][k] = lo[3 * b + k];
      B.hi[b][k] = hi[3 * b + k];
    }
    B.sv[b] = sv[b];
  }
  for (int s = 0; s < 27; ++s)
    for (int k = 0; k < 3; ++k) B.shift[s][k] = sh[3 * s + k];
  // the region no box reaches
  T dlo[3], dhi[3];
  for (int k = 0; k < 3; ++k) {
    dlo[k] = std::numeric_limits<T>::max();
    dhi[k] = std::numeric_limits<T>::lowest();
    B.in_lo[k] = std::numeric_limits<T>::lowest();
    B.in_hi[k] = std::numeric_limits<T>::max();
  }
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) {
      dlo[k] = B.lo[b][k] < dlo[k] ? B.lo[b][k] : dlo[k];
      dhi[k] = B.hi[b][k] > dhi[k] ? B.hi[b][k] : dhi[k];
    }
  B.filt = nb > 0 ? 1 : 0;
  for (int b = 0; b < nb; ++b) {
    bool slab = false;
    for (int k = 0; k < 3; ++k) {
      const bool at_lo = B.lo[b][k] <= dlo[k], at_hi = B.hi[b][k] >= dhi[k];
      if (at_lo && at_hi) continue;  // spans the dimension
      if (at_lo) {
        B.in_lo[k] = B.hi[b][k] > B.in_lo[k] ? B.hi[b][k] : B.in_lo[k];
        slab = true;
      } else if (at_hi) {
        B.in_hi[k] = B.lo[b][k] < B.in_hi[k] ? B.lo[b][k] : B.in_hi[k];
        slab = true;
      }
    }
    if (!slab) B.filt = 0;
  }
}

// distance2 as the reference's Point::distance2: ((dx^2 + dy^2) + dz^2), unfused,
// in the position dtype -- the Verlet test must round exactly like the oracle.
template <class T>
__device__ __forceinline__ T dist2_unfused(T dx, T dy, T dz) {
#pragma clang fp contract(off)
  T r = dx * dx;
  r = r + dy * dy;
  r = r + dz * dz;
  return r;
}

// Cell-relative quantised coordinates of the FP32 pair kernels: per dim,
// u = rint((x - cell_low) / q) with q = spacing / 2^28, stored as
// g = u + c 2^28 (mod 2^32), c the cell coordinate.  The separation of two
// particles whose cells are at most kMaxBoxNN apart is the exact integer
//   d_int = u_q - u_p + (c_q - c_p) 2^28 = int32(g_q - g_p),
// |d_int| < (kMaxBoxNN + 1) 2^28 <= 2^30, so the wrapped subtraction is exact.
constexpr int kQuantBits = 28;
constexpr int kMaxBoxNN = 3;
constexpr int32_t kQuantOne = int32_t(1) << kQuantBits;
// cell coordinates packed in the 4th int: cx | cy << 11 | cz << 22.
__host__ __device__ __forceinline__ int32_t pack_cell(uint32_t cx, uint32_t cy, uint32_t cz) {
  return int32_t(cx | (cy << 11) | (cz << 22));
}

// g = u + c 2^28 (mod 2^32) with u = rint((x - cell_low) / q), u clamped to
// +-2^29 (a particle may drift past its cell between rebuilds).  qlo, qsp:
// cell-list box low corner and spacing, qinv = 2^28 / spacing, cc = cell coords.
__device__ __forceinline__ int4 quantise(const double xs[3], const uint32_t cc[3], const double qlo[3],
                                         const double qsp[3], const double qinv[3]) {
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double org = qlo[k] + double(int(cc[k]) - 1) * qsp[k];
    double u = rint((xs[k] - org) * qinv[k]);
    u = u < -536870912.0 ? -536870912.0 : (u > 536870912.0 ? 536870912.0 : u);
    q[k] = uint32_t(int(u)) + (cc[k] << kQuantBits);
  }
  return make_int4(int(q[0]), int(q[1]), int(q[2]), pack_cell(cc[0], cc[1], cc[2]));
}

// ------------------------------------------------------------- launch
constexpr int kWave = 64;

inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  return unsigned(g == 0 ? 1 : g);
}

// ------------------------------------------------------------- scans
// Exclusive scans with the total appended at out[n]: out has n + 1 entries.
// tmp must hold scan_tmp_bytes(n) bytes.  No host synchronisation.
size_t scan_tmp_bytes(size_t n);
int scan_excl_u32(const uint32_t* in, uint32_t* out, size_t n, void* tmp, hipStream_t s);
int scan_excl_u64(const uint64_t* in, uint64_t* out, size_t n, void* tmp, hipStream_t s);
// 32-bit counts widened to 64-bit prefix sums (Verlet offsets can exceed 2^32).
int scan_excl_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, void* tmp,
                         hipStream_t s);

// Fill helper (graph-safe replacement for hipMemsetAsync).
int fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
int fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s);

// Wave-level inclusive scan (64 lanes).
template <class V>
__device__ __forceinline__ V wave_incl_scan(V x) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    V y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

template <class V>
__device__ __forceinline__ V wave_sum(V x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// w * 2 + hit, per lane: the compare's lane mask enters v_addc as its carry-in,
// so appending a hit bit costs one VALU op after the compare.
__device__ __forceinline__ uint32_t shift_in(uint32_t w, bool hit) {
  const unsigned long long m = __ballot(hit);
  uint32_t r;
  asm("v_addc_co_u32 %0, vcc, %1, %1, %2" : "=v"(r) : "v"(w), "s"(m) : "vcc");
  return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
// block b runs on XCD b % 8.  This maps the blocks of XCD x onto one contiguous
// range of tiles, so the tiles that run side by side on one XCD are neighbours
// and their writes to adjacent output runs meet in the same L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n_blocks) {
  const uint32_t x = b % kXcds, i = b / kXcds;
  const uint32_t per = n_blocks / kXcds, rem = n_blocks % kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

// ------------------------------------------------------ cell-list view
// What the walkers need of a built cell list, passed to kernels by value.
struct ClDev {
  const unsigned long long* cs;  // packed starts: low = sorted index, high = domain position
  const uint32_t* dom;           // domain sort ids
  const int4* quant;             // cell-relative stream (or null)
  uint32_t divc[3];
  uint32_t n_dom;
  uint32_t box_nn;  // stencil half-width in cells
  uint64_t n_cells;
};

// Target items of the Verlet build: up to 64 consecutive domain rows whose
// cells lie on one x-row of the grid (cells [cx0, cx0 + ncx) at (cy, cz)).
// An item spans at most kItemMaxCells cells (the FP32C cell-tile kernels rely
// on it for their separations).
constexpr uint32_t kItemMaxCells = 8;
struct TargetItem {
  uint32_t row0;
  uint32_t n;
  uint32_t cell0;  // linear id of the first cell
  uint32_t ncx;
};

// Grid, stencil origin and radius of one target item: the cell list's own grid,
// or (ensembles) the grid of the item's system, whose cells start at cbase.
template <class T>
struct ItemGeom {
  uint32_t d0, d1, d2, cbase;
  int cx0, cy, cz, ncx;
  T r2;
};

template <class T>
__device__ __forceinline__ ItemGeom<T> item_geom(const ClDev& c, const TargetItem& ti, T r2,
                                                 const SysGrid<T>* sys, const T* r2_sys) {
  ItemGeom<T> g;
  g.ncx = int(ti.ncx & 0xFFu);
  g.r2 = r2;
  if (sys) {
    const uint32_t s = ti.ncx >> 8;
    g.d0 = sys[s].divc[0];
    g.d1 = sys[s].divc[1];
    g.d2 = sys[s].divc[2];
    g.cbase = sys[s].cbase;
    if (r2_sys) g.r2 = r2_sys[s];
  } else {
    g.d0 = c.divc[0];
    g.d1 = c.divc[1];
    g.d2 = c.divc[2];
    g.cbase = 0;
  }
  const uint32_t cl = ti.cell0 - g.cbase;
  g.cx0 = int(cl % g.d0);
  g.cy = int((cl / g.d0) % g.d1);
  g.cz = int(cl / (g.d0 * g.d1));
  return g;
}

// The item's (2B+1)^2 x-runs in traversal order: fn(ja, jb) with wave-uniform
// bounds of sorted indices.
template <class T, class F>
__device__ __forceinline__ void for_each_run(const ClDev& c, const ItemGeom<T>& G, F&& fn) {
  const int B = int(c.box_nn);
  const int xa = G.cx0 - B > 0 ? G.cx0 - B : 0;
  const int xe = G.cx0 + G.ncx - 1 + B;
  const int xb = xe < int(G.d0) - 1 ? xe : int(G.d0) - 1;
  for (int dz = -B; dz <= B; ++dz) {
    const int zz = G.cz + dz;
    if (zz < 0 || zz >= int(G.d2)) continue;
    for (int dy = -B; dy <= B; ++dy) {
      const int yy = G.cy + dy;
      if (yy < 0 || yy >= int(G.d1)) continue;
      const uint32_t rowc = G.cbase + G.d0 * (uint32_t(yy) + G.d1 * uint32_t(zz));
      const uint32_t ja = __builtin_amdgcn_readfirstlane(uint32_t(c.cs[rowc + uint32_t(xa)]));
      const uint32_t jb = __builtin_amdgcn_readfirstlane(uint32_t(c.cs[rowc + uint32_t(xb) + 1]));
      if (ja < jb) fn(ja, jb);
    }
  }
}

// Ensemble cell lists: one sort over the grids of all systems (global cell ids).
// sys_dev = SysGrid<T>[n_sys], rbase_dev = first global x-row of each system
// (n_sys + 1 entries); both stay owned by the caller.
int cl_create_many(ofp_cl** out, int dtype, uint32_t n_sys, const void* sys_dev, const uint32_t* rbase_dev,
                   uint32_t n_rows, uint64_t n_cells, uint32_t opt);
// sid = system of every particle; the particle count is *n_dev (<= n_cap).
int cl_construct_many(ofp_cl* h, const void* pos, size_t stride, const uint32_t* sid, size_t n_cap,
                      const uint32_t* n_dev, size_t gm, void* pos_sorted, size_t sorted_stride, hipStream_t s);
int cl_sid_sorted(const ofp_cl* h, const uint32_t** sid_sorted);
// Ensemble Verlet build (no host synchronisation); the entry count stays on the
// device at offsets[n_rows] and the list is written only if it fits e_cap.
// r2_sys = squared Verlet radius per system in the position dtype.
// sizing != 0: the count is read back (synchronous), *e_cap grows to at least
// count x headroom and nbr with it.  grid = persistent workgroups (0: default).
// rows_mode: 0 = chosen by the mean candidates per target, 1 = the item
// kernels, 2 = thread per row.
int vl_build_many(ofp_vl* h, const ofp_cl* cl, int dtype, const void* pos_sorted, size_t stride,
                  const void* sys_dev, const void* r2_sys_dev, size_t* e_cap, double headroom, int sizing,
                  unsigned grid, uint64_t* e_count, hipStream_t s, int rows_mode);
int vl_offsets(const ofp_vl* h, const unsigned long long** off);
// [0] masks used by the last count pass, [1] != 0 if they did not fit
int vl_mask_state(const ofp_vl* h, const unsigned long long** mstate);
// Replace the mask buffer by one of exactly mask_cap masks (the next build
// that needs more flags the overflow); *old_cap (may be null) = the previous one.
int vl_set_mask_cap(ofp_vl* h, size_t mask_cap, size_t* old_cap);
// Ensemble LJ over the Verlet rows: per-system constant tables (built once by
// lj_tables_many for fixed output arrays), `lanes` lanes per row (8, 16, 32);
// with_u = 0 writes the forces only (energy and n_neigh are left as they are).
// all_cubic (may be null) tells whether every system's cells are cubic; pass it
// to pair_lj_many as `cubic`.
int lj_tables_many(int precision, uint32_t n_sys, const ofp_lj_params* prm, const double* spacing, void* fx,
                   void* fy, void* fz, void* energy, uint32_t* n_neigh, const void* pos_sorted, size_t pstride,
                   int pos_f64, DevBuf* tab, DevBuf* qtab, int* all_cubic);
// Per-system dynamics constants of an ensemble.
template <class T>
struct SysDyn {
  T lo[3], hi[3], L[3];  // periodic box
  T lim2;                // (skin r_cut / 2)^2, rebuild threshold on the squared displacement
  T inv_m;
};

// Velocity Verlet between two force evaluations, for particle i (storage
// order) with force a: the second half kick v += dt/(2m) a of the step that
// ends, the next step's first half kick and its drift x += dt v, and the
// displacement check against the position at the last build (flags[0] |= 1
// when the squared displacement exceeds lim2).
// The arithmetic on values: v, x updated in place; returns whether the
// displacement exceeds the threshold.
template <class T>
__device__ __forceinline__ bool kick_kick_drift_v(T (&v)[3], T (&x)[3], const T (&xb)[3], const SysDyn<T>& D, T dt,
                                                  const T (&a)[3]) {
#pragma clang fp contract(off)
  const T h = (T(0.5) * dt) * D.inv_m;
  T dd = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T v1 = v[d] + h * a[d];  // end of this step
    v[d] = v1 + h * a[d];          // first half kick of the next
    x[d] = x[d] + dt * v[d];
    const T r = x[d] - xb[d];
    dd = dd + r * r;
  }
  return dd > D.lim2;
}
// xo: the new position.
template <class T>
__device__ inline void kick_kick_drift(uint32_t i, size_t cap, T* __restrict__ pos, T* __restrict__ vel,
                                       const T* __restrict__ xb, const SysDyn<T>& D, T dt, const T (&a)[3],
                                       uint32_t* __restrict__ flags, T (&xo)[3]) {
  T v[3], b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v[d] = vel[d * cap + i];
    xo[d] = pos[d * cap + i];
    b[d] = xb[d * cap + i];
  }
  const bool far = kick_kick_drift_v<T>(v, xo, b, D, dt, a);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel[d * cap + i] = v[d];
    pos[d * cap + i] = xo[d];
  }
  if (far) atomicOr(&flags[0], 1u);
}

// The pair kernels that end a step of an ensemble run apply kick_kick_drift to
// the particle of each row as soon as its force is summed (the force is not
// stored).  s2n maps a row's sorted slot p to its storage index; per_sys = 0:
// one system, constants dyn[0] / sys[0], else those of the row's system.  The
// new position also goes to slot p of the next step's sorted positions (snext)
// and, for FP32C, of its quantised records (qnext, against the cell of the
// last build, w of the row's current record): the pair pass of the next step
// reads those, while this one reads the current ones, which no lane changes.
struct NoKick {
  static constexpr bool kOn = false;
  struct Pre {};
  __device__ __forceinline__ uint32_t index(uint32_t) const { return 0u; }
  __device__ __forceinline__ Pre prefetch(uint32_t) const { return Pre{}; }
};
template <class T>
struct StepKick {
  static constexpr bool kOn = true;
  using Tv = T;
  const uint32_t* s2n;
  T* pos;
  T* vel;
  const T* xb;
  size_t cap;
  const SysDyn<T>* dyn;
  int per_sys;
  T dt;
  uint32_t* flags;
  T* snext;
  int4* qnext;  // null: no quantised records (FP64 pass)
  const SysGrid<T>* sys;
  // the particle's storage index and state, loaded ahead of the kick (the
  // kernels issue these loads before their last gathers, so they wait once)
  struct Pre {
    uint32_t i;
    T v[3], x[3], xb[3];
  };
  __device__ __forceinline__ uint32_t index(uint32_t p) const { return s2n[p]; }
  __device__ __forceinline__ Pre prefetch(uint32_t i) const {
    Pre q;
    q.i = i;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      q.v[d] = vel[d * cap + i];
      q.x[d] = pos[d * cap + i];
      q.xb[d] = xb[d * cap + i];
    }
    return q;
  }
  __device__ __forceinline__ void apply(const Pre& q, uint32_t p, uint32_t sy, int w, const T (&a)[3]) const {
    const uint32_t s = per_sys ? sy : 0u;
    T v[3] = {q.v[0], q.v[1], q.v[2]};
    T x[3] = {q.x[0], q.x[1], q.x[2]};
    const bool far = kick_kick_drift_v<T>(v, x, q.xb, dyn[s], dt, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      vel[d * cap + q.i] = v[d];
      pos[d * cap + q.i] = x[d];
      snext[d * cap + p] = x[d];
    }
    if (far) atomicOr(&flags[0], 1u);
    if (qnext) {
      const uint32_t cc[3] = {uint32_t(w) & 2047u, (uint32_t(w) >> 11) & 2047u, (uint32_t(w) >> 22) & 1023u};
      const double xs[3] = {double(x[0]), double(x[1]), double(x[2])};
      const SysGrid<T>& G = sys[s];
      qnext[p] = quantise(xs, cc, G.qlo, G.qsp, G.qinv);
    }
  }
};

// vl == null: the cell-tile pass over the cell list's item queue (the cells of
// the last build, whose stencil covers r_cut + skin), sysgrid = SysGrid<T>[n_sys]
// of the position dtype.  kick (Verlet rows only, with_u = 0): a StepKick<T>
// of the position dtype, applied to every row in place of storing its force.
// Over Verlet rows, pos_sorted (sorted positions, plane stride `stride`) and
// quant (FP32C records; null: the cell list's own) are what the pass reads.
int pair_lj_many(const ofp_cl* cl, const ofp_vl* vl, int dtype, const void* pos_sorted, size_t stride,
                 int precision, const void* tab, const void* qtab, int lanes, size_t e_cap, int with_u, int cubic,
                 hipStream_t s, const void* sysgrid = nullptr, const void* kick = nullptr,
                 const void* quant = nullptr);

}  // namespace ofp

// Internal cell-list accessor shared by the pair and Verlet sources.
extern "C" int ofp_cl_dev_(const ofp_cl* cl, ofp::ClDev* d, ofp_cl_view* v);
// A device buffer of at least `bytes` owned by the cell list, for the pair
// kernels' per-particle records (its contents belong to the last caller).
extern "C" int ofp_cl_aux_(const ofp_cl* cl, size_t bytes, void** p);
// Work queue of target items of a built cell list (built on first use).
extern "C" int ofp_cl_items_(const ofp_cl* cl, const ofp::TargetItem** items,
                             const uint32_t** n_items, uint32_t** counter, size_t* cap,
                             hipStream_t s);
