// ofp_internal.hpp -- shared helpers of libofp_nn.so (gfx950 only).
//
// Error reporting, grow-only device buffers, the plane-copy argument block, the
// cell-key formula shared by every kernel, and the wave/block scan primitives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
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
// Grow-only device allocation owned by a handle.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap && p) return OFP_OK;
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
    if (p) (void)hipFree(p);
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

// Broadcast of lane k's value (k uniform) through v_readlane.
__device__ __forceinline__ float bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ double bcast(double v, int k) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffll), k);
  int hi = __builtin_amdgcn_readlane(int(b >> 32), k);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ __forceinline__ uint32_t bcast(uint32_t v, int k) {
  return uint32_t(__builtin_amdgcn_readlane(int(v), k));
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
struct TargetItem {
  uint32_t row0;
  uint32_t n;
  uint32_t cell0;  // linear id of the first cell
  uint32_t ncx;
};

// ------------------------------------------------- compensated sums
// Two-sum (Knuth): s + c carries the sum with the rounding error of every
// addition accumulated in c.
struct Comp32 {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    float t = s + x;
    float z = t - s;
    c += (s - (t - z)) + (x - z);
    s = t;
  }
  // merge another compensated sum (used by the cross-lane reductions)
  __device__ __forceinline__ void merge(float s2, float c2) {
    add(s2);
    c += c2;
  }
  __device__ __forceinline__ double value() const { return double(s) + double(c); }
};

}  // namespace ofp

// Internal cell-list accessor shared by the pair and Verlet sources.
extern "C" int ofp_cl_dev_(const ofp_cl* cl, ofp::ClDev* d, ofp_cl_view* v);
// Work queue of target items of a built cell list (built on first use).
extern "C" int ofp_cl_items_(const ofp_cl* cl, const ofp::TargetItem** items,
                             const uint32_t** n_items, uint32_t** counter, size_t* cap,
                             hipStream_t s);
