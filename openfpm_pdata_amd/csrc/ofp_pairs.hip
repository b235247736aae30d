// ofp_pairs.hip -- pair-interaction loops (LJ, SPH density, SPH pressure force)
// in plain FP64 and in FP32 on quantised cell-relative coordinates with FP64
// per-particle sums (FP32C), plus the Tait EOS and the FP64 reductions.
//
// Two walkers:
//  * Verlet-driven (vl != NULL): 16 lanes per row, 4 rows per wave; the lanes of
//    a row read 16 consecutive entries (coalesced index stream), gather the
//    neighbours, and the partial sums are combined across the 16 lanes at the
//    end.
//  * cell tiles (vl == NULL): persistent waves over the cell list's target
//    items (up to 64 domain particles of one x-row run of cells, one per lane);
//    the item's (2B+1)^2 stencil x-runs are staged through LDS and every lane
//    tests and evaluates its own target against them.
//
// FP32C arithmetic: x_q - x_p per dim is the exact int32 u_q - u_p + (c_q -
// c_p) 2^28 (quantum q_k = spacing_k / 2^28, see the cell-list build); all
// lengths are in units of q_0, so no rounded scale factor touches a distance.
// Every pair gets the exact square (integer products summed in FP64, split once
// into an FP32 hi/lo pair); the cut-off test compares its hi part.  Physical
// constants in q_0 units that enter a power are carried as (hi, lo) FP32 pairs.
// Each pair's FP32 factor times its FP64 separation (an exact product) is summed
// in FP64 (on this GPU an FP64 add or FMA issues at the rate of an FP32 one, so
// this costs less than a compensated FP32 sum), and the remaining scale factors
// are applied once per particle in FP64.
#include "ofp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

extern "C" int ofp_vl_dev_(const ofp_vl* h, const ofp_cl* cl, int symmetric, const unsigned long long** off,
                           const uint32_t** nbr);
extern "C" int ofp_vl_aux_(const ofp_vl* h, size_t bytes, void** p);

namespace ofp {

constexpr int kRowLanes = 16;
constexpr int kUnroll = 4;  // Verlet entries in flight per lane

// ----------------------------------------------------------------- walkers
// exact x_q - x_p in quanta from the wrapped per-particle coordinates
__device__ __forceinline__ int wrap_diff(int gq, int gp) { return int(uint32_t(gq) - uint32_t(gp)); }

// -------------------------------------------------- FP64 plain physics
template <class R>
__device__ __forceinline__ R w3poly(R q) {
  if (q < R(1)) return R(1) - R(1.5) * q * q + R(0.75) * q * q * q;
  R t = R(2) - q;
  return q < R(2) ? R(0.25) * t * t * t : R(0);
}
template <class R>
__device__ __forceinline__ R dw3poly(R q) {
  if (q < R(1)) return R(-3) * q + R(2.25) * q * q;
  R t = R(2) - q;
  return q < R(2) ? R(-0.75) * t * t : R(0);
}

// kU: also the per-particle energy and neighbour count (en / nn); the forms
// share one layout, so one constant table serves both.
template <bool kU>
struct LJ64T {
  double rc2, sig2, eps24, eps4;
  double* fx;
  double* fy;
  double* fz;
  double* en;
  uint32_t* nn;
  struct St {
    double fx = 0, fy = 0, fz = 0, u = 0;
    uint32_t n = 0;
  };
  __device__ void start(St&, uint32_t) const {}
  __device__ __forceinline__ double aux(uint32_t) const { return 0.0; }
  __device__ __forceinline__ void pair(St& s, double, double rx, double ry, double rz, double d2) const {
    if constexpr (kU) ++s.n;
    if (d2 == 0.0) return;
    double inv = 1.0 / d2;
    double s2 = sig2 * inv;
    double s6 = s2 * s2 * s2;
    double s12 = s6 * s6;
    double f = eps24 * (2.0 * s12 - s6) * inv;
    s.fx += f * rx;
    s.fy += f * ry;
    s.fz += f * rz;
    if constexpr (kU) s.u += eps4 * (s12 - s6);
  }
  __device__ __forceinline__ void reduce(St& s, int w) const {
    for (int d = w / 2; d > 0; d >>= 1) {
      s.fx += __shfl_xor(s.fx, d, w);
      s.fy += __shfl_xor(s.fy, d, w);
      s.fz += __shfl_xor(s.fz, d, w);
      if constexpr (kU) {
        s.u += __shfl_xor(s.u, d, w);
        s.n += __shfl_xor(s.n, d, w);
      }
    }
  }
  // the force finish stores, in the position dtype (what a kick reads back)
  template <class T>
  __device__ __forceinline__ void force_as(const St& s, T (&a)[3]) const {
    a[0] = T(s.fx);
    a[1] = T(s.fy);
    a[2] = T(s.fz);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    if (fx) {
      fx[p] = s.fx;
      fy[p] = s.fy;
      fz[p] = s.fz;
    }
    if constexpr (kU) {
      if (en) en[p] = s.u;
      if (nn) nn[p] = s.n;
    }
  }
};
using LJ64 = LJ64T<true>;

struct Dens64 {
  double rc2, inv_h, sk_m;  // sk_m = m / (pi h^3)
  double* rho;
  struct St {
    double acc = 0;
  };
  __device__ void start(St&, uint32_t) const {}
  __device__ __forceinline__ double aux(uint32_t) const { return 0.0; }
  __device__ __forceinline__ void pair(St& s, double, double, double, double, double d2) const {
    s.acc += w3poly(sqrt(d2) * inv_h);
  }
  __device__ __forceinline__ void reduce(St& s, int w) const {
    for (int d = w / 2; d > 0; d >>= 1) s.acc += __shfl_xor(s.acc, d, w);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const { rho[p] = sk_m * (s.acc + 1.0); }
};

struct Force64 {
  double rc2, inv_h, sd_m;  // sd_m = m / (pi h^4)
  const double* pterm;      // P / rho^2 per sorted particle
  double* fx;
  double* fy;
  double* fz;
  struct St {
    double fx = 0, fy = 0, fz = 0, pi = 0;
  };
  __device__ __forceinline__ void start(St& s, uint32_t p) const { s.pi = pterm[p]; }
  __device__ __forceinline__ double aux(uint32_t j) const { return pterm[j]; }
  __device__ __forceinline__ void pair(St& s, double pj, double rx, double ry, double rz, double d2) const {
    if (d2 == 0.0) return;
    double r = sqrt(d2);
    double coef = (s.pi + pj) * dw3poly(r * inv_h) / r;
    s.fx += coef * rx;
    s.fy += coef * ry;
    s.fz += coef * rz;
  }
  __device__ __forceinline__ void reduce(St& s, int w) const {
    for (int d = w / 2; d > 0; d >>= 1) {
      s.fx += __shfl_xor(s.fx, d, w);
      s.fy += __shfl_xor(s.fy, d, w);
      s.fz += __shfl_xor(s.fz, d, w);
    }
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    fx[p] = -sd_m * s.fx;
    fy[p] = -sd_m * s.fy;
    fz[p] = -sd_m * s.fz;
  }
};

// ------------------------------------------------- FP32C physics (q_0 units)
// FP32 storage (16-byte quantised records, FP32 P/rho^2 and forces); the pair
// is evaluated in FP64 from the FP64 separation.  On this GPU, with packed FP32
// off, an FP64 add, multiply or FMA issues at the rate of an FP32 one, and the
// FP64 reciprocal / reciprocal square root estimates at half the rate of the
// FP32 ones, so one FP64 Newton step on the estimate replaces the (hi, lo) FP32
// pairs an FP32 evaluation needed to stay near 1e-6: the pair costs about the
// same time (the kernels are bound by their gathers), and the error left is
// that of the FP32 inputs (P/rho^2, the quantum), near 1e-7 or below.
//
// One pair in q_0 units: x_p - x_q and its square in FP64 (integer components;
// for non-cubic cells times the FP64 axis ratio), and the square rounded to
// FP32 (d2), on which the cut-off test decides.
struct QPair {
  double rx, ry, rz, r2;
  float d2;
};

// 1/x and 1/sqrt(x) (x > 0) from the hardware estimates (about 2^-23 relative)
// and one Newton step each (about 2^-46).
__device__ __forceinline__ double rcp_refined(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double rsqrt_refined(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);  // 1 - x y^2
  return __builtin_fma(0.5 * y, e, y);
}

// Butterfly sum over groups of w lanes (every lane of a group ends with the
// same value: FP addition is commutative).
__device__ __forceinline__ void lane_sum(double& a, int w) {
  for (int d = w / 2; d > 0; d >>= 1) a += __shfl_xor(a, d, w);
}

// Cut-off decisions identical to the FP64 form.  The quantised square differs
// from the exact one by at most 2 sqrt(3) q_m r + 3 q_m^2 (q_m = the largest
// quantum, in q_0 units) and is then rounded to FP32, so a pair whose FP32
// square lies in [lo2, hi2) around the cut-off is decided again as the FP64
// kernels decide it: x_p - x_q per component in FP64 from the stored (sorted)
// positions, the squares summed unfused, compared with the FP64 form's r_c^2.
// Outside the band both forms agree.  A few pairs per million fall in the band,
// so the divergent path costs nothing measurable.
struct ExactCut {
  const void* pos;  // sorted positions, 3 planes pstride apart
  size_t pstride;
  int pos_f64;
  float lo2, hi2;  // band of the FP32 square (q_0 units)
  double rc2;      // r_c^2 in position units, the FP64 form's constant
  __device__ __forceinline__ bool in_band(float d2) const { return d2 >= lo2 && d2 < hi2; }
  __device__ bool hit(uint32_t p, uint32_t j) const {
    double dx, dy, dz;
    if (pos_f64) {
      const double* P = static_cast<const double*>(pos);
      dx = P[p] - P[j];
      dy = P[pstride + p] - P[pstride + j];
      dz = P[2 * pstride + p] - P[2 * pstride + j];
    } else {
      const float* P = static_cast<const float*>(pos);
      dx = double(P[p]) - double(P[j]);
      dy = double(P[pstride + p]) - double(P[pstride + j]);
      dz = double(P[2 * pstride + p]) - double(P[2 * pstride + j]);
    }
    return dist2_unfused<double>(dx, dy, dz) < rc2;
  }
  // the same decision from the two stored positions (dtype T), loaded by the caller
  template <class T>
  __device__ __forceinline__ bool hit_from(const T (&a)[3], const T (&b)[3]) const {
    const double dx = double(a[0]) - double(b[0]);
    const double dy = double(a[1]) - double(b[1]);
    const double dz = double(a[2]) - double(b[2]);
    return dist2_unfused<double>(dx, dy, dz) < rc2;
  }
};

// kU as for LJ64T.
template <bool kU>
struct LJ32T {
  static constexpr bool kExact = true;  // the (sigma/r)^12 chain needs the exact square
  static constexpr bool kRefine = true;
  static constexpr bool kAuxInQuant = false;
  float rc2;      // (r_c / q0)^2
  double sig2;    // (sigma / q0)^2
  double fscale;  // 24 eps / q0
  double uscale;  // 4 eps
  float* fx;
  float* fy;
  float* fz;
  float* en;
  uint32_t* nn;
  ExactCut xc;
  // Close pairs (r < sigma, where the repulsion is steep and strong): the
  // quantised separation (error < 1 quantum) is replaced by the FP64
  // difference of the stored positions, since the r^-13 repulsion multiplies
  // its relative error by 13 and a pair force there exceeds 24 eps / sigma.
  double inv_q0;  // 1 / q0
  float close2;   // (sigma / q0)^2
  struct St {
    double fx = 0, fy = 0, fz = 0, u = 0;
    uint32_t n = 0;
  };
  // The refined separation from the two stored positions, loaded by the
  // caller (same arithmetic as refine), so that a group of pairs can issue all
  // its loads before the first use.
  template <class T>
  __device__ __forceinline__ void refine_from(QPair& r, const T (&a)[3], const T (&b)[3]) const {
    const double dx = (double(a[0]) - double(b[0])) * inv_q0;
    const double dy = (double(a[1]) - double(b[1])) * inv_q0;
    const double dz = (double(a[2]) - double(b[2])) * inv_q0;
    r.rx = dx;
    r.ry = dy;
    r.rz = dz;
    r.r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
  }
  __device__ void refine(QPair& r, uint32_t p, uint32_t j) const {
    double dx, dy, dz;
    if (xc.pos_f64) {
      const double* P = static_cast<const double*>(xc.pos);
      dx = P[p] - P[j];
      dy = P[xc.pstride + p] - P[xc.pstride + j];
      dz = P[2 * xc.pstride + p] - P[2 * xc.pstride + j];
    } else {
      const float* P = static_cast<const float*>(xc.pos);
      dx = double(P[p]) - double(P[j]);
      dy = double(P[xc.pstride + p]) - double(P[xc.pstride + j]);
      dz = double(P[2 * xc.pstride + p]) - double(P[2 * xc.pstride + j]);
    }
    dx *= inv_q0;
    dy *= inv_q0;
    dz *= inv_q0;
    r.rx = dx;
    r.ry = dy;
    r.rz = dz;
    r.r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
  }
  __device__ void start(St&, uint32_t) const {}
  __device__ __forceinline__ float aux(uint32_t) const { return 0.f; }
  // (sigma/r)^6 and ^12 and the force factor (2 (sigma/r)^12 - (sigma/r)^6) / r^2
  // from 1/r^2 in FP64.
  __device__ __forceinline__ void eval(double r2, double* f, double* u) const {
    const double inv = rcp_refined(r2);
    const double s2 = sig2 * inv;
    const double s6 = s2 * s2 * s2;
    const double s12 = s6 * s6;
    *f = __builtin_fma(2.0, s12, -s6) * inv;
    *u = s12 - s6;
  }
  __device__ __forceinline__ void pair(St& s, float, const QPair& r) const {
    if constexpr (kU) ++s.n;
    if (r.r2 == 0.0) return;
    double f, u;
    eval(r.r2, &f, &u);
    s.fx = __builtin_fma(f, r.rx, s.fx);
    s.fy = __builtin_fma(f, r.ry, s.fy);
    s.fz = __builtin_fma(f, r.rz, s.fz);
    if constexpr (kU) s.u += u;
  }
  __device__ __forceinline__ void pair_masked(St& s, float, const QPair& r, bool hit) const {
    if constexpr (kU) s.n += hit ? 1u : 0u;
    const bool live = hit && r.r2 != 0.0;
    // a miss is evaluated at r = sigma (every power finite) and weighted by 0
    // (a product rather than a select keeps the evaluation branch-free)
    double f, u;
    eval(live ? r.r2 : sig2, &f, &u);
    const double m = live ? 1.0 : 0.0;
    f *= m;
    s.fx = __builtin_fma(f, r.rx, s.fx);
    s.fy = __builtin_fma(f, r.ry, s.fy);
    s.fz = __builtin_fma(f, r.rz, s.fz);
    if constexpr (kU) s.u = __builtin_fma(u, m, s.u);
  }
  __device__ __forceinline__ void reduce(St& s, int w) const {
    lane_sum(s.fx, w);
    lane_sum(s.fy, w);
    lane_sum(s.fz, w);
    if constexpr (kU) {
      lane_sum(s.u, w);
      for (int d = w / 2; d > 0; d >>= 1) s.n += __shfl_xor(s.n, d, w);
    }
  }
  template <class T>
  __device__ __forceinline__ void force_as(const St& s, T (&a)[3]) const {
    a[0] = T(float(fscale * s.fx));
    a[1] = T(float(fscale * s.fy));
    a[2] = T(float(fscale * s.fz));
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    if (fx) {
      fx[p] = float(fscale * s.fx);
      fy[p] = float(fscale * s.fy);
      fz[p] = float(fscale * s.fz);
    }
    if constexpr (kU) {
      if (en) en[p] = float(uscale * s.u);
      if (nn) nn[p] = s.n;
    }
  }
};
using LJ32 = LJ32T<true>;

// Cubic-spline argument q = r/h in FP64 from the FP64 square (q_0 units): r and
// 1/r from the refined reciprocal square root; r2 = 0 (coincident particles)
// gives q = 0.
struct SplineQ {
  double qh;  // q_0 / h
  __device__ __forceinline__ double q_and_inv(double r2, double* inv_r) const {
    const bool pos = r2 > 0.0;
    const double y = rsqrt_refined(pos ? r2 : 1.0);
    *inv_r = y;
    return (pos ? r2 * y : 0.0) * qh;
  }
  // W(q) and dW/dq of the cubic spline without its normalisation
  __device__ __forceinline__ static double w(double q) {
    const double t = 2.0 - q;
    return q < 1.0 ? __builtin_fma(q * q, __builtin_fma(0.75, q, -1.5), 1.0) : (t > 0.0 ? 0.25 * t * t * t : 0.0);
  }
  __device__ __forceinline__ static double dw(double q) {
    const double t = 2.0 - q;
    return q < 1.0 ? q * __builtin_fma(2.25, q, -3.0) : (t > 0.0 ? -0.75 * t * t : 0.0);
  }
};

struct Dens32 {
  static constexpr bool kExact = true;
  static constexpr bool kRefine = false;
  static constexpr bool kAuxInQuant = false;
  float rc2;   // (2h / q0)^2
  ExactCut xc;
  SplineQ sp;
  double sk_m;  // m / (pi h^3)
  double* rho;  // output in FP64
  struct St {
    double acc = 0;
  };
  __device__ void start(St&, uint32_t) const {}
  __device__ __forceinline__ float aux(uint32_t) const { return 0.f; }
  __device__ __forceinline__ void pair(St& s, float, const QPair& rp) const {
    double inv_r;
    s.acc += SplineQ::w(sp.q_and_inv(rp.r2, &inv_r));
  }
  __device__ __forceinline__ void pair_masked(St& s, float, const QPair& rp, bool hit) const {
    // W is finite at every argument, so a miss adds W * 0 = 0 (a product, not a
    // select: the compiler keeps the evaluation unconditional, without branches)
    double inv_r;
    s.acc = __builtin_fma(SplineQ::w(sp.q_and_inv(rp.r2, &inv_r)), hit ? 1.0 : 0.0, s.acc);
  }
  __device__ __forceinline__ void reduce(St& s, int w) const { lane_sum(s.acc, w); }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const { rho[p] = sk_m * (s.acc + 1.0); }
};

struct Force32 {
  static constexpr bool kExact = true;
  static constexpr bool kRefine = false;
  // Verlet rows gather (q_x, q_y, q_z, pterm) as one 16-byte record per entry
  // from qaux instead of the quantised stream and pterm separately.
  static constexpr bool kAuxInQuant = true;
  float rc2;
  ExactCut xc;
  SplineQ sp;
  double fscale;       // -m / (pi h^4)
  const float* pterm;  // P / rho^2
  const int4* qaux;    // sorted: quantised position with pterm as the 4th word
  float* fx;
  float* fy;
  float* fz;
  struct St {
    double fx = 0, fy = 0, fz = 0;
    double pi = 0;
  };
  __device__ __forceinline__ void start(St& s, uint32_t p) const { s.pi = pterm[p]; }
  __device__ __forceinline__ float aux(uint32_t j) const { return pterm[j]; }
  // (P_p/rho_p^2 + P_q/rho_q^2) dW/dq / r; coincident particles (r2 = 0) give 0
  __device__ __forceinline__ double coef(const St& s, float pj, double r2) const {
    double inv_r;
    const double q = sp.q_and_inv(r2, &inv_r);
    return (s.pi + double(pj)) * SplineQ::dw(q) * inv_r;
  }
  __device__ __forceinline__ void pair(St& s, float pj, const QPair& rp) const {
    if (rp.r2 == 0.0) return;
    const double c = coef(s, pj, rp.r2);
    s.fx = __builtin_fma(c, rp.rx, s.fx);
    s.fy = __builtin_fma(c, rp.ry, s.fy);
    s.fz = __builtin_fma(c, rp.rz, s.fz);
  }
  __device__ __forceinline__ void pair_masked(St& s, float pj, const QPair& rp, bool hit) const {
    // the factor is finite at every argument, so a miss contributes factor * 0;
    // a product rather than a select keeps the evaluation branch-free
    const double c = coef(s, pj, rp.r2) * (hit ? 1.0 : 0.0);
    s.fx = __builtin_fma(c, rp.rx, s.fx);
    s.fy = __builtin_fma(c, rp.ry, s.fy);
    s.fz = __builtin_fma(c, rp.rz, s.fz);
  }
  __device__ __forceinline__ void reduce(St& s, int w) const {
    lane_sum(s.fx, w);
    lane_sum(s.fy, w);
    lane_sum(s.fz, w);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    fx[p] = float(fscale * s.fx);
    fy[p] = float(fscale * s.fy);
    fz[p] = float(fscale * s.fz);
  }
};

// x_p - x_q in q_0 units; exact squared distance for hits.
struct QGeom {
  float ry, rz;       // q_y / q_0, q_z / q_0 (1 for cubic cells)
  double ryd, rzd;
};

// A pair gets x_p - x_q in FP64 (integer components; for non-cubic cells times
// the FP64 axis ratio, rounded once) and its square in FP64 (products summed
// with FMAs); the cut-off test compares the square rounded to FP32 (physics
// without kExact: the FP32 square of the rounded components), and a square in
// the ExactCut band is decided on the stored positions as the FP64 form decides.
// The cell walk first drops candidates whose FP32 square exceeds the cut-off by
// more than any rounding and the band (pre_cut), then decides like the Verlet
// rows.
// kCubic: cubic cells (q_y = q_z = q_0), no component scaling.
// Physics with close-pair refinement (LJ) replaces the quantised separation of
// a hit closer than close2 by the exact one from the stored positions.
constexpr float kPreCut = 1.0f + 1.0f / 65536.0f;

template <class Ph, bool kCubic>
__device__ __forceinline__ QPair quant_geom(const QGeom& qg, int ix, int iy, int iz, float* d2f) {
  QPair p;
  if constexpr (Ph::kExact) {
    const double dx = double(ix);
    const double dy = kCubic ? double(iy) : double(iy) * qg.ryd;
    const double dz = kCubic ? double(iz) : double(iz) * qg.rzd;
    const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
    p.rx = -dx;
    p.ry = -dy;
    p.rz = -dz;
    p.r2 = r2;
    p.d2 = float(r2);
    *d2f = p.d2;
  } else {
    const float fx = float(ix);
    const float fy = kCubic ? float(iy) : float(iy) * qg.ry;
    const float fz = kCubic ? float(iz) : float(iz) * qg.rz;
    p.rx = -double(fx);
    p.ry = -double(fy);
    p.rz = -double(fz);
    p.d2 = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
    p.r2 = p.d2;
    *d2f = p.d2;
  }
  return p;
}
template <class Ph>
__device__ __forceinline__ void refine_close(const Ph& ph, QPair& p, bool hit, uint32_t ip, uint32_t jq) {
  if constexpr (Ph::kRefine) {
    if (hit && p.d2 < ph.close2) ph.refine(p, ip, jq);  // rare, divergent
  }
}

template <class Ph, bool kCubic>
__device__ __forceinline__ void quant_pair(const Ph& ph, typename Ph::St& st, const QGeom& qg, float aux,
                                           int ix, int iy, int iz, uint32_t ip, uint32_t jq) {
  const float fx = float(ix);
  const float fy = kCubic ? float(iy) : float(iy) * qg.ry;
  const float fz = kCubic ? float(iz) : float(iz) * qg.rz;
  const float pre = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
  if (pre < ph.rc2 * kPreCut) {
    float d2;
    QPair p = quant_geom<Ph, kCubic>(qg, ix, iy, iz, &d2);
    bool hit = d2 < ph.rc2;
    if (ph.xc.in_band(d2)) hit = ph.xc.hit(ip, jq);  // rare, divergent
    if (hit) {
      refine_close(ph, p, true, ip, jq);
      ph.pair(st, aux, p);
    }
  }
}

// Branch-free variant for Verlet rows (most entries are hits): the pair is
// evaluated for every entry with a safe distance and its contributions are
// multiplied by the hit mask.  j == p enters with hit = false.
template <class Ph, bool kCubic>
__device__ __forceinline__ void quant_pair_masked(const Ph& ph, typename Ph::St& st, const QGeom& qg, float aux,
                                                  int ix, int iy, int iz, bool not_self, uint32_t ip,
                                                  uint32_t jq) {
  float d2;
  QPair p = quant_geom<Ph, kCubic>(qg, ix, iy, iz, &d2);
  bool hit = not_self && d2 < ph.rc2;
  if (not_self && ph.xc.in_band(d2)) hit = ph.xc.hit(ip, jq);  // rare, divergent
  refine_close(ph, p, hit, ip, jq);
  ph.pair_masked(st, aux, p, hit);
}

// Four Verlet entries of one row, for physics with close-pair refinement, in
// three passes: the decisions of all four from their quantised separations
// (hit, in the ExactCut band, possibly close); the stored positions (sref,
// dtype T, the positions the records were made from) of the entries that need
// them gathered together, so a group waits for
// memory once rather than once per such pair; then each entry decided and
// evaluated in turn, its separation recomputed from the records (only the
// decisions and the gathered positions stay live across the passes, which
// keeps down the register count, and with it the waves per SIMD that hide the
// gathers).  Same results as four quant_pair_masked calls.  pp: the row
// particle's stored position.
template <class T, class Ph, bool kCubic>
__device__ __forceinline__ void quant_group_masked(const Ph& ph, typename Ph::St& st, const QGeom& qg,
                                                   const float (&aux)[4], const int4 (&Q)[4], const int4& P,
                                                   const uint32_t (&j)[4], const bool (&not_self)[4],
                                                   const T (&pp)[3], const T* __restrict__ sref, size_t sstride) {
  static_assert(Ph::kRefine, "grouped evaluation is for physics with close-pair refinement");
  bool hit[4], band[4], need[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float d2;
    const QPair r =
        quant_geom<Ph, kCubic>(qg, wrap_diff(Q[u].x, P.x), wrap_diff(Q[u].y, P.y), wrap_diff(Q[u].z, P.z), &d2);
    hit[u] = not_self[u] && d2 < ph.rc2;
    band[u] = not_self[u] && ph.xc.in_band(d2);
    // close pairs are refined; a band pair's final decision may still be a hit
    need[u] = band[u] || (hit[u] && r.d2 < ph.close2);
  }
  T xq[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (need[u]) {
      xq[u][0] = sref[j[u]];
      xq[u][1] = sref[sstride + j[u]];
      xq[u][2] = sref[2 * sstride + j[u]];
    }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (band[u]) hit[u] = ph.xc.hit_from(pp, xq[u]);  // rare
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int ix = wrap_diff(Q[u].x, P.x), iy = wrap_diff(Q[u].y, P.y), iz = wrap_diff(Q[u].z, P.z);
    // fresh values for the compiler: the separation is recomputed here rather
    // than kept from the first pass
    asm volatile("" : "+v"(ix), "+v"(iy), "+v"(iz));
    float d2;
    QPair r = quant_geom<Ph, kCubic>(qg, ix, iy, iz, &d2);
    if (hit[u] && r.d2 < ph.close2) ph.refine_from(r, pp, xq[u]);
    ph.pair_masked(st, aux[u], r, hit[u]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// One lane per row: close pairs are not refined inside the loop (where one
// lane's close pair costs every lane of the wave a second memory round trip
// per group) but set aside, up to kDefer of them, and evaluated after the
// row's last group, their stored positions gathered together.  A row with more
// close pairs refines the others in place.  The row's sum takes its terms in
// another order than quant_group_masked's; the pairs and their separations
// are the same.
constexpr int kDefer = 4;
struct Deferred {
  uint32_t j[kDefer];
  float a[kDefer];
  int n = 0;
  // j[0] is the newest (a shift register: no dynamic indexing of registers)
  __device__ __forceinline__ void push(uint32_t jj, float aa) {
#pragma unroll
    for (int k = kDefer - 1; k > 0; --k) {
      j[k] = j[k - 1];
      a[k] = a[k - 1];
    }
    j[0] = jj;
    a[0] = aa;
    ++n;
  }
};

template <class T, class Ph, bool kCubic>
__device__ __forceinline__ void quant_group_defer(const Ph& ph, typename Ph::St& st, const QGeom& qg,
                                                  const float (&aux)[4], const int4 (&Q)[4], const int4& P,
                                                  const uint32_t (&j)[4], const bool (&not_self)[4],
                                                  const T (&pp)[3], const T* __restrict__ sref, size_t sstride,
                                                  Deferred& df) {
  static_assert(Ph::kRefine, "deferred refinement is for physics with close-pair refinement");
  bool hit[4], need[4], close[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float d2;
    const QPair r =
        quant_geom<Ph, kCubic>(qg, wrap_diff(Q[u].x, P.x), wrap_diff(Q[u].y, P.y), wrap_diff(Q[u].z, P.z), &d2);
    hit[u] = not_self[u] && d2 < ph.rc2;
    const bool band = not_self[u] && ph.xc.in_band(d2);
    close[u] = hit[u] && !band && r.d2 < ph.close2;
    if (close[u] && df.n < kDefer) {
      df.push(j[u], aux[u]);
      hit[u] = false;
      close[u] = false;
    }
    need[u] = band || close[u];  // rare: the band, or a row past kDefer close pairs
  }
  T xq[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (need[u]) {
      xq[u][0] = sref[j[u]];
      xq[u][1] = sref[sstride + j[u]];
      xq[u][2] = sref[2 * sstride + j[u]];
      if (!close[u]) hit[u] = ph.xc.hit_from(pp, xq[u]);  // the band decision
    }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int ix = wrap_diff(Q[u].x, P.x), iy = wrap_diff(Q[u].y, P.y), iz = wrap_diff(Q[u].z, P.z);
    asm volatile("" : "+v"(ix), "+v"(iy), "+v"(iz));  // recomputed, not kept from the first pass
    float d2;
    QPair r = quant_geom<Ph, kCubic>(qg, ix, iy, iz, &d2);
    if (close[u]) ph.refine_from(r, pp, xq[u]);
    ph.pair_masked(st, aux[u], r, hit[u]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The set-aside close pairs of a row: positions gathered together, then each
// evaluated from the refined separation.
template <class T, class Ph>
__device__ __forceinline__ void flush_deferred(const Ph& ph, typename Ph::St& st, const T (&pp)[3],
                                               const T* __restrict__ sref, size_t sstride, const Deferred& df) {
  if (df.n == 0) return;
  T xq[kDefer][3];
#pragma unroll
  for (int k = 0; k < kDefer; ++k)
    if (k < df.n) {
      xq[k][0] = sref[df.j[k]];
      xq[k][1] = sref[sstride + df.j[k]];
      xq[k][2] = sref[2 * sstride + df.j[k]];
    }
#pragma unroll
  for (int k = 0; k < kDefer; ++k)
    if (k < df.n) {
      QPair r;
      ph.refine_from(r, pp, xq[k]);
      ph.pair(st, df.a[k], r);
    }
}

// ---------------------------------------------------------- FP64 kernels
template <class T, class Ph>
__global__ __launch_bounds__(256) void k_pairs_f64(ClDev c, const T* __restrict__ pos,
                                                   size_t stride, const unsigned long long* __restrict__ voff,
                                                   const uint32_t* __restrict__ vnbr, Ph ph) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = tid / kRowLanes;
  const bool valid = k < c.n_dom;
  const uint32_t p = c.dom[valid ? k : 0];
  const T tx = pos[p], ty = pos[stride + p], tz = pos[2 * stride + p];
  const double xp = tx, yp = ty, zp = tz;
  typename Ph::St st;
  ph.start(st, p);
  // x_p - x_q; unfused square, as the oracle, so the neighbour decision is
  // bit-identical
  auto body = [&](uint32_t j, T qx, T qy, T qz, double a) {
    if (j == p) return;
    double rx = xp - double(qx);
    double ry = yp - double(qy);
    double rz = zp - double(qz);
    double d2 = dist2_unfused<double>(rx, ry, rz);
    if (d2 < ph.rc2) ph.pair(st, a, rx, ry, rz, d2);
  };
  const unsigned lane = threadIdx.x & (kRowLanes - 1);
  if (valid) {
    const unsigned long long e1 = voff[k + 1];
    for (unsigned long long e = voff[k] + lane; e < e1; e += kUnroll * kRowLanes) {
      uint32_t j[kUnroll];
      T qx[kUnroll], qy[kUnroll], qz[kUnroll];
      double a[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const unsigned long long ee = e + u * kRowLanes;
        j[u] = ee < e1 ? vnbr[ee] : p;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        qx[u] = pos[j[u]];
        qy[u] = pos[stride + j[u]];
        qz[u] = pos[2 * stride + j[u]];
        a[u] = ph.aux(j[u]);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) body(j[u], qx[u], qy[u], qz[u], a[u]);
    }
  }
  ph.reduce(st, kRowLanes);
  if (valid && lane == 0) ph.finish(st, p);
}

// --------------------------------------------------------- FP32C kernels
template <class Ph, bool kCubic>
__global__ __launch_bounds__(256) void k_pairs_f32c(ClDev c, QGeom qg, const unsigned long long* __restrict__ voff,
                                                    const uint32_t* __restrict__ vnbr, Ph ph) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = tid / kRowLanes;
  const bool valid = k < c.n_dom;
  const uint32_t p = c.dom[valid ? k : 0];
  const int4 P = c.quant[p];
  typename Ph::St st;
  ph.start(st, p);
  const unsigned lane = threadIdx.x & (kRowLanes - 1);
  if (valid) {
    const unsigned long long e1 = voff[k + 1];
    // the row's entries of the next step are loaded before the current step's
    // gathers are used, so the row stream's latency hides behind the pairs.
    // The loads are unconditional, so the compiler can count them in its
    // waits: past the row end a lane reads the row's last entry (entry 0 when
    // the list's first row is empty; the buffer holds at least one entry) and
    // its pair is masked off.
    const unsigned long long elast = e1 > 0 ? e1 - 1 : 0;
    uint32_t jn[kUnroll];
    bool okn[kUnroll];
    auto load_j = [&](unsigned long long e) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const unsigned long long ee = e + u * kRowLanes;
        okn[u] = ee < e1;
        jn[u] = vnbr[okn[u] ? ee : elast];
      }
    };
    load_j(voff[k] + lane);
    for (unsigned long long e = voff[k] + lane; e < e1; e += kUnroll * kRowLanes) {
      uint32_t j[kUnroll];
      bool ok[kUnroll];
      int4 Q[kUnroll];
      float a[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        j[u] = jn[u];
        ok[u] = okn[u] && jn[u] != p;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if constexpr (Ph::kAuxInQuant) {
          Q[u] = ph.qaux[j[u]];
          a[u] = __int_as_float(Q[u].w);
        } else {
          Q[u] = c.quant[j[u]];
          a[u] = ph.aux(j[u]);
        }
      }
      load_j(e + kUnroll * kRowLanes);
      // one entry's evaluation at a time (its FP64 temporaries are what sets
      // the kernel's register count, and with it the waves per SIMD that
      // hide the gathers' latency)
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        quant_pair_masked<Ph, kCubic>(ph, st, qg, a[u], wrap_diff(Q[u].x, P.x), wrap_diff(Q[u].y, P.y),
                                      wrap_diff(Q[u].z, P.z), ok[u], p, j[u]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  ph.reduce(st, kRowLanes);
  if (valid && lane == 0) ph.finish(st, p);
}

// ---------------------------------------------- LDS-staged cell tiles
// The pair pass over a cell list without a Verlet list.  Persistent waves take
// target items from the cell list's item queue: up to 64 domain particles of
// consecutive cells on one x-row of the grid, one per lane.  The item's
// candidates are the (2B+1)^2 x-runs of its stencil, each a contiguous range
// of sorted indices; they pass through the wave's LDS region kTileChunk at a
// time (coalesced loads).  Check: every lane tests its own target against every
// staged candidate (one broadcast LDS read per candidate), appending one bit
// per candidate to its hit mask (the compare's lane mask as v_addc's carry).
// Evaluate: each lane walks its own set bits in ascending order and evaluates
// those pairs from the staged records.  A target's pairs are summed in one
// lane in ascending sorted index, so the results do not depend on the order in
// which the items are taken.  Candidates that lie in the union of the item's
// stencils but outside a target's own stencil are more than B cells away in x
// from it, so they are misses for it.
//
// FP32C separations: the union spans at most kItemMaxCells - 1 + B cells in x
// from any of the item's targets, so a candidate is either within 2^31 quanta
// (the wrapped subtraction is exact) or its wrapped separation is still more
// than B cells (one cell of slack for positions that have drifted since the
// cells were built): no far candidate passes the check.
constexpr int kTileChunk = 128;  // candidates staged per round (4 mask words)
constexpr int kTileBlock = 256;  // four independent waves
constexpr int kTileWaves = kTileBlock / kWave;
static_assert(16 - (int(kItemMaxCells) - 1 + kMaxBoxNN) - 2 > kMaxBoxNN, "wrapped FP32C separations");

// Hit bits of staged candidates k0 .. k0 + cnt - 1 (cnt <= 32, uniform) for
// this lane's target, candidate k0 + k at bit k; `test(Q)` is the lane's check.
template <class R, class F>
__device__ __forceinline__ uint32_t staged_hits(const R* buf, uint32_t k0, uint32_t cnt, F&& test) {
  if (cnt == 0) return 0u;
  uint32_t w = 0, k = 0;
  for (; k + 8 <= cnt; k += 8) {
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) w = shift_in(w, test(buf[k0 + k + i]));
  }
  for (; k < cnt; ++k) w = shift_in(w, test(buf[k0 + k]));
  return __builtin_bitreverse32(w << (32u - cnt));
}

// Calls fn(k) for the set bits k of the 128-bit mask (m[0] bits 0..31, ...),
// lowest first, one bit per lane per round, until no lane has bits left.
template <class F>
__device__ __forceinline__ void for_each_hit(uint32_t m[4], F&& fn) {
  unsigned long long lo = (static_cast<unsigned long long>(m[1]) << 32) | m[0];
  unsigned long long hi = (static_cast<unsigned long long>(m[3]) << 32) | m[2];
  while (__ballot((lo | hi) != 0ull) != 0ull) {
    if ((lo | hi) != 0ull) {
      uint32_t k;
      if (lo != 0ull) {
        k = uint32_t(__builtin_ctzll(lo));
        lo &= lo - 1ull;
      } else {
        k = 64u + uint32_t(__builtin_ctzll(hi));
        hi &= hi - 1ull;
      }
      fn(k);
    }
  }
}

// Staging: records j0 .. j0 + m - 1 into the wave's buffer.  The wave's LDS
// operations execute in issue order; the barriers keep the compiler from
// moving the reads of one round across the writes of the next.
template <class R>
__device__ __forceinline__ void stage(R* buf, const R* __restrict__ src, uint32_t j0, uint32_t m, int lane) {
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k = uint32_t(lane); k < m; k += kWave) buf[k] = src[j0 + k];
  __builtin_amdgcn_wave_barrier();
}

// FP32C.  rec: per sorted particle the quantised record; with Ph::kAuxInQuant
// its fourth word is the physics' per-particle value (P / rho^2), otherwise
// the fourth word is not read.  kMany: ensembles, the physics and geometry of
// an item come from the tables of its system (ncx >> 8).
template <class T, class Ph, bool kCubic, bool kMany>
__global__ __launch_bounds__(kTileBlock) void k_tile_f32c(ClDev c, QGeom qg1, Ph ph1, const Ph* __restrict__ tab,
                                                          const QGeom* __restrict__ qtab,
                                                          const SysGrid<T>* __restrict__ sys,
                                                          const int4* __restrict__ rec,
                                                          const TargetItem* __restrict__ items,
                                                          const uint32_t* __restrict__ n_items_p,
                                                          uint32_t* __restrict__ counter) {
  __shared__ int4 lds[kTileWaves][kTileChunk];
  const int lane = threadIdx.x & (kWave - 1);
  int4* buf = lds[threadIdx.x / kWave];
  const uint32_t n_items = *n_items_p;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const TargetItem ti = items[it];
    const uint32_t sy = ti.ncx >> 8;
    const Ph ph = kMany ? tab[sy] : ph1;
    const QGeom qg = kMany ? qtab[sy] : qg1;
    const ItemGeom<T> G = item_geom(c, ti, T(0), kMany ? sys : nullptr, static_cast<const T*>(nullptr));
    const bool valid = uint32_t(lane) < ti.n;
    const uint32_t p = c.dom[ti.row0 + (valid ? uint32_t(lane) : 0u)];
    const int4 P = c.quant[p];
    const float pre2 = ph.rc2 * kPreCut;
    typename Ph::St st;
    ph.start(st, p);
    auto test = [&](const int4& Q) {
      const float fx = float(wrap_diff(Q.x, P.x));
      const float fy = kCubic ? float(wrap_diff(Q.y, P.y)) : float(wrap_diff(Q.y, P.y)) * qg.ry;
      const float fz = kCubic ? float(wrap_diff(Q.z, P.z)) : float(wrap_diff(Q.z, P.z)) * qg.rz;
      return __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)) < pre2;
    };
    for_each_run(c, G, [&](uint32_t ja, uint32_t jb) {
      for (uint32_t j0 = ja; j0 < jb; j0 += kTileChunk) {
        const uint32_t m = jb - j0 < uint32_t(kTileChunk) ? jb - j0 : uint32_t(kTileChunk);
        stage(buf, rec, j0, m, lane);
        uint32_t mk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t k0 = uint32_t(q) * 32u;
          mk[q] = k0 < m ? staged_hits(buf, k0, m - k0 < 32u ? m - k0 : 32u, test) : 0u;
          if (!valid) mk[q] = 0u;
        }
        for_each_hit(mk, [&](uint32_t k) {
          const uint32_t j = j0 + k;
          if (j == p) return;
          const int4 Q = buf[k];
          float a;
          if constexpr (Ph::kAuxInQuant)
            a = __int_as_float(Q.w);
          else
            a = ph.aux(j);
          quant_pair<Ph, kCubic>(ph, st, qg, a, wrap_diff(Q.x, P.x), wrap_diff(Q.y, P.y), wrap_diff(Q.z, P.z), p,
                                 j);
        });
      }
    });
    if (valid) ph.finish(st, p);
  }
}

// Plain FP64 form: positions (three planes) staged per round; the check is the
// pair's own cut-off decision (FP64 separation of the stored positions, squares
// summed unfused), so a set bit is a pair to evaluate.
template <class T>
struct Pos3 {
  T x, y, z;
};

template <class T, class Ph, bool kMany>
__global__ __launch_bounds__(kTileBlock) void k_tile_f64(ClDev c, const T* __restrict__ pos, size_t stride,
                                                         Ph ph1, const Ph* __restrict__ tab,
                                                         const SysGrid<T>* __restrict__ sys,
                                                         const TargetItem* __restrict__ items,
                                                         const uint32_t* __restrict__ n_items_p,
                                                         uint32_t* __restrict__ counter) {
  __shared__ Pos3<T> lds[kTileWaves][kTileChunk];
  const int lane = threadIdx.x & (kWave - 1);
  Pos3<T>* buf = lds[threadIdx.x / kWave];
  const uint32_t n_items = *n_items_p;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const TargetItem ti = items[it];
    const Ph ph = kMany ? tab[ti.ncx >> 8] : ph1;
    const ItemGeom<T> G = item_geom(c, ti, T(0), kMany ? sys : nullptr, static_cast<const T*>(nullptr));
    const bool valid = uint32_t(lane) < ti.n;
    const uint32_t p = c.dom[ti.row0 + (valid ? uint32_t(lane) : 0u)];
    const double xp = pos[p], yp = pos[stride + p], zp = pos[2 * stride + p];
    typename Ph::St st;
    ph.start(st, p);
    auto test = [&](const Pos3<T>& Q) {
      return dist2_unfused<double>(xp - double(Q.x), yp - double(Q.y), zp - double(Q.z)) < ph.rc2;
    };
    for_each_run(c, G, [&](uint32_t ja, uint32_t jb) {
      for (uint32_t j0 = ja; j0 < jb; j0 += kTileChunk) {
        const uint32_t m = jb - j0 < uint32_t(kTileChunk) ? jb - j0 : uint32_t(kTileChunk);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = uint32_t(lane); k < m; k += kWave) {
          const uint32_t j = j0 + k;
          buf[k] = Pos3<T>{pos[j], pos[stride + j], pos[2 * stride + j]};
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t mk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t k0 = uint32_t(q) * 32u;
          mk[q] = k0 < m ? staged_hits(buf, k0, m - k0 < 32u ? m - k0 : 32u, test) : 0u;
          if (!valid) mk[q] = 0u;
        }
        for_each_hit(mk, [&](uint32_t k) {
          const uint32_t j = j0 + k;
          if (j == p) return;
          const Pos3<T> Q = buf[k];
          const double rx = xp - double(Q.x), ry = yp - double(Q.y), rz = zp - double(Q.z);
          ph.pair(st, ph.aux(j), rx, ry, rz, dist2_unfused<double>(rx, ry, rz));
        });
      }
    });
    if (valid) ph.finish(st, p);
  }
}

// ------------------------------------------------------ ensemble kernels
// One lane per Verlet row: the row's entries [e0, e1) are read in groups of
// four from the row's first entry, one 16-byte load per group (4-byte aligned;
// gfx950 takes dword-aligned multi-dword loads), so a wave's entry loads touch
// a quarter of the cache lines that four 4-byte loads per lane would and a row
// of L entries takes ceil(L / 4) groups.  The list buffer holds at least three
// entries past the list's capacity, so the last group read stays inside it.
// Slots of the last group past the row's end (the next row's entries, or
// unwritten ones past the list's end) read as p, the row's own particle, which
// the pair evaluation skips.  body(j, next) gets the group and a callable that
// issues the next group's load; body calls it once its gathers are issued, so
// that load is in flight during the evaluation.
template <class F>
__device__ __forceinline__ void row_groups(const uint32_t* __restrict__ vnbr, unsigned long long e0,
                                           unsigned long long e1, uint32_t p, F&& body) {
  auto load = [&](unsigned long long e) {
    uint4 q;
    __builtin_memcpy(&q, vnbr + e, sizeof(q));
    return q;
  };
  uint4 Jn = load(e0 < e1 ? e0 : e1);
  for (unsigned long long b = e0; b < e1; b += 4) {
    const uint32_t raw[4] = {Jn.x, Jn.y, Jn.z, Jn.w};
    uint32_t j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = b + u < e1 ? raw[u] : p;
    body(j, [&] { Jn = load(b + 4 < e1 ? b + 4 : b); });
  }
}

// The systems of a wave's rows one at a time, so that a system's constants are
// wave-uniform (scalar registers, not a copy per lane): a wave's rows are
// consecutive in cell order and the systems' cells are numbered one system
// after another, so a wave nearly always holds one system (two where one
// system's rows end).  f(sy) runs on the lanes whose rows belong to system sy;
// the lanes of a row share its system, so a row is never split.
template <class F>
__device__ __forceinline__ void by_system(bool valid, uint32_t sy_lane, F&& f) {
  unsigned long long todo = __ballot(valid);
  while (todo) {
    uint32_t sy = __builtin_amdgcn_readlane(sy_lane, int(__builtin_ctzll(todo)));
    // both sides of the test opaque to the compiler, the lane's copy made
    // last: inside the branch it would otherwise use the equal per-lane value
    // for sy (and load the system's constants per lane)
    asm volatile("" : "+s"(sy));
    uint32_t sy_mine = sy_lane;
    asm volatile("" : "+v"(sy_mine));
    const bool mine = valid && sy_mine == sy;
    todo &= ~__ballot(mine);
    if (mine) f(sy);
  }
}

// The end of a row: its force stored (NoKick), or applied to the row's particle
// by the step's kick (StepKick).
template <class Ph, class K>
__device__ __forceinline__ void finish_row(const Ph& ph, const typename Ph::St& st, uint32_t p, uint32_t sy, int w,
                                           const K& kick, const typename K::Pre& pre) {
  if constexpr (K::kOn) {
    typename K::Tv a[3];
    ph.force_as(st, a);
    // opaque copies: the kick's addresses and the cell terms of its records
    // are computed here, not hoisted above the row's loop where they would
    // hold registers throughout
    uint32_t pk = p;
    int wk = w;
    asm volatile("" : "+v"(pk), "+v"(wk));
    kick.apply(pre, pk, sy, wk, a);
  } else {
    ph.finish(st, p);
  }
}

// Verlet rows of an ensemble: the constants of row k are the table entry of its
// system (sid_sorted of the row's particle); kLanes lanes per row.  K: NoKick
// (forces stored) or StepKick<T> (the row's particle kicked and drifted).
template <class T, class Ph, int kLanes, bool kCubic, class K>
__global__ __launch_bounds__(256) void k_pairs_many_f32c(ClDev c, const Ph* __restrict__ tab,
                                                         const QGeom* __restrict__ qtab,
                                                         const uint32_t* __restrict__ sid_sorted,
                                                         const unsigned long long* __restrict__ voff,
                                                         const uint32_t* __restrict__ vnbr,
                                                         unsigned long long e_lim, const T* __restrict__ sref,
                                                         size_t sstride, K kick) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = tid / kLanes;
  const bool valid = k < c.n_dom;
  const uint32_t p = c.dom[valid ? k : 0];
  const int4 P = c.quant[p];
  const uint32_t ik = kick.index(p);
  const unsigned lane = threadIdx.x & (kLanes - 1);
  by_system(valid, sid_sorted[p], [&](uint32_t sy) {
  const Ph ph = tab[sy];
  const QGeom qg = qtab[sy];
  // the stored position, for close pairs
  const T pp[3] = {sref[p], sref[sstride + p], sref[2 * sstride + p]};
  typename Ph::St st;
  ph.start(st, p);
  typename K::Pre pre;
  // the kick's loads after the row's loop (an opaque index, so they are not
  // hoisted above it, where their registers would be held throughout)
  auto prefetch = [&] {
    uint32_t i = ik;
    asm volatile("" : "+v"(i));
    pre = kick.prefetch(i);
  };
  if (kLanes == 1) {
    // e_lim: entries the list holds (a list that overflowed its capacity was not
    // written; its rows are cut at the capacity and the caller rebuilds)
    const unsigned long long e0 = voff[k];
    const unsigned long long e1 = voff[k + 1] < e_lim ? voff[k + 1] : e_lim;
    Deferred df;
    row_groups(vnbr, e0, e1, p, [&](const uint32_t (&j)[4], auto&& next) {
      int4 Q[4];
      float a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        // slots outside the row (j = p) gather nothing: each gathered record
        // costs the wave a cache-line lookup
        Q[u] = P;
        a[u] = 0.f;
        if (j[u] != p) {
          Q[u] = c.quant[j[u]];
          a[u] = ph.aux(j[u]);
        }
      }
      next();
      const bool ns[4] = {j[0] != p, j[1] != p, j[2] != p, j[3] != p};
      quant_group_defer<T, Ph, kCubic>(ph, st, qg, a, Q, P, j, ns, pp, sref, sstride, df);
    });
    prefetch();
    flush_deferred<T, Ph>(ph, st, pp, sref, sstride, df);
  } else {
    const unsigned long long e1 = voff[k + 1] < e_lim ? voff[k + 1] : e_lim;
    // next step's entries in flight during the current step, unconditional
    // loads (as in k_pairs_f32c): a lane past the row end reads the row's last
    // entry (entry 0 if the first row is empty) and its pair is masked off
    const unsigned long long elast = e1 > 0 ? e1 - 1 : 0;
    uint32_t jn[kUnroll];
    bool okn[kUnroll];
    auto load_j = [&](unsigned long long e) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const unsigned long long ee = e + u * kLanes;
        okn[u] = ee < e1;
        jn[u] = vnbr[okn[u] ? ee : elast];
      }
    };
    load_j(voff[k] + lane);
    for (unsigned long long e = voff[k] + lane; e < e1; e += kUnroll * kLanes) {
      uint32_t j[kUnroll];
      bool ok[kUnroll];
      int4 Q[kUnroll];
      float a[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        j[u] = jn[u];
        ok[u] = okn[u] && jn[u] != p;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        Q[u] = c.quant[j[u]];
        a[u] = ph.aux(j[u]);
      }
      load_j(e + kUnroll * kLanes);
      static_assert(kUnroll == 4, "groups of four entries");
      quant_group_masked<T, Ph, kCubic>(ph, st, qg, a, Q, P, j, ok, pp, sref, sstride);
    }
    prefetch();
  }
  ph.reduce(st, kLanes);
  if (lane == 0) finish_row(ph, st, p, sy, P.w, kick, pre);
  });
}

template <class T, class Ph, int kLanes, class K>
__global__ __launch_bounds__(256) void k_pairs_many_f64(ClDev c, const Ph* __restrict__ tab,
                                                        const uint32_t* __restrict__ sid_sorted,
                                                        const T* __restrict__ pos, size_t stride,
                                                        const unsigned long long* __restrict__ voff,
                                                        const uint32_t* __restrict__ vnbr,
                                                        unsigned long long e_lim, K kick) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = tid / kLanes;
  const bool valid = k < c.n_dom;
  const uint32_t p = c.dom[valid ? k : 0];
  const double xp = pos[p], yp = pos[stride + p], zp = pos[2 * stride + p];
  const uint32_t ik = kick.index(p);
  const unsigned lane = threadIdx.x & (kLanes - 1);
  by_system(valid, sid_sorted[p], [&](uint32_t sy) {
  const Ph ph = tab[sy];
  typename Ph::St st;
  ph.start(st, p);
  auto eval = [&](uint32_t j, T qx, T qy, T qz) {
    const double rx = xp - double(qx);
    const double ry = yp - double(qy);
    const double rz = zp - double(qz);
    const double d2 = dist2_unfused<double>(rx, ry, rz);
    if (d2 < ph.rc2) ph.pair(st, 0.0, rx, ry, rz, d2);
  };
  if (kLanes == 1) {
    const unsigned long long e1 = voff[k + 1] < e_lim ? voff[k + 1] : e_lim;
    row_groups(vnbr, voff[k], e1, p, [&](const uint32_t (&j)[4], auto&& next) {
      T qx[4], qy[4], qz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        qx[u] = pos[j[u]];
        qy[u] = pos[stride + j[u]];
        qz[u] = pos[2 * stride + j[u]];
      }
      next();
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (j[u] != p) eval(j[u], qx[u], qy[u], qz[u]);
    });
  } else {
    const unsigned long long e1 = voff[k + 1] < e_lim ? voff[k + 1] : e_lim;
    for (unsigned long long e = voff[k] + lane; e < e1; e += kUnroll * kLanes) {
      uint32_t j[kUnroll];
      T qx[kUnroll], qy[kUnroll], qz[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const unsigned long long ee = e + u * kLanes;
        j[u] = ee < e1 ? vnbr[ee] : p;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        qx[u] = pos[j[u]];
        qy[u] = pos[stride + j[u]];
        qz[u] = pos[2 * stride + j[u]];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        if (j[u] != p) eval(j[u], qx[u], qy[u], qz[u]);
    }
  }
  uint32_t i = ik;  // opaque: the kick's loads stay after the loop
  asm volatile("" : "+v"(i));
  const typename K::Pre pre = kick.prefetch(i);
  ph.reduce(st, kLanes);
  if (lane == 0) finish_row(ph, st, p, sy, 0, kick, pre);
  });
}

// ---------------------------------------------------------------- EOS
template <class PT>
__global__ __launch_bounds__(256) void k_tait(const double* __restrict__ rho, size_t n, double inv_rho0,
                                              double B, double* __restrict__ prs, PT* __restrict__ pterm) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double rr = rho[i];
  double r = rr * inv_rho0;
  double r2 = r * r, r4 = r2 * r2;
  double P = B * (r4 * r2 * r - 1.0);
  if (prs) prs[i] = P;
  pterm[i] = PT(P / (rr * rr));
}

// ------------------------------------------------------------- reductions
constexpr int kRedBlock = 256;
constexpr unsigned kRedGrid = 1024;

__device__ __forceinline__ double red_op(int op, double a, double b) {
  return op == OFP_RED_SUM ? a + b : (op == OFP_RED_MAX ? fmax(a, b) : fmin(a, b));
}

template <class T>
__global__ __launch_bounds__(kRedBlock) void k_reduce_partial(const T* __restrict__ x, const T* __restrict__ comp,
                                                              size_t n, int op, double ident,
                                                              double* __restrict__ partial) {
  __shared__ double lds[kRedBlock / kWave];
  double acc = ident;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    double v = double(x[i]);
    if (comp) v += double(comp[i]);
    acc = red_op(op, acc, v);
  }
  for (int d = kWave / 2; d > 0; d >>= 1) acc = red_op(op, acc, __shfl_xor(acc, d, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int w = 1; w < kRedBlock / kWave; ++w) r = red_op(op, r, lds[w]);
    partial[blockIdx.x] = r;
  }
}

// Combines the nb partials and writes the result to *out (for ofp_reduce the
// slot after the partials, for ofp_reduce_noalloc the caller's device word).
__global__ __launch_bounds__(kRedBlock) void k_reduce_final(const double* partial, unsigned nb, int op, double ident,
                                                            double* __restrict__ out) {
  __shared__ double lds[kRedBlock / kWave];
  double acc = ident;
  for (unsigned i = threadIdx.x; i < nb; i += blockDim.x) acc = red_op(op, acc, partial[i]);
  for (int d = kWave / 2; d > 0; d >>= 1) acc = red_op(op, acc, __shfl_xor(acc, d, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int w = 1; w < kRedBlock / kWave; ++w) r = red_op(op, r, lds[w]);
    *out = r;
  }
}

// Partials of ofp_reduce_noalloc: one file-scope array per GPU, the most the
// partial kernel's grid can write (kRedGrid) plus one.  In host code the name
// stands for the host-side shadow, never for the device copy: kernels and
// copies get the device address that reduce_scratch() looks up.
constexpr size_t kRedScratch = size_t(kRedGrid) + 1;
static __device__ __align__(256) double scratch[kRedScratch];

// ------------------------------------------------------------- launchers
struct PairCtx {
  const ofp_cl* cl = nullptr;
  ClDev c;
  ofp_cl_view v;
  const unsigned long long* voff = nullptr;
  const uint32_t* vnbr = nullptr;
};

static int pair_ctx(const ofp_cl* cl, const ofp_vl* vl, int precision, double cutoff, PairCtx* x) {
  OFP_CHECK(cl, OFP_E_ARG, "null cell list");
  x->cl = cl;
  OFP_TRY(ofp_cl_dev_(cl, &x->c, &x->v));
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG, "bad precision %d", precision);
  if (precision == OFP_PREC_F32C)
    OFP_CHECK(x->c.quant != nullptr, OFP_E_STATE, "FP32C pair kernels need a cell list built with OFP_CL_QUANT");
  if (vl) {
    OFP_TRY(ofp_vl_dev_(vl, cl, 0, &x->voff, &x->vnbr));
  } else {
    for (int k = 0; k < 3; ++k)
      OFP_CHECK(cutoff <= x->v.spacing[k] * x->c.box_nn * (1 + 1e-6), OFP_E_ARG,
                "cut-off %g larger than the stencil (%u x %g)", cutoff, x->c.box_nn, x->v.spacing[k]);
  }
  return OFP_OK;
}

// Persistent cell-tile launches: 8 workgroups of 4 waves per CU.
static unsigned tile_grid() {
  static unsigned g = 0;
  if (!g) {
    int dev = 0, ncu = 256;
    if (ofp_device_info(&dev, &ncu, nullptr, 0) != OFP_OK || ncu <= 0) ncu = 256;
    g = unsigned(ncu) * 8u;
  }
  return g;
}

struct TileQueue {
  const TargetItem* items = nullptr;
  const uint32_t* n_items = nullptr;
  uint32_t* counter = nullptr;
};

// The cell list's item queue with its consumer head cleared (by a kernel, so a
// caller may capture the launch in a graph).
static int tile_queue(const PairCtx& x, TileQueue* q, hipStream_t s) {
  size_t cap = 0;
  OFP_TRY(ofp_cl_items_(x.cl, &q->items, &q->n_items, &q->counter, &cap, s));
  return fill_u32(q->counter, 0, 1, s);
}

// Verlet-row launches: kRowLanes lanes per domain row.
static unsigned pair_grid(const PairCtx& x) { return grid_for(size_t(x.c.n_dom) * kRowLanes, 256); }

// With a Verlet list: its rows; without: the cell tiles of the cell list.
template <class Ph>
static int launch_f64(const PairCtx& x, const void* pos, size_t stride, const Ph& ph, hipStream_t s) {
  if (x.c.n_dom == 0) return OFP_OK;
  OFP_CHECK(pos, OFP_E_ARG, "null sorted positions");
  OFP_CHECK(stride >= x.v.n_part, OFP_E_ARG, "stride < n_part");
  auto go = [&](auto tag) -> int {
    using T = decltype(tag);
    const T* P = static_cast<const T*>(pos);
    if (x.voff) {
      k_pairs_f64<T, Ph><<<pair_grid(x), 256, 0, s>>>(x.c, P, stride, x.voff, x.vnbr, ph);
    } else {
      TileQueue q;
      OFP_TRY(tile_queue(x, &q, s));
      k_tile_f64<T, Ph, false><<<tile_grid(), kTileBlock, 0, s>>>(x.c, P, stride, ph, nullptr, nullptr, q.items,
                                                                   q.n_items, q.counter);
    }
    OFP_LAUNCHED();
    return OFP_OK;
  };
  return x.v.dtype == OFP_F32 ? go(float(0)) : go(double(0));
}

static QGeom qgeom(const PairCtx& x) {
  QGeom q;
  q.ryd = x.v.spacing[1] / x.v.spacing[0];
  q.rzd = x.v.spacing[2] / x.v.spacing[0];
  q.ry = float(q.ryd);
  q.rz = float(q.rzd);
  return q;
}

static double q0_of(const PairCtx& x) { return x.v.spacing[0] / double(kQuantOne); }

static SplineQ spline_q(double h, double q0) {
  SplineQ sp;
  sp.qh = q0 / h;
  return sp;
}

// (q_x, q_y, q_z, aux) per sorted particle: the force rows gather one record
// per entry instead of two.
__global__ void k_quant_aux(const int4* __restrict__ quant, const float* __restrict__ aux, uint32_t n,
                            int4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 q = quant[i];
  out[i] = make_int4(q.x, q.y, q.z, __float_as_int(aux[i]));
}

template <class Ph>
static int launch_f32c(const PairCtx& x, const Ph& ph, hipStream_t s) {
  if (x.c.n_dom == 0) return OFP_OK;
  const unsigned grid = pair_grid(x);
  const QGeom qg = qgeom(x);
  const bool cubic = qg.ryd == 1.0 && qg.rzd == 1.0;
  if (x.voff) {
    if (cubic)
      k_pairs_f32c<Ph, true><<<grid, 256, 0, s>>>(x.c, qg, x.voff, x.vnbr, ph);
    else
      k_pairs_f32c<Ph, false><<<grid, 256, 0, s>>>(x.c, qg, x.voff, x.vnbr, ph);
  } else {
    TileQueue q;
    OFP_TRY(tile_queue(x, &q, s));
    const int4* rec = x.c.quant;
    if constexpr (Ph::kAuxInQuant) rec = ph.qaux;
    OFP_CHECK(rec, OFP_E_STATE, "cell-tile pass without its record stream");
    if (cubic)
      k_tile_f32c<float, Ph, true, false><<<tile_grid(), kTileBlock, 0, s>>>(
          x.c, qg, ph, nullptr, nullptr, nullptr, rec, q.items, q.n_items, q.counter);
    else
      k_tile_f32c<float, Ph, false, false><<<tile_grid(), kTileBlock, 0, s>>>(
          x.c, qg, ph, nullptr, nullptr, nullptr, rec, q.items, q.n_items, q.counter);
  }
  OFP_LAUNCHED();
  return OFP_OK;
}

// pos: sorted positions of the pass (pstride apart, FP64 if pos_f64) for the
// close-pair refinement, or null.
// The forms of one physics with and without energy share their layout.
static_assert(sizeof(LJ64T<true>) == sizeof(LJ64T<false>) && sizeof(LJ32T<true>) == sizeof(LJ32T<false>));
template <class To, class From>
static To same_layout(const From& f) {
  static_assert(sizeof(To) == sizeof(From) && std::is_trivially_copyable_v<To>);
  To t;
  std::memcpy(&t, &f, sizeof(To));
  return t;
}

// The band of ExactCut for a cut-off r_cut (position units) compared with the
// FP32 square rc2f (q_0 units); qm = the largest quantum in q_0 units (1 for
// cubic cells).  Relative half-width: the quantisation bound at r = r_c with a
// 25 % margin, plus 2^-22 for the FP32 roundings of the square and of rc2f.
static int exact_cut(const void* pos, size_t pstride, int pos_f64, double r_cut, double rc2_pos, double q0,
                     double qm, float rc2f, ExactCut* xc) {
  OFP_CHECK(pos, OFP_E_ARG, "FP32C pair passes need the sorted positions (cut-off ties are decided on them)");
  const double R = r_cut / q0;
  const double rel = 1.25 * (2.0 * std::sqrt(3.0) * qm / R + 3.0 * qm * qm / (R * R)) + std::ldexp(1.0, -22);
  OFP_CHECK(rel < std::ldexp(1.0, -18), OFP_E_LIMIT, "cells too large for the cut-off %g (quantum band %g)", r_cut,
            rel);
  xc->pos = pos;
  xc->pstride = pstride;
  xc->pos_f64 = pos_f64;
  xc->lo2 = std::nextafter(float(double(rc2f) * (1.0 - rel)), 0.0f);
  xc->hi2 = std::nextafter(float(double(rc2f) * (1.0 + rel)), INFINITY);
  xc->rc2 = rc2_pos;
  return OFP_OK;
}

static double quantum_ratio(double ryd, double rzd) { return std::max(1.0, std::max(ryd, rzd)); }

static int lj32_of(const ofp_lj_params& prm, double q0, double qm, const void* pos, size_t pstride, int pos_f64,
                   LJ32* out) {
  LJ32 ph;
  std::memset(&ph, 0, sizeof(ph));
  const double s = prm.sigma / q0;
  ph.rc2 = float((prm.r_cut / q0) * (prm.r_cut / q0));
  ph.sig2 = s * s;
  ph.fscale = 24.0 * prm.eps / q0;
  ph.uscale = 4.0 * prm.eps;
  OFP_TRY(exact_cut(pos, pstride, pos_f64, prm.r_cut, prm.r_cut * prm.r_cut, q0, qm, ph.rc2, &ph.xc));
  ph.inv_q0 = 1.0 / q0;
  ph.close2 = float(s * s);
  *out = ph;
  return OFP_OK;
}

// Device address of this GPU's `scratch` array, looked up once per device
// (with that device current) by hipGetSymbolAddress and kept.  Each GPU holds
// its own copy of the array at its own address.
constexpr int kMaxDevices = 64;
static int reduce_scratch(double** out) {
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  OFP_HIP(hipGetDevice(&dev));
  OFP_CHECK(dev >= 0 && dev < kMaxDevices, OFP_E_ARG, "device %d beyond the %d this library serves", dev,
            kMaxDevices);
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(OFP_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return OFP_OK;
}

}  // namespace ofp

using namespace ofp;

// Per-system constant tables of the ensemble LJ pass (device memory).
// spacing = 3 cell spacings per system; outputs are shared by all systems.
int ofp::lj_tables_many(int precision, uint32_t n_sys, const ofp_lj_params* prm, const double* spacing, void* fx,
                        void* fy, void* fz, void* energy, uint32_t* n_neigh, const void* pos_sorted,
                        size_t pstride, int pos_f64, DevBuf* tab, DevBuf* qtab, int* all_cubic) {
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG, "bad precision %d", precision);
  if (all_cubic) {
    *all_cubic = 1;
    for (uint32_t s = 0; s < n_sys; ++s)
      if (spacing[3 * s + 1] != spacing[3 * s] || spacing[3 * s + 2] != spacing[3 * s]) *all_cubic = 0;
  }
  if (precision == OFP_PREC_F64) {
    std::vector<LJ64> t(n_sys);
    for (uint32_t s = 0; s < n_sys; ++s)
      t[s] = LJ64{prm[s].r_cut * prm[s].r_cut, prm[s].sigma * prm[s].sigma, 24.0 * prm[s].eps, 4.0 * prm[s].eps,
                  (double*)fx, (double*)fy, (double*)fz, (double*)energy, n_neigh};
    OFP_TRY(tab->ensure(t.size() * sizeof(LJ64)));
    OFP_HIP(hipMemcpy(tab->p, t.data(), t.size() * sizeof(LJ64), hipMemcpyHostToDevice));
    return OFP_OK;
  }
  std::vector<LJ32> t(n_sys);
  std::vector<QGeom> q(n_sys);
  for (uint32_t s = 0; s < n_sys; ++s) {
    const double q0 = spacing[3 * s] / double(kQuantOne);
    const double qm = quantum_ratio(spacing[3 * s + 1] / spacing[3 * s], spacing[3 * s + 2] / spacing[3 * s]);
    OFP_TRY(lj32_of(prm[s], q0, qm, pos_sorted, pstride, pos_f64, &t[s]));
    t[s].fx = (float*)fx;
    t[s].fy = (float*)fy;
    t[s].fz = (float*)fz;
    t[s].en = (float*)energy;
    t[s].nn = n_neigh;
    q[s].ryd = spacing[3 * s + 1] / spacing[3 * s];
    q[s].rzd = spacing[3 * s + 2] / spacing[3 * s];
    q[s].ry = float(q[s].ryd);
    q[s].rz = float(q[s].rzd);
  }
  OFP_TRY(tab->ensure(t.size() * sizeof(LJ32)));
  OFP_TRY(qtab->ensure(q.size() * sizeof(QGeom)));
  OFP_HIP(hipMemcpy(tab->p, t.data(), t.size() * sizeof(LJ32), hipMemcpyHostToDevice));
  OFP_HIP(hipMemcpy(qtab->p, q.data(), q.size() * sizeof(QGeom), hipMemcpyHostToDevice));
  return OFP_OK;
}

// Lanes per row as a compile-time constant: f(std::integral_constant<int, L>).
template <class F>
static void with_lanes(int lanes, F&& f) {
  switch (lanes) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: break;
  }
}

// A kick applies to positions of its own dtype only.
template <class K, class T>
constexpr bool kick_fits() {
  if constexpr (K::kOn)
    return std::is_same<typename K::Tv, T>::value;
  else
    return true;
}

// The table memory holds LJ32 / LJ64 (kU = true); the forces-only forms read
// the same bytes (same layout).
template <bool kU, class K>
static void launch_lj_many(const PairCtx& x, int dtype, int precision, const void* pos_sorted, size_t stride,
                           const void* tab, const void* qtab, const uint32_t* sid, int lanes, unsigned long long el,
                           bool cubic, const K& kick, hipStream_t s) {
  const unsigned grid = grid_for(size_t(x.c.n_dom) * unsigned(lanes), 256);
  const auto* vo = x.voff;
  const auto* vn = x.vnbr;
  with_lanes(lanes, [&](auto L) {
    constexpr int kL = decltype(L)::value;
    if (precision == OFP_PREC_F32C) {
      using Ph = LJ32T<kU>;
      const Ph* t = static_cast<const Ph*>(tab);
      const QGeom* q = static_cast<const QGeom*>(qtab);
      auto go = [&](auto T0) {
        using T = decltype(T0);
        const T* sr = static_cast<const T*>(pos_sorted);
        if constexpr (kick_fits<K, T>()) {
          if (cubic)
            k_pairs_many_f32c<T, Ph, kL, true, K><<<grid, 256, 0, s>>>(x.c, t, q, sid, vo, vn, el, sr, stride, kick);
          else
            k_pairs_many_f32c<T, Ph, kL, false, K><<<grid, 256, 0, s>>>(x.c, t, q, sid, vo, vn, el, sr, stride,
                                                                        kick);
        }
      };
      if (dtype == OFP_F32)
        go(float(0));
      else
        go(double(0));
    } else {
      using Ph = LJ64T<kU>;
      const Ph* t = static_cast<const Ph*>(tab);
      if (dtype == OFP_F32) {
        if constexpr (kick_fits<K, float>())
          k_pairs_many_f64<float, Ph, kL, K><<<grid, 256, 0, s>>>(x.c, t, sid, static_cast<const float*>(pos_sorted),
                                                                  stride, vo, vn, el, kick);
      } else {
        if constexpr (kick_fits<K, double>())
          k_pairs_many_f64<double, Ph, kL, K><<<grid, 256, 0, s>>>(
              x.c, t, sid, static_cast<const double*>(pos_sorted), stride, vo, vn, el, kick);
      }
    }
  });
}

// Ensemble LJ over the cell tiles of the last build's cells (no Verlet list):
// one persistent launch over the (system, cell tile) item queue.
template <bool kU, class T>
static int launch_lj_tiles(const PairCtx& x, int precision, const void* pos_sorted, size_t stride, const void* tab,
                           const void* qtab, const void* sysgrid, bool cubic, hipStream_t s) {
  TileQueue q;
  OFP_TRY(tile_queue(x, &q, s));
  const SysGrid<T>* sys = static_cast<const SysGrid<T>*>(sysgrid);
  if (precision == OFP_PREC_F32C) {
    using Ph = LJ32T<kU>;
    const Ph* t = static_cast<const Ph*>(tab);
    const QGeom* qt = static_cast<const QGeom*>(qtab);
    if (cubic)
      k_tile_f32c<T, Ph, true, true><<<tile_grid(), kTileBlock, 0, s>>>(x.c, QGeom{}, Ph{}, t, qt, sys, x.c.quant,
                                                                         q.items, q.n_items, q.counter);
    else
      k_tile_f32c<T, Ph, false, true><<<tile_grid(), kTileBlock, 0, s>>>(x.c, QGeom{}, Ph{}, t, qt, sys, x.c.quant,
                                                                          q.items, q.n_items, q.counter);
  } else {
    using Ph = LJ64T<kU>;
    k_tile_f64<T, Ph, true><<<tile_grid(), kTileBlock, 0, s>>>(x.c, static_cast<const T*>(pos_sorted), stride, Ph{},
                                                               static_cast<const Ph*>(tab), sys, q.items, q.n_items,
                                                               q.counter);
  }
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp::pair_lj_many(const ofp_cl* cl, const ofp_vl* vl, int dtype, const void* pos_sorted, size_t stride,
                      int precision, const void* tab, const void* qtab, int lanes, size_t e_cap, int with_u,
                      int cubic, hipStream_t s, const void* sysgrid, const void* kick, const void* quant) {
  PairCtx x;
  OFP_TRY(pair_ctx(cl, vl, precision, 0.0, &x));
  if (quant) x.c.quant = static_cast<const int4*>(quant);
  const uint32_t* sid = nullptr;
  OFP_TRY(cl_sid_sorted(cl, &sid));
  if (x.c.n_dom == 0) return OFP_OK;
  if (!vl) {
    OFP_CHECK(sysgrid, OFP_E_ARG, "ensemble cell-tile pass needs the systems' grids");
    OFP_CHECK(!kick, OFP_E_ARG, "the cell-tile pass stores its forces (no fused kick)");
    if (dtype == OFP_F32)
      return with_u ? launch_lj_tiles<true, float>(x, precision, pos_sorted, stride, tab, qtab, sysgrid, cubic, s)
                    : launch_lj_tiles<false, float>(x, precision, pos_sorted, stride, tab, qtab, sysgrid, cubic, s);
    return with_u ? launch_lj_tiles<true, double>(x, precision, pos_sorted, stride, tab, qtab, sysgrid, cubic, s)
                  : launch_lj_tiles<false, double>(x, precision, pos_sorted, stride, tab, qtab, sysgrid, cubic, s);
  }
  OFP_CHECK(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32, OFP_E_ARG,
            "lanes per row must be 1, 2, 4, 8, 16 or 32 (got %d)", lanes);
  if (kick) {
    OFP_CHECK(!with_u, OFP_E_ARG, "a fused kick writes no energies");
    if (dtype == OFP_F32)
      launch_lj_many<false>(x, dtype, precision, pos_sorted, stride, tab, qtab, sid, lanes, e_cap, cubic != 0,
                            *static_cast<const StepKick<float>*>(kick), s);
    else
      launch_lj_many<false>(x, dtype, precision, pos_sorted, stride, tab, qtab, sid, lanes, e_cap, cubic != 0,
                            *static_cast<const StepKick<double>*>(kick), s);
  } else if (with_u) {
    launch_lj_many<true>(x, dtype, precision, pos_sorted, stride, tab, qtab, sid, lanes, e_cap, cubic != 0, NoKick{},
                         s);
  } else {
    launch_lj_many<false>(x, dtype, precision, pos_sorted, stride, tab, qtab, sid, lanes, e_cap, cubic != 0, NoKick{},
                          s);
  }
  OFP_LAUNCHED();
  return OFP_OK;
}

// LJ over a symmetric list (each interaction once): one wave per row, lanes over
// the row's entries; the pair force goes to the row's particle (+f, summed over
// the wave) and to the neighbour (-f) by FP64 atomic adds, and both neighbour
// counts are incremented.  fx/fy/fz/nn hold every particle, ghosts included
// (sorted index), and start at zero.
template <class T>
__global__ __launch_bounds__(256) void k_lj_sym(const T* __restrict__ pos, size_t stride,
                                                const unsigned long long* __restrict__ off,
                                                const uint32_t* __restrict__ nbr, uint64_t n_rows,
                                                const uint32_t* __restrict__ dom, double rc2, double sig2,
                                                double eps24, double* __restrict__ fx, double* __restrict__ fy,
                                                double* __restrict__ fz, uint32_t* __restrict__ nn) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t r = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (r >= n_rows) return;  // wave-uniform
  const uint32_t i = dom[r];
  const double xi = double(pos[i]), yi = double(pos[stride + i]), zi = double(pos[2 * stride + i]);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  uint32_t cnt = 0;
  const uint64_t e1 = off[r + 1];
  for (uint64_t e = off[r] + uint64_t(lane); e < e1; e += kWave) {
    const uint32_t j = nbr[e];
    const double rx = xi - double(pos[j]), ry = yi - double(pos[stride + j]), rz = zi - double(pos[2 * stride + j]);
    const double d2 = rx * rx + ry * ry + rz * rz;
    if (d2 < rc2) {
      ++cnt;
      if (nn) atomicAdd(nn + j, 1u);
      if (d2 > 0.0) {
        const double inv = 1.0 / d2;
        const double s2 = sig2 * inv;
        const double s6 = s2 * s2 * s2;
        const double f = eps24 * (2.0 * s6 * s6 - s6) * inv;
        sx += f * rx;
        sy += f * ry;
        sz += f * rz;
        unsafeAtomicAdd(fx + j, -f * rx);
        unsafeAtomicAdd(fy + j, -f * ry);
        unsafeAtomicAdd(fz + j, -f * rz);
      }
    }
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    sx += __shfl_xor(sx, d);
    sy += __shfl_xor(sy, d);
    sz += __shfl_xor(sz, d);
    cnt += __shfl_xor(cnt, d);
  }
  if (lane == 0) {
    unsafeAtomicAdd(fx + i, sx);
    unsafeAtomicAdd(fy + i, sy);
    unsafeAtomicAdd(fz + i, sz);
    if (nn) atomicAdd(nn + i, cnt);
  }
}

__global__ void k_zero_f64(double* __restrict__ p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

extern "C" {

int ofp_pair_lj_sym(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, const ofp_lj_params* prm,
                    double* fx, double* fy, double* fz, uint32_t* n_neigh, void* stream) {
  OFP_CHECK(cl && vl && pos && prm && fx && fy && fz, OFP_E_ARG, "ofp_pair_lj_sym: null argument");
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(ofp_cl_dev_(cl, &c, &v));
  const unsigned long long* off = nullptr;
  const uint32_t* nbr = nullptr;
  OFP_TRY(ofp_vl_dev_(vl, cl, 1, &off, &nbr));
  OFP_CHECK(stride >= v.n_part, OFP_E_ARG, "stride < particles");
  hipStream_t s = as_stream(stream);
  const size_t n = v.n_part;
  if (n == 0) return OFP_OK;
  double* outs[3] = {fx, fy, fz};
  for (double* o : outs) {
    k_zero_f64<<<grid_for(n, 256), 256, 0, s>>>(o, n);
    OFP_LAUNCHED();
  }
  if (n_neigh) OFP_TRY(fill_u32(n_neigh, 0u, n, s));
  const uint64_t rows = v.ghost_marker;
  if (rows == 0) return OFP_OK;
  const unsigned grid = grid_for(rows * kWave, 256);
  const double rc2 = prm->r_cut * prm->r_cut, sig2 = prm->sigma * prm->sigma, eps24 = 24.0 * prm->eps;
  if (v.dtype == OFP_F32)
    k_lj_sym<float><<<grid, 256, 0, s>>>(static_cast<const float*>(pos), stride, off, nbr, rows, v.dom_sort_ids,
                                         rc2, sig2, eps24, fx, fy, fz, n_neigh);
  else
    k_lj_sym<double><<<grid, 256, 0, s>>>(static_cast<const double*>(pos), stride, off, nbr, rows, v.dom_sort_ids,
                                          rc2, sig2, eps24, fx, fy, fz, n_neigh);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_pair_lj(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, int precision,
                const ofp_lj_params* prm, void* fx, void* fy, void* fz, void* energy, uint32_t* n_neigh,
                void* stream) {
  OFP_CHECK(prm, OFP_E_ARG, "ofp_pair_lj: null parameters");
  OFP_CHECK((fx && fy && fz) || (!fx && !fy && !fz), OFP_E_ARG, "force planes: all or none");
  PairCtx x;
  OFP_TRY(pair_ctx(cl, vl, precision, prm->r_cut, &x));
  hipStream_t s = as_stream(stream);
  const bool with_u = energy || n_neigh;  // otherwise the forces-only kernels
  if (precision == OFP_PREC_F64) {
    LJ64 ph{prm->r_cut * prm->r_cut, prm->sigma * prm->sigma, 24.0 * prm->eps, 4.0 * prm->eps,
            (double*)fx, (double*)fy, (double*)fz, (double*)energy, n_neigh};
    return with_u ? launch_f64(x, pos, stride, ph, s)
                  : launch_f64(x, pos, stride, same_layout<LJ64T<false>>(ph), s);
  }
  const double q0 = q0_of(x);
  const QGeom qg = qgeom(x);
  LJ32 ph;
  OFP_TRY(lj32_of(*prm, q0, quantum_ratio(qg.ryd, qg.rzd), pos, stride, x.v.dtype == OFP_F64 ? 1 : 0, &ph));
  ph.fx = (float*)fx;
  ph.fy = (float*)fy;
  ph.fz = (float*)fz;
  ph.en = (float*)energy;
  ph.nn = n_neigh;
  return with_u ? launch_f32c(x, ph, s) : launch_f32c(x, same_layout<LJ32T<false>>(ph), s);
}

int ofp_sph_density(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, int precision,
                    const ofp_sph_params* prm, void* rho, void* stream) {
  OFP_CHECK(prm && rho, OFP_E_ARG, "ofp_sph_density: null argument");
  PairCtx x;
  const double h = prm->h;
  OFP_TRY(pair_ctx(cl, vl, precision, 2 * h, &x));
  hipStream_t s = as_stream(stream);
  const double sk_m = prm->mass / (M_PI * h * h * h);
  if (precision == OFP_PREC_F64) {
    Dens64 ph{4 * h * h, 1.0 / h, sk_m, (double*)rho};
    return launch_f64(x, pos, stride, ph, s);
  }
  const double q0 = q0_of(x);
  const QGeom qg = qgeom(x);
  Dens32 ph;
  ph.rc2 = float((2 * h / q0) * (2 * h / q0));
  OFP_TRY(exact_cut(pos, stride, x.v.dtype == OFP_F64 ? 1 : 0, 2 * h, 4 * h * h, q0, quantum_ratio(qg.ryd, qg.rzd),
                    ph.rc2, &ph.xc));
  ph.sp = spline_q(h, q0);
  ph.sk_m = sk_m;
  ph.rho = (double*)rho;
  return launch_f32c(x, ph, s);
}

int ofp_sph_eos(int precision, const void* rho, size_t n, const ofp_sph_params* prm, void* prs, void* pterm,
                void* stream) {
  OFP_CHECK(prm && (n == 0 || (rho && pterm)), OFP_E_ARG, "ofp_sph_eos: null argument");
  if (n == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  const double B = prm->c0 * prm->c0 * prm->rho0 / 7.0;
  if (precision == OFP_PREC_F64)
    k_tait<double><<<grid_for(n, 256), 256, 0, s>>>((const double*)rho, n, 1.0 / prm->rho0, B, (double*)prs,
                                                    (double*)pterm);
  else if (precision == OFP_PREC_F32C)
    k_tait<float><<<grid_for(n, 256), 256, 0, s>>>((const double*)rho, n, 1.0 / prm->rho0, B, (double*)prs,
                                                   (float*)pterm);
  else
    return set_error(OFP_E_ARG, "bad precision %d", precision);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_sph_force(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, int precision,
                  const ofp_sph_params* prm, const void* pterm, void* fx, void* fy, void* fz, void* stream) {
  OFP_CHECK(prm && pterm && fx && fy && fz, OFP_E_ARG, "ofp_sph_force: null argument");
  PairCtx x;
  const double h = prm->h;
  OFP_TRY(pair_ctx(cl, vl, precision, 2 * h, &x));
  hipStream_t s = as_stream(stream);
  const double sd_m = prm->mass / (M_PI * h * h * h * h);
  if (precision == OFP_PREC_F64) {
    Force64 ph{4 * h * h, 1.0 / h, sd_m, (const double*)pterm, (double*)fx, (double*)fy, (double*)fz};
    return launch_f64(x, pos, stride, ph, s);
  }
  const double q0 = q0_of(x);
  const QGeom qg = qgeom(x);
  Force32 ph;
  ph.rc2 = float((2 * h / q0) * (2 * h / q0));
  OFP_TRY(exact_cut(pos, stride, x.v.dtype == OFP_F64 ? 1 : 0, 2 * h, 4 * h * h, q0, quantum_ratio(qg.ryd, qg.rzd),
                    ph.rc2, &ph.xc));
  ph.sp = spline_q(h, q0);
  ph.fscale = -sd_m;
  ph.pterm = (const float*)pterm;
  ph.qaux = nullptr;
  ph.fx = (float*)fx;
  ph.fy = (float*)fy;
  ph.fz = (float*)fz;
  if (x.c.n_dom > 0) {
    void* qa = nullptr;
    if (x.voff)
      OFP_TRY(ofp_vl_aux_(vl, (x.v.n_part + 1) * sizeof(int4), &qa));
    else
      OFP_TRY(ofp_cl_aux_(cl, (x.v.n_part + 1) * sizeof(int4), &qa));
    k_quant_aux<<<grid_for(x.v.n_part, 256), 256, 0, s>>>(x.c.quant, (const float*)pterm, uint32_t(x.v.n_part),
                                                          static_cast<int4*>(qa));
    OFP_LAUNCHED();
    ph.qaux = static_cast<const int4*>(qa);
  }
  return launch_f32c(x, ph, s);
}

int ofp_reduce(int dtype, const void* x, const void* comp, size_t n, int op, double* result, void* stream) {
  OFP_CHECK(result, OFP_E_ARG, "ofp_reduce: null result");
  OFP_CHECK(n == 0 || x, OFP_E_ARG, "ofp_reduce: null input");
  OFP_CHECK(op == OFP_RED_SUM || op == OFP_RED_MAX || op == OFP_RED_MIN, OFP_E_ARG, "bad op %d", op);
  const double ident = op == OFP_RED_SUM ? 0.0 : (op == OFP_RED_MAX ? -INFINITY : INFINITY);
  if (n == 0) {
    *result = ident;
    return OFP_OK;
  }
  hipStream_t s = as_stream(stream);
  unsigned grid = grid_for(n, kRedBlock);
  if (grid > kRedGrid) grid = kRedGrid;
  double* partial = nullptr;
  OFP_HIP(hipMallocAsync(reinterpret_cast<void**>(&partial), (grid + 1) * sizeof(double), s));
  if (dtype == OFP_F32)
    k_reduce_partial<float><<<grid, kRedBlock, 0, s>>>((const float*)x, (const float*)comp, n, op, ident, partial);
  else if (dtype == OFP_F64)
    k_reduce_partial<double><<<grid, kRedBlock, 0, s>>>((const double*)x, (const double*)comp, n, op, ident,
                                                        partial);
  else {
    (void)hipFreeAsync(partial, s);
    return set_error(OFP_E_ARG, "bad dtype %d", dtype);
  }
  OFP_LAUNCHED();
  k_reduce_final<<<1, kRedBlock, 0, s>>>(partial, grid, op, ident, partial + grid);
  OFP_LAUNCHED();
  double r = 0;
  OFP_HIP(hipMemcpyAsync(&r, partial + grid, sizeof(double), hipMemcpyDeviceToHost, s));
  OFP_HIP(hipFreeAsync(partial, s));
  OFP_HIP(hipStreamSynchronize(s));
  *result = r;
  return OFP_OK;
}

int ofp_reduce_noalloc(int dtype, const void* x, const void* comp, size_t n, int op, double* result_dev,
                       void* stream) {
  OFP_CHECK(result_dev, OFP_E_ARG, "ofp_reduce_noalloc: null result");
  OFP_CHECK(n == 0 || x, OFP_E_ARG, "ofp_reduce_noalloc: null input");
  OFP_CHECK(op == OFP_RED_SUM || op == OFP_RED_MAX || op == OFP_RED_MIN, OFP_E_ARG, "bad op %d", op);
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  OFP_CHECK(n <= OFP_REDUCE_NOALLOC_MAX_N, OFP_E_ARG,
            "ofp_reduce_noalloc: %zu elements exceed the maximum of %llu", n,
            static_cast<unsigned long long>(OFP_REDUCE_NOALLOC_MAX_N));
  double* part = nullptr;
  OFP_TRY(reduce_scratch(&part));
  hipStream_t s = as_stream(stream);
  const double ident = op == OFP_RED_SUM ? 0.0 : (op == OFP_RED_MAX ? -INFINITY : INFINITY);
  // the same grid as ofp_reduce (so the same partials and the same result)
  unsigned grid = n ? grid_for(n, kRedBlock) : 0;
  if (grid > kRedGrid) grid = kRedGrid;
  OFP_CHECK(size_t(grid) + 1 <= kRedScratch, OFP_E_ARG, "ofp_reduce_noalloc: %u partials exceed the scratch", grid);
  if (grid > 0) {
    if (dtype == OFP_F32)
      k_reduce_partial<float><<<grid, kRedBlock, 0, s>>>((const float*)x, (const float*)comp, n, op, ident, part);
    else
      k_reduce_partial<double><<<grid, kRedBlock, 0, s>>>((const double*)x, (const double*)comp, n, op, ident,
                                                          part);
    OFP_LAUNCHED();
  }
  // n == 0: no partials, the final kernel writes the identity
  k_reduce_final<<<1, kRedBlock, 0, s>>>(part, grid, op, ident, result_dev);
  OFP_LAUNCHED();
  return OFP_OK;
}

}  // extern "C"
