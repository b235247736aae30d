// ofp_pairs.hip -- Verlet-list build and the pair-interaction loops (LJ, SPH
// density, SPH pressure force) in plain FP64 and in FP32 on quantised
// cell-relative coordinates with two-sum (compensated) accumulation.
//
// Thread mapping: one thread per domain particle, taken in domain-sort-id order
// (GET_PARTICLE_SORT, Vector/vector_dist_kernel.hpp:16-17), so the lanes of a wave
// share their neighbour cells and the neighbour loads are served by L1/L2.  The
// 27-cell walk is done as 9 contiguous x-runs: cells (x-1..x+1, y+dy, z+dz) are
// consecutive in the x-fastest linearisation, hence one sorted-index range each.
#include "ofp_internal.hpp"

#include <cmath>
#include <cstring>

struct ofp_vl {
  ofp::DevBuf cnt, off, nbr, scantmp;
  uint64_t n_rows = 0, n_entries = 0;
  bool built = false;
  const ofp_cl* cl = nullptr;
};

namespace ofp {

// Cell-list fields the walkers need (by value).
struct ClDev {
  const unsigned long long* cs;  // packed cell starts
  const uint32_t* dom;           // domain sort ids
  const int4* quant;
  uint32_t divc[3];
  uint32_t n_dom;
};

static int cl_dev(const ofp_cl* cl, ClDev* d, ofp_cl_view* v) {
  OFP_TRY(ofp_cl_get_view(cl, v));
  d->cs = reinterpret_cast<const unsigned long long*>(v->cell_start);
  d->dom = v->dom_sort_ids;
  d->quant = reinterpret_cast<const int4*>(v->quant);
  for (int k = 0; k < 3; ++k) d->divc[k] = v->divc[k];
  d->n_dom = uint32_t(v->ghost_marker);
  return OFP_OK;
}

// Sorted-index range of the x-run (x0..x1, yy, zz).
__device__ __forceinline__ void run_range(const ClDev& c, uint32_t x0, uint32_t x1, uint32_t yy,
                                          uint32_t zz, uint32_t* b, uint32_t* e) {
  uint32_t c0 = x0 + c.divc[0] * (yy + c.divc[1] * zz);
  uint32_t c1 = x1 + c.divc[0] * (yy + c.divc[1] * zz);
  *b = uint32_t(c.cs[c0]);
  *e = uint32_t(c.cs[c1 + 1]);
}

// Calls f(j, oy, oz) for every sorted index j of the 27 cells around (cx, cy, cz)
// in ascending sorted index (= reference traversal order), with the y/z cell
// offsets of the run it belongs to.
template <class F>
__device__ __forceinline__ void walk27(const ClDev& c, uint32_t cx, uint32_t cy, uint32_t cz,
                                       F&& f) {
  const uint32_t x0 = cx > 0 ? cx - 1 : 0;
  const uint32_t x1 = cx + 1 < c.divc[0] ? cx + 1 : c.divc[0] - 1;
  for (int dz = -1; dz <= 1; ++dz) {
    int zz = int(cz) + dz;
    if (zz < 0 || zz >= int(c.divc[2])) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int yy = int(cy) + dy;
      if (yy < 0 || yy >= int(c.divc[1])) continue;
      uint32_t b, e;
      run_range(c, x0, x1, uint32_t(yy), uint32_t(zz), &b, &e);
      for (uint32_t j = b; j < e; ++j) f(j, dy, dz);
    }
  }
}

// ----------------------------------------------------------- Verlet build
template <class T, bool kFill>
__global__ __launch_bounds__(256) void k_vl_pass(ClDev c, CellGeom<T> g, const T* __restrict__ pos,
                                                 size_t stride, T r2, uint32_t* __restrict__ cnt,
                                                 const unsigned long long* __restrict__ off,
                                                 uint32_t* __restrict__ nbr) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c.n_dom) return;
  uint32_t p = c.dom[k];
  const T xp = pos[p], yp = pos[stride + p], zp = pos[2 * stride + p];
  uint32_t cx, cy, cz;
  (void)cell_of(g, xp, yp, zp, &cx, &cy, &cz);
  uint32_t n = 0;
  unsigned long long base = kFill ? off[k] : 0ull;
  walk27(c, cx, cy, cz, [&](uint32_t j, int, int) {
    T d2 = dist2_unfused<T>(pos[j] - xp, pos[stride + j] - yp, pos[2 * stride + j] - zp);
    if (d2 < r2) {
      if (kFill) nbr[base + n] = j;
      ++n;
    }
  });
  if (!kFill) cnt[k] = n;
}

// ----------------------------------------------------------- accumulators
// Two-sum (Knuth) compensated accumulator: s + c carries the sum with the
// rounding error of every addition folded into c.
struct Comp32 {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    float t = s + x;
    float z = t - s;
    c += (s - (t - z)) + (x - z);
    s = t;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// ------------------------------------------------------------ physics FP64
struct LJ64 {
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
  __device__ __forceinline__ void pair(St& s, uint32_t, double rx, double ry, double rz,
                                       double d2) const {
    ++s.n;
    if (d2 == 0.0) return;
    double inv = 1.0 / d2;
    double s2 = sig2 * inv;
    double s6 = s2 * s2 * s2;
    double s12 = s6 * s6;
    double f = eps24 * (2.0 * s12 - s6) * inv;
    s.fx += f * rx;
    s.fy += f * ry;
    s.fz += f * rz;
    s.u += eps4 * (s12 - s6);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    if (fx) {
      fx[p] = s.fx;
      fy[p] = s.fy;
      fz[p] = s.fz;
    }
    if (en) en[p] = s.u;
    if (nn) nn[p] = s.n;
  }
};

// Monaghan cubic spline; sk = 1/(pi h^3), sd = 1/(pi h^4)
template <class R>
__device__ __forceinline__ R w3(R q, R sk) {
  if (q < R(1)) return sk * (R(1) - R(1.5) * q * q + R(0.75) * q * q * q);
  R t = R(2) - q;
  return q < R(2) ? sk * R(0.25) * t * t * t : R(0);
}
template <class R>
__device__ __forceinline__ R dw3(R q, R sd) {
  if (q < R(1)) return sd * (R(-3) * q + R(2.25) * q * q);
  R t = R(2) - q;
  return q < R(2) ? sd * (R(-0.75) * t * t) : R(0);
}

struct Dens64 {
  double rc2, inv_h, sk, mass;
  double* rho;
  struct St {
    double acc = 0;
  };
  __device__ __forceinline__ void pair(St& s, uint32_t, double, double, double, double d2) const {
    s.acc += mass * w3(sqrt(d2) * inv_h, sk);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    rho[p] = s.acc + mass * sk;  // self term m W(0)
  }
};

struct Force64 {
  double rc2, inv_h, sd, mass;
  const double* rho;
  const double* prs;
  double* fx;
  double* fy;
  double* fz;
  struct St {
    double fx = 0, fy = 0, fz = 0, pi = 0;
  };
  __device__ __forceinline__ void pair(St& s, uint32_t q, double rx, double ry, double rz,
                                       double d2) const {
    if (d2 == 0.0) return;
    double r = sqrt(d2);
    double rq = rho[q];
    double coef = -mass * (s.pi + prs[q] / (rq * rq)) * dw3(r * inv_h, sd) / r;
    s.fx += coef * rx;
    s.fy += coef * ry;
    s.fz += coef * rz;
  }
  __device__ __forceinline__ void start(St& s, uint32_t p) const {
    double rp = rho[p];
    s.pi = prs[p] / (rp * rp);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    fx[p] = s.fx;
    fy[p] = s.fy;
    fz[p] = s.fz;
  }
};

// ------------------------------------------------------------ physics FP32C
__device__ __forceinline__ float rcp_refined(float x) {
  float y = __builtin_amdgcn_rcpf(x);
  float e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(y, e, y);
}

struct LJ32 {
  float rc2, sig2, eps24, eps4;
  float* fx;
  float* fy;
  float* fz;
  float* en;
  uint32_t* nn;
  struct St {
    Comp32 fx, fy, fz, u;
    uint32_t n = 0;
  };
  __device__ __forceinline__ void pair(St& s, uint32_t, float rx, float ry, float rz,
                                       float d2) const {
    ++s.n;
    if (d2 == 0.f) return;
    float inv = rcp_refined(d2);
    float s2 = sig2 * inv;
    float s6 = s2 * s2 * s2;
    float s12 = s6 * s6;
    float f = eps24 * (2.0f * s12 - s6) * inv;
    s.fx.add(f * rx);
    s.fy.add(f * ry);
    s.fz.add(f * rz);
    s.u.add(eps4 * (s12 - s6));
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    if (fx) {
      fx[p] = s.fx.value();
      fy[p] = s.fy.value();
      fz[p] = s.fz.value();
    }
    if (en) en[p] = s.u.value();
    if (nn) nn[p] = s.n;
  }
};

struct Dens32 {
  float rc2, inv_h, sk, mass;
  float* rho;
  struct St {
    Comp32 acc;
  };
  __device__ __forceinline__ void pair(St& s, uint32_t, float, float, float, float d2) const {
    s.acc.add(mass * w3(__builtin_sqrtf(d2) * inv_h, sk));
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    Comp32 a = s.acc;
    a.add(mass * sk);
    rho[p] = a.value();
  }
};

struct Force32 {
  float rc2, inv_h, sd, mass;
  const float* rho;
  const float* prs;
  float* fx;
  float* fy;
  float* fz;
  struct St {
    Comp32 fx, fy, fz;
    float pi = 0;
  };
  __device__ __forceinline__ void start(St& s, uint32_t p) const {
    float rp = rho[p];
    s.pi = prs[p] / (rp * rp);
  }
  __device__ __forceinline__ void pair(St& s, uint32_t q, float rx, float ry, float rz,
                                       float d2) const {
    if (d2 == 0.f) return;
    float r = __builtin_sqrtf(d2);
    float rq = rho[q];
    float coef = -mass * (s.pi + prs[q] / (rq * rq)) * dw3(r * inv_h, sd) / r;
    s.fx.add(coef * rx);
    s.fy.add(coef * ry);
    s.fz.add(coef * rz);
  }
  __device__ __forceinline__ void finish(const St& s, uint32_t p) const {
    fx[p] = s.fx.value();
    fy[p] = s.fy.value();
    fz[p] = s.fz.value();
  }
};

template <class Ph>
struct has_start {
  template <class U>
  static auto test(int) -> decltype(&U::start, char());
  template <class U>
  static int test(...);
  static constexpr bool value = sizeof(test<Ph>(0)) == 1;
};

template <class Ph>
__device__ __forceinline__ void ph_start(const Ph& ph, typename Ph::St& s, uint32_t p) {
  if constexpr (has_start<Ph>::value) ph.start(s, p);
}

// ------------------------------------------------------------ FP64 kernel
template <class T, class Ph, bool kVerlet>
__global__ __launch_bounds__(256) void k_pairs_f64(ClDev c, CellGeom<T> g, const T* __restrict__ pos,
                                                   size_t stride,
                                                   const unsigned long long* __restrict__ voff,
                                                   const uint32_t* __restrict__ vnbr, Ph ph) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c.n_dom) return;
  const uint32_t p = c.dom[k];
  const T tx = pos[p], ty = pos[stride + p], tz = pos[2 * stride + p];
  const double xp = tx, yp = ty, zp = tz;
  typename Ph::St st;
  ph_start(ph, st, p);
  auto body = [&](uint32_t j) {
    if (j == p) return;
    double rx = xp - double(pos[j]);
    double ry = yp - double(pos[stride + j]);
    double rz = zp - double(pos[2 * stride + j]);
    // unfused, as the oracle, so the neighbour decision is bit-identical
    double d2 = dist2_unfused<double>(rx, ry, rz);
    if (d2 < ph.rc2) ph.pair(st, j, rx, ry, rz, d2);
  };
  if constexpr (kVerlet) {
    for (unsigned long long e = voff[k]; e < voff[k + 1]; ++e) body(vnbr[e]);
  } else {
    uint32_t cx, cy, cz;
    (void)cell_of(g, tx, ty, tz, &cx, &cy, &cz);
    walk27(c, cx, cy, cz, [&](uint32_t j, int, int) { body(j); });
  }
  ph.finish(st, p);
}

// ------------------------------------------------------------ FP32C kernel
struct QScale {
  float q[3];  // spacing / 2^29 per dim
};

__device__ __forceinline__ void unpack_cell(int32_t w, int* cx, int* cy, int* cz) {
  *cx = w & 2047;
  *cy = (w >> 11) & 2047;
  *cz = (w >> 22) & 1023;
}

template <class Ph, bool kVerlet>
__global__ __launch_bounds__(256) void k_pairs_f32c(ClDev c, QScale qs,
                                                    const unsigned long long* __restrict__ voff,
                                                    const uint32_t* __restrict__ vnbr, Ph ph) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c.n_dom) return;
  const uint32_t p = c.dom[k];
  const int4 P = c.quant[p];
  int pcx, pcy, pcz;
  unpack_cell(P.w, &pcx, &pcy, &pcz);
  typename Ph::St st;
  ph_start(ph, st, p);
  // x_p - x_q from the exact integer difference u_q - u_p + (c_q - c_p) 2^29
  auto body = [&](uint32_t j, int ox, int oy, int oz, const int4& Q) {
    if (j == p) return;
    int ix = (Q.x - P.x) + ox * kQuantOne;
    int iy = (Q.y - P.y) + oy * kQuantOne;
    int iz = (Q.z - P.z) + oz * kQuantOne;
    float rx = -float(ix) * qs.q[0];
    float ry = -float(iy) * qs.q[1];
    float rz = -float(iz) * qs.q[2];
    float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz));
    if (d2 < ph.rc2) ph.pair(st, j, rx, ry, rz, d2);
  };
  if constexpr (kVerlet) {
    for (unsigned long long e = voff[k]; e < voff[k + 1]; ++e) {
      uint32_t j = vnbr[e];
      int4 Q = c.quant[j];
      int qx, qy, qz;
      unpack_cell(Q.w, &qx, &qy, &qz);
      body(j, qx - pcx, qy - pcy, qz - pcz, Q);
    }
  } else {
    walk27(c, uint32_t(pcx), uint32_t(pcy), uint32_t(pcz), [&](uint32_t j, int oy, int oz) {
      int4 Q = c.quant[j];
      body(j, (Q.w & 2047) - pcx, oy, oz, Q);
    });
  }
  ph.finish(st, p);
}

// ---------------------------------------------------------------- EOS
template <class R>
__global__ __launch_bounds__(256) void k_tait(const R* __restrict__ rho, size_t n, R inv_rho0, R B,
                                              R* __restrict__ prs) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R r = rho[i] * inv_rho0;
  R r2 = r * r, r4 = r2 * r2;
  prs[i] = B * (r4 * r2 * r - R(1));
}

// ------------------------------------------------------------- reductions
constexpr int kRedBlock = 256;
constexpr unsigned kRedGrid = 1024;

__device__ __forceinline__ double red_op(int op, double a, double b) {
  return op == OFP_RED_SUM ? a + b : (op == OFP_RED_MAX ? fmax(a, b) : fmin(a, b));
}

template <class T>
__global__ __launch_bounds__(kRedBlock) void k_reduce_partial(const T* __restrict__ x,
                                                              const T* __restrict__ comp, size_t n,
                                                              int op, double ident,
                                                              double* __restrict__ partial) {
  __shared__ double lds[kRedBlock / kWave];
  double acc = ident;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
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

__global__ __launch_bounds__(kRedBlock) void k_reduce_final(double* partial, unsigned nb, int op,
                                                            double ident) {
  __shared__ double lds[kRedBlock / kWave];
  double acc = ident;
  for (unsigned i = threadIdx.x; i < nb; i += blockDim.x) acc = red_op(op, acc, partial[i]);
  for (int d = kWave / 2; d > 0; d >>= 1) acc = red_op(op, acc, __shfl_xor(acc, d, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (int w = 1; w < kRedBlock / kWave; ++w) r = red_op(op, r, lds[w]);
    partial[nb] = r;
  }
}

// ------------------------------------------------------------- launchers
struct PairCtx {
  ClDev c;
  ofp_cl_view v;
  const unsigned long long* voff = nullptr;
  const uint32_t* vnbr = nullptr;
};

static int pair_ctx(const ofp_cl* cl, const ofp_vl* vl, int precision, PairCtx* x) {
  OFP_CHECK(cl, OFP_E_ARG, "null cell list");
  OFP_TRY(cl_dev(cl, &x->c, &x->v));
  OFP_CHECK(precision == OFP_PREC_F64 || precision == OFP_PREC_F32C, OFP_E_ARG,
            "bad precision %d", precision);
  if (precision == OFP_PREC_F32C)
    OFP_CHECK(x->c.quant != nullptr, OFP_E_STATE,
              "FP32C pair kernels need a cell list built with OFP_CL_QUANT");
  if (vl) {
    OFP_CHECK(vl->built && vl->cl == cl, OFP_E_STATE, "Verlet list not built from this cell list");
    x->voff = vl->off.as<unsigned long long>();
    x->vnbr = vl->nbr.as<uint32_t>();
  }
  return OFP_OK;
}

template <class Ph>
static int launch_f64(const ofp_cl* cl, const PairCtx& x, const void* pos, size_t stride, const Ph& ph,
                      hipStream_t s) {
  if (x.c.n_dom == 0) return OFP_OK;
  OFP_CHECK(pos, OFP_E_ARG, "null sorted positions");
  OFP_CHECK(stride >= x.v.n_part, OFP_E_ARG, "stride < n_part");
  unsigned grid = grid_for(x.c.n_dom, 256);
  if (x.v.dtype == OFP_F32) {
    CellGeom<float> g;
    for (int k = 0; k < 3; ++k) {
      g.lo[k] = float(x.v.lo[k]);
      g.sp[k] = float(x.v.spacing[k]);
      g.divc[k] = x.v.divc[k];
    }
    if (x.voff)
      k_pairs_f64<float, Ph, true><<<grid, 256, 0, s>>>(x.c, g, (const float*)pos, stride, x.voff, x.vnbr, ph);
    else
      k_pairs_f64<float, Ph, false><<<grid, 256, 0, s>>>(x.c, g, (const float*)pos, stride, x.voff, x.vnbr, ph);
  } else {
    CellGeom<double> g;
    for (int k = 0; k < 3; ++k) {
      g.lo[k] = x.v.lo[k];
      g.sp[k] = x.v.spacing[k];
      g.divc[k] = x.v.divc[k];
    }
    if (x.voff)
      k_pairs_f64<double, Ph, true><<<grid, 256, 0, s>>>(x.c, g, (const double*)pos, stride, x.voff, x.vnbr, ph);
    else
      k_pairs_f64<double, Ph, false><<<grid, 256, 0, s>>>(x.c, g, (const double*)pos, stride, x.voff, x.vnbr, ph);
  }
  OFP_LAUNCHED();
  return OFP_OK;
}

template <class Ph>
static int launch_f32c(const PairCtx& x, const Ph& ph, hipStream_t s) {
  if (x.c.n_dom == 0) return OFP_OK;
  QScale qs;
  for (int k = 0; k < 3; ++k) qs.q[k] = float(x.v.spacing[k] / double(kQuantOne));
  unsigned grid = grid_for(x.c.n_dom, 256);
  if (x.voff)
    k_pairs_f32c<Ph, true><<<grid, 256, 0, s>>>(x.c, qs, x.voff, x.vnbr, ph);
  else
    k_pairs_f32c<Ph, false><<<grid, 256, 0, s>>>(x.c, qs, x.voff, x.vnbr, ph);
  OFP_LAUNCHED();
  return OFP_OK;
}

template <class T>
static int vl_build(ofp_vl* h, const ofp_cl* cl, const void* pos, size_t stride, double r_cut,
                    hipStream_t s) {
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(cl_dev(cl, &c, &v));
  CellGeom<T> g;
  for (int k = 0; k < 3; ++k) {
    g.lo[k] = T(v.lo[k]);
    g.sp[k] = T(v.spacing[k]);
    g.divc[k] = v.divc[k];
    OFP_CHECK(v.spacing[k] >= r_cut * (1.0 - 1e-6), OFP_E_ARG,
              "cell spacing %g smaller than the Verlet radius %g", v.spacing[k], r_cut);
  }
  const size_t n = c.n_dom;
  OFP_TRY(h->cnt.ensure((n + 1) * 4));
  OFP_TRY(h->off.ensure((n + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(n)));
  const T rc = T(r_cut);
  const T r2 = rc * rc;
  if (n > 0) {
    k_vl_pass<T, false><<<grid_for(n, 256), 256, 0, s>>>(c, g, (const T*)pos, stride, r2,
                                                         h->cnt.as<uint32_t>(), nullptr, nullptr);
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
  uint64_t E = 0;
  OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  OFP_TRY(h->nbr.ensure((E + 1) * 4));
  if (n > 0) {
    k_vl_pass<T, true><<<grid_for(n, 256), 256, 0, s>>>(c, g, (const T*)pos, stride, r2, nullptr,
                                                        h->off.as<unsigned long long>(),
                                                        h->nbr.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_rows = n;
  h->n_entries = E;
  h->built = true;
  h->cl = cl;
  return OFP_OK;
}

}  // namespace ofp

using namespace ofp;

extern "C" {

int ofp_vl_create(ofp_vl** h) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_vl_create: null handle pointer");
  *h = new ofp_vl();
  return OFP_OK;
}

int ofp_vl_destroy(ofp_vl* h) {
  if (!h) return OFP_OK;
  h->cnt.release();
  h->off.release();
  h->nbr.release();
  h->scantmp.release();
  delete h;
  return OFP_OK;
}

int ofp_vl_build(ofp_vl* h, const ofp_cl* cl, const void* pos_sorted, size_t stride, double r_cut,
                 void* stream) {
  OFP_CHECK(h && cl, OFP_E_ARG, "ofp_vl_build: null handle");
  OFP_CHECK(r_cut > 0, OFP_E_ARG, "r_cut must be positive");
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(cl, &v));
  OFP_CHECK(v.n_part == 0 || pos_sorted, OFP_E_ARG, "null sorted positions");
  OFP_CHECK(stride >= v.n_part, OFP_E_ARG, "stride < n_part");
  if (v.dtype == OFP_F32) return vl_build<float>(h, cl, pos_sorted, stride, r_cut, as_stream(stream));
  return vl_build<double>(h, cl, pos_sorted, stride, r_cut, as_stream(stream));
}

int ofp_vl_get_view(const ofp_vl* h, ofp_vl_view* v) {
  OFP_CHECK(h && v, OFP_E_ARG, "ofp_vl_get_view: null argument");
  OFP_CHECK(h->built, OFP_E_STATE, "Verlet list not built");
  v->offsets = h->off.as<uint64_t>();
  v->nbr = h->nbr.as<uint32_t>();
  v->n_rows = h->n_rows;
  v->n_entries = h->n_entries;
  return OFP_OK;
}

int ofp_pair_lj(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, int precision,
                const ofp_lj_params* prm, void* fx, void* fy, void* fz, void* energy,
                uint32_t* n_neigh, void* stream) {
  OFP_CHECK(prm, OFP_E_ARG, "ofp_pair_lj: null parameters");
  OFP_CHECK((fx && fy && fz) || (!fx && !fy && !fz), OFP_E_ARG, "force planes: all or none");
  PairCtx x;
  OFP_TRY(pair_ctx(cl, vl, precision, &x));
  OFP_CHECK(prm->r_cut <= x.v.spacing[0] * (1 + 1e-6) || vl, OFP_E_ARG,
            "cut-off larger than the cell size");
  hipStream_t s = as_stream(stream);
  if (precision == OFP_PREC_F64) {
    LJ64 ph{prm->r_cut * prm->r_cut, prm->sigma * prm->sigma, 24.0 * prm->eps, 4.0 * prm->eps,
            (double*)fx, (double*)fy, (double*)fz, (double*)energy, n_neigh};
    return launch_f64(cl, x, pos, stride, ph, s);
  }
  LJ32 ph{float(prm->r_cut * prm->r_cut), float(prm->sigma * prm->sigma), float(24.0 * prm->eps),
          float(4.0 * prm->eps), (float*)fx, (float*)fy, (float*)fz, (float*)energy, n_neigh};
  return launch_f32c(x, ph, s);
}

int ofp_sph_density(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride,
                    int precision, const ofp_sph_params* prm, void* rho, void* stream) {
  OFP_CHECK(prm && rho, OFP_E_ARG, "ofp_sph_density: null argument");
  PairCtx x;
  OFP_TRY(pair_ctx(cl, vl, precision, &x));
  const double h = prm->h;
  OFP_CHECK(2 * h <= x.v.spacing[0] * (1 + 1e-6) || vl, OFP_E_ARG, "support 2h larger than the cell size");
  hipStream_t s = as_stream(stream);
  const double sk = 1.0 / (M_PI * h * h * h);
  if (precision == OFP_PREC_F64) {
    Dens64 ph{4 * h * h, 1.0 / h, sk, prm->mass, (double*)rho};
    return launch_f64(cl, x, pos, stride, ph, s);
  }
  Dens32 ph{float(4 * h * h), float(1.0 / h), float(sk), float(prm->mass), (float*)rho};
  return launch_f32c(x, ph, s);
}

int ofp_sph_eos(int precision, const void* rho, size_t n, const ofp_sph_params* prm, void* prs,
                void* stream) {
  OFP_CHECK(prm && (n == 0 || (rho && prs)), OFP_E_ARG, "ofp_sph_eos: null argument");
  if (n == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  const double B = prm->c0 * prm->c0 * prm->rho0 / 7.0;
  if (precision == OFP_PREC_F64)
    k_tait<double><<<grid_for(n, 256), 256, 0, s>>>((const double*)rho, n, 1.0 / prm->rho0, B, (double*)prs);
  else if (precision == OFP_PREC_F32C)
    k_tait<float><<<grid_for(n, 256), 256, 0, s>>>((const float*)rho, n, float(1.0 / prm->rho0), float(B),
                                                   (float*)prs);
  else
    return set_error(OFP_E_ARG, "bad precision %d", precision);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_sph_force(const ofp_cl* cl, const ofp_vl* vl, const void* pos, size_t stride, int precision,
                  const ofp_sph_params* prm, const void* rho, const void* prs, void* fx, void* fy,
                  void* fz, void* stream) {
  OFP_CHECK(prm && rho && prs && fx && fy && fz, OFP_E_ARG, "ofp_sph_force: null argument");
  PairCtx x;
  OFP_TRY(pair_ctx(cl, vl, precision, &x));
  const double h = prm->h;
  OFP_CHECK(2 * h <= x.v.spacing[0] * (1 + 1e-6) || vl, OFP_E_ARG, "support 2h larger than the cell size");
  hipStream_t s = as_stream(stream);
  const double sd = 1.0 / (M_PI * h * h * h * h);
  if (precision == OFP_PREC_F64) {
    Force64 ph{4 * h * h, 1.0 / h, sd, prm->mass, (const double*)rho, (const double*)prs,
               (double*)fx, (double*)fy, (double*)fz};
    return launch_f64(cl, x, pos, stride, ph, s);
  }
  Force32 ph{float(4 * h * h), float(1.0 / h), float(sd), float(prm->mass), (const float*)rho,
             (const float*)prs, (float*)fx, (float*)fy, (float*)fz};
  return launch_f32c(x, ph, s);
}

int ofp_reduce(int dtype, const void* x, const void* comp, size_t n, int op, double* result,
               void* stream) {
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
  k_reduce_final<<<1, kRedBlock, 0, s>>>(partial, grid, op, ident);
  OFP_LAUNCHED();
  double r = 0;
  OFP_HIP(hipMemcpyAsync(&r, partial + grid, sizeof(double), hipMemcpyDeviceToHost, s));
  OFP_HIP(hipFreeAsync(partial, s));
  OFP_HIP(hipStreamSynchronize(s));
  *result = r;
  return OFP_OK;
}

}  // extern "C"
