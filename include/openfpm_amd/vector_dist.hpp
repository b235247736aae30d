// openfpm_amd/vector_dist.hpp -- vector_dist_gpu surface for the one-GPU
// neighbour-search path on MI355X, header-only, over the C ABI of libofp_nn.so.
//
// User code written against openfpm_pdata's GPU API recompiles with hipcc:
//   vector_dist_gpu<3,float,aggregate<float,float[3]>> vd(n, domain, bc, ghost);
//   vd.hostToDevicePos(); vd.template ghost_get<0>(RUN_ON_DEVICE);
//   auto NN = vd.getCellListGPU(r_cut, CL_NON_SYMMETRIC | CL_GPU_REORDER);
//   vd.template updateCellListGPU<0>(NN);
//   CUDA_LAUNCH(kernel, vd.getDomainIteratorGPU(64), vd.toKernel(), NN.toKernel());
//   vd.template restoreOrder<1>(NN);
// Reference surface (src/Vector/vector_dist.hpp): ctor :498, map :2158,
// ghost_get :2203, ghost_put :2304, getCellListGPU<CellType> :1222 (dense
// CellList_gpu or CELLLIST_GPU_SPARSE :49), updateCellListGPU<prp...,
// CellList_type> :1334, restoreOrder :3042, getVerletSym :1412, toKernel :2961,
// getDomainIteratorGPU :1958, host<->device copies :2984-3030; device views
// src/Vector/vector_dist_kernel.hpp (GET_PARTICLE, GET_PARTICLE_SORT :15-17).
//
// Storage: positions are 3 SoA planes of `capacity` elements; each property
// of aggregate<P...> is stored as SoA planes too (a T[n] property is n planes,
// T[n][m] is n*m planes), the memory_traits_inte layout of the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ofp_io.h"
#include "ofp_nn.h"

// ------------------------------------------------------------ options
// ghost_get options: NO_POSITION, WITH_POSITION, NO_CHANGE_ELEMENTS
// (Vector/cuda/vector_dist_comm_util_funcs.cuh:13-15); SKIP_LABELLING and
// KEEP_PROPERTIES share one value (util/common_pdata.hpp:12-13), so either name
// asks for the refresh of the cached ghost list.
#ifndef NO_POSITION
#define NO_POSITION 1
#define WITH_POSITION 2
#define NO_CHANGE_ELEMENTS 4
#define SKIP_LABELLING 512
#define KEEP_PROPERTIES 512
#define RUN_ON_DEVICE 1024
#endif
// cell-list options (vector_dist.hpp:1340-1372)
#ifndef CL_NON_SYMMETRIC
#define CL_NON_SYMMETRIC 1
#define CL_GPU_REORDER_POSITION 256
#define CL_GPU_REORDER_PROPERTY 512
#define CL_GPU_REORDER (CL_GPU_REORDER_POSITION | CL_GPU_REORDER_PROPERTY)
#endif
#ifndef PERIODIC
#define NON_PERIODIC 0
#define PERIODIC 1
#endif

namespace openfpm_amd {

inline void check(int rc, const char* what) {
  if (rc != OFP_OK)
    throw std::runtime_error(std::string(what) + ": " + ofp_last_error());
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
constexpr int dtype_of() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "positions must be float or double");
  return std::is_same<T, float>::value ? OFP_F32 : OFP_F64;
}

// property type -> element type and number of SoA planes
template <class P>
struct prop_traits {
  using elem = P;
  static constexpr int n1 = 1, n2 = 1, planes = 1;
};
template <class E, size_t N>
struct prop_traits<E[N]> {
  using elem = E;
  static constexpr int n1 = int(N), n2 = 1, planes = int(N);
};
template <class E, size_t N, size_t M>
struct prop_traits<E[N][M]> {
  using elem = E;
  static constexpr int n1 = int(N), n2 = int(M), planes = int(N * M);
};

}  // namespace openfpm_amd

// --------------------------------------------------------- small types
template <typename... T>
struct aggregate {
  static constexpr int max_prop = int(sizeof...(T));
};

template <unsigned int dim, typename T>
struct Point {
  T x[dim];
  __host__ __device__ Point() {
    for (unsigned i = 0; i < dim; ++i) x[i] = T(0);
  }
  __host__ __device__ Point(std::initializer_list<T> l) {
    unsigned i = 0;
    for (T v : l) x[i++] = v;
  }
  __host__ __device__ T& operator[](unsigned i) { return x[i]; }
  __host__ __device__ const T& operator[](unsigned i) const { return x[i]; }
  __host__ __device__ T get(unsigned i) const { return x[i]; }
  __host__ __device__ Point operator-(const Point& o) const {
    Point r;
    for (unsigned i = 0; i < dim; ++i) r.x[i] = x[i] - o.x[i];
    return r;
  }
  __host__ __device__ Point operator+(const Point& o) const {
    Point r;
    for (unsigned i = 0; i < dim; ++i) r.x[i] = x[i] + o.x[i];
    return r;
  }
  __host__ __device__ Point& operator+=(const Point& o) {
    for (unsigned i = 0; i < dim; ++i) x[i] += o.x[i];
    return *this;
  }
  __host__ __device__ Point& operator/=(T v) {
    for (unsigned i = 0; i < dim; ++i) x[i] /= v;
    return *this;
  }
  __host__ __device__ T norm2() const {
    T r = 0;
    for (unsigned i = 0; i < dim; ++i) r += x[i] * x[i];
    return r;
  }
  __host__ __device__ T norm() const { return sqrt(norm2()); }
  __host__ __device__ T distance(const Point& o) const { return (*this - o).norm(); }
};
template <unsigned int dim, typename T>
__host__ __device__ Point<dim, T> operator*(T s, const Point<dim, T>& p) {
  Point<dim, T> r;
  for (unsigned i = 0; i < dim; ++i) r.x[i] = s * p.x[i];
  return r;
}

template <unsigned int dim, typename T>
struct Box {
  T lo[dim], hi[dim];
  Box() {
    for (unsigned i = 0; i < dim; ++i) lo[i] = hi[i] = T(0);
  }
  Box(std::initializer_list<T> l, std::initializer_list<T> h) {
    unsigned i = 0;
    for (T v : l) lo[i++] = v;
    i = 0;
    for (T v : h) hi[i++] = v;
  }
  T getLow(unsigned i) const { return lo[i]; }
  T getHigh(unsigned i) const { return hi[i]; }
};

// Ghost<dim,T> g(r): the box [-r, r] (Decomposition ghost).
template <unsigned int dim, typename T>
struct Ghost : public Box<dim, T> {
  explicit Ghost(T r) {
    for (unsigned i = 0; i < dim; ++i) {
      this->lo[i] = -r;
      this->hi[i] = r;
    }
  }
};

// ------------------------------------------------------------ launches
struct ite_gpu {
  dim3 wthr, thr;
};

#define CUDA_LAUNCH(kernel, ite, ...) hipLaunchKernelGGL(kernel, (ite).wthr, (ite).thr, 0, 0, __VA_ARGS__)
#define CUDA_LAUNCH_DIM3(kernel, wthr_, thr_, ...) hipLaunchKernelGGL(kernel, wthr_, thr_, 0, 0, __VA_ARGS__)

// thread -> particle (Vector/vector_dist_kernel.hpp:15-17)
#define GET_PARTICLE(vd)                                                                  \
  blockDim.x* blockIdx.x + threadIdx.x;                                                   \
  if (blockDim.x * blockIdx.x + threadIdx.x >= static_cast<unsigned int>(vd.size_local())) { \
    return;                                                                               \
  };
#define GET_PARTICLE_SORT(p, NN)                                                          \
  if (blockDim.x * blockIdx.x + threadIdx.x >= NN.getGhostMarker()) {                     \
    return;                                                                               \
  } else {                                                                                \
    p = NN.getDomainSortIds().template get<0>(blockDim.x * blockIdx.x + threadIdx.x);     \
  }

// --------------------------------------------------------- device views
namespace openfpm_amd {

template <class E>
struct plane_ref1 {  // T[n] property of one particle
  E* base;
  size_t stride, p;
  __host__ __device__ E& operator[](int k) const { return base[size_t(k) * stride + p]; }
};
template <class E, int M>
struct plane_ref2 {  // T[n][m] property of one particle
  E* base;
  size_t stride, p;
  __host__ __device__ plane_ref1<E> operator[](int k) const {
    return plane_ref1<E>{base + size_t(k) * M * stride, stride, p};
  }
};

template <class P>
struct prop_access {
  using tr = prop_traits<P>;
  using E = typename tr::elem;
  __host__ __device__ static E& get(E* base, size_t, size_t p) { return base[p]; }
};
template <class E, size_t N>
struct prop_access<E[N]> {
  __host__ __device__ static plane_ref1<E> get(E* base, size_t stride, size_t p) {
    return plane_ref1<E>{base, stride, p};
  }
};
template <class E, size_t N, size_t M>
struct prop_access<E[N][M]> {
  __host__ __device__ static plane_ref2<E, int(M)> get(E* base, size_t stride, size_t p) {
    return plane_ref2<E, int(M)>{base, stride, p};
  }
};

template <class... P>
struct type_list {};
template <int i, class L>
struct nth;
template <int i, class H, class... T>
struct nth<i, type_list<H, T...>> : nth<i - 1, type_list<T...>> {};
template <class H, class... T>
struct nth<0, type_list<H, T...>> {
  using type = H;
};

template <class A>
struct agg_list;
template <class... P>
struct agg_list<aggregate<P...>> {
  using type = type_list<P...>;
  static constexpr int n = int(sizeof...(P));
};

template <class St>
struct pos_ref {
  St* base;
  size_t stride, p;
  __host__ __device__ St& operator[](int k) const { return base[size_t(k) * stride + p]; }
  template <unsigned dim>
  __host__ __device__ operator Point<dim, St>() const {
    Point<dim, St> r;
    for (unsigned k = 0; k < dim; ++k) r.x[k] = base[size_t(k) * stride + p];
    return r;
  }
};

}  // namespace openfpm_amd

// Device view of a vector_dist (vector_dist_ker, Vector/vector_dist_kernel.hpp:57-387).
template <unsigned int dim, typename St, typename prop>
struct vector_dist_ker {
  static constexpr int n_prop = openfpm_amd::agg_list<prop>::n;
  St* pos;
  size_t stride;
  int gm;
  int n_all;
  void* prp[n_prop > 0 ? n_prop : 1];

  __host__ __device__ int size_local() const { return gm; }
  __host__ __device__ int size() const { return n_all; }
  __host__ __device__ openfpm_amd::pos_ref<St> getPos(size_t p) const {
    return openfpm_amd::pos_ref<St>{pos, stride, p};
  }
  template <unsigned int id>
  __host__ __device__ decltype(auto) getProp(size_t p) const {
    using P = typename openfpm_amd::nth<int(id), typename openfpm_amd::agg_list<prop>::type>::type;
    using E = typename openfpm_amd::prop_traits<P>::elem;
    return openfpm_amd::prop_access<P>::get(static_cast<E*>(prp[id]), stride, p);
  }
};

// Cell key returned by getCell (grid_key_dx<3> in the reference); m is the
// cell's entry in a sparse list (-1: dense list, or a cell holding no particle).
struct ofp_cell_key {
  int c[3];
  int m;
};

// Template arguments of the reference's CellList_gpu<dim, T, Memory, transform,
// is_sparse> (openfpm_data): tags here, so user code naming the full type compiles.
struct CudaMemory {};
template <unsigned int dim, typename St>
struct shift_only {};

// ids list view with the reference's `.template get<0>(i)` accessor
struct ofp_ids_view {
  const uint32_t* ids;
  template <unsigned int>
  __host__ __device__ uint32_t get(size_t i) const {
    return ids[i];
  }
};

template <unsigned int dim, typename St>
struct CellList_gpu_ker;

// Neighbour iterator of getNNIteratorBox: the (2B+1)^3 stencil as (2B+1)^2
// contiguous runs of sorted indices (cells of one x-row are consecutive).
template <unsigned int dim, typename St>
struct NN_gpu_it {
  const CellList_gpu_ker<dim, St>* cl;
  int cx0, cx1, cy, cz;  // run x-range and centre
  int dy, dz, B;
  uint32_t j, e;
  __device__ void load_run();
  __device__ bool isNext() const { return j < e; }
  __device__ uint32_t get_sort() const { return j; }
  __device__ uint32_t get() const;
  __device__ NN_gpu_it& operator++() {
    ++j;
    if (j >= e) next_run();
    return *this;
  }
  __device__ void next_run() {
    while (j >= e) {
      if (++dy > B) {
        dy = -B;
        if (++dz > B) return;
      }
      load_run();
    }
  }
};

// CellList_gpu::toKernel() view.
template <unsigned int dim, typename St>
struct CellList_gpu_ker {
  const unsigned long long* cs;
  const uint32_t* s2n;
  const uint32_t* dom;
  St lo[3], sp[3];
  uint32_t divc[3];
  uint32_t gm;
  int box_nn;

  __device__ ofp_cell_key getCell(const Point<dim, St>& x) const {
    ofp_cell_key k;
    for (int d = 0; d < 3; ++d) {
#pragma clang fp contract(off)
      St f = floor((x.x[d] - lo[d]) / sp[d]);
      int v = (f != f) ? 0 : (f < St(-1) ? 0 : (f > St(divc[d]) - St(2) ? int(divc[d]) - 1 : int(f) + 1));
      k.c[d] = v;
    }
    k.m = -1;
    return k;
  }
  __device__ NN_gpu_it<dim, St> getNNIteratorBox(const ofp_cell_key& k) const {
    NN_gpu_it<dim, St> it;
    it.cl = this;
    it.B = box_nn;
    it.cx0 = k.c[0] - box_nn > 0 ? k.c[0] - box_nn : 0;
    it.cx1 = k.c[0] + box_nn < int(divc[0]) - 1 ? k.c[0] + box_nn : int(divc[0]) - 1;
    it.cy = k.c[1];
    it.cz = k.c[2];
    it.dy = -box_nn;
    it.dz = -box_nn;
    it.j = it.e = 0;
    it.load_run();
    it.next_run();
    return it;
  }
  __device__ ofp_ids_view getDomainSortIds() const { return ofp_ids_view{dom}; }
  __device__ unsigned int getGhostMarker() const { return gm; }
  __device__ uint32_t getNCells() const { return divc[0] * divc[1] * divc[2]; }
  __device__ uint32_t getNelements(uint32_t c) const { return uint32_t(cs[c + 1]) - uint32_t(cs[c]); }
  __device__ uint32_t get(uint32_t c, uint32_t k) const { return s2n[uint32_t(cs[c]) + k]; }
};

template <unsigned int dim, typename St>
__device__ void NN_gpu_it<dim, St>::load_run() {
  const int yy = cy + dy, zz = cz + dz;
  if (dz > B || yy < 0 || zz < 0 || yy >= int(cl->divc[1]) || zz >= int(cl->divc[2])) {
    j = e = 0;
    return;
  }
  const uint32_t row = cl->divc[0] * (uint32_t(yy) + cl->divc[1] * uint32_t(zz));
  j = uint32_t(cl->cs[row + uint32_t(cx0)]);
  e = uint32_t(cl->cs[row + uint32_t(cx1) + 1]);
}

template <unsigned int dim, typename St>
__device__ uint32_t NN_gpu_it<dim, St>::get() const {
  return cl->s2n[j];
}

// Neighbour iterator of a sparse list: the non-empty cells of the centre's
// stencil table, each a run of sorted indices.
struct NN_gpu_it_sparse {
  const int32_t* nn;  // the stencil's entries of the centre cell (k cells)
  const uint32_t* start;
  const uint32_t* s2n;
  int k, K;
  uint32_t j, e;
  __device__ bool isNext() const { return j < e; }
  __device__ uint32_t get_sort() const { return j; }
  __device__ uint32_t get() const { return s2n[j]; }
  __device__ void next_cell() {
    while (j >= e && ++k < K) {
      const int m = nn[k];
      if (m >= 0) {
        j = start[m];
        e = start[m + 1];
      }
    }
  }
  __device__ NN_gpu_it_sparse& operator++() {
    ++j;
    if (j >= e) next_cell();
    return *this;
  }
};

// toKernel() view of CELLLIST_GPU_SPARSE: only the non-empty cells (ascending
// id) with their (2B + 1)^3 stencil tables.  getCell finds the entry of the
// position's cell by binary search; a cell holding no particle has none, and
// its iterator is empty (pair loops ask for the cells of their own particles).
template <unsigned int dim, typename St>
struct CellList_gpu_ker_sparse {
  const uint32_t* key;    // cell id of each entry
  const uint32_t* start;  // n_nz + 1 sorted-index starts
  const int32_t* nn;      // K entries per cell
  const uint32_t* s2n;
  const uint32_t* dom;
  St lo[3], sp[3];
  uint32_t divc[3];
  uint32_t gm, n_nz;
  int K;

  __device__ ofp_cell_key getCell(const Point<dim, St>& x) const {
    ofp_cell_key k;
    for (int d = 0; d < 3; ++d) {
#pragma clang fp contract(off)
      St f = floor((x.x[d] - lo[d]) / sp[d]);
      int v = (f != f) ? 0 : (f < St(-1) ? 0 : (f > St(divc[d]) - St(2) ? int(divc[d]) - 1 : int(f) + 1));
      k.c[d] = v;
    }
    const uint32_t c = uint32_t(k.c[0]) + divc[0] * (uint32_t(k.c[1]) + divc[1] * uint32_t(k.c[2]));
    uint32_t a = 0, b = n_nz;
    while (a < b) {
      const uint32_t h = (a + b) / 2;
      if (key[h] < c)
        a = h + 1;
      else
        b = h;
    }
    k.m = (a < n_nz && key[a] == c) ? int(a) : -1;
    return k;
  }
  __device__ NN_gpu_it_sparse getNNIteratorBox(const ofp_cell_key& k) const {
    NN_gpu_it_sparse it;
    it.nn = k.m >= 0 ? nn + size_t(k.m) * size_t(K) : nullptr;
    it.K = k.m >= 0 ? K : 0;
    it.start = start;
    it.s2n = s2n;
    it.k = -1;
    it.j = it.e = 0;
    it.next_cell();
    return it;
  }
  __device__ ofp_ids_view getDomainSortIds() const { return ofp_ids_view{dom}; }
  __device__ unsigned int getGhostMarker() const { return gm; }
  __device__ uint32_t getNCells() const { return n_nz; }
  __device__ uint32_t getNelements(uint32_t m) const { return start[m + 1] - start[m]; }
  __device__ uint32_t get(uint32_t m, uint32_t i) const { return s2n[start[m] + i]; }
};

// ------------------------------------------------------ host containers
// CellList_gpu<dim, St, Memory, transform, is_sparse>: the dense list, or with
// is_sparse its non-empty cells only (built after each construct).
template <unsigned int dim, typename St, typename Memory = CudaMemory, typename transform = shift_only<dim, St>,
          bool is_sparse_ = false>
class CellList_gpu {
 public:
  static constexpr bool is_sparse = is_sparse_;
  using ker_type = typename std::conditional<is_sparse_, CellList_gpu_ker_sparse<dim, St>,
                                             CellList_gpu_ker<dim, St>>::type;
  ofp_cl* h = nullptr;
  size_t opt = 0;
  int box_nn = 1;
  St lo[3], hi[3];
  uint32_t div[3];
  St r_cut;

  CellList_gpu() = default;
  CellList_gpu(const St* lo_, const St* hi_, const uint32_t* div_, size_t opt_, int box_nn_, St rc)
      : opt(opt_), box_nn(box_nn_), r_cut(rc) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = lo_[k];
      hi[k] = hi_[k];
      div[k] = div_[k];
    }
    openfpm_amd::check(ofp_cl_create(&h, openfpm_amd::dtype_of<St>(), lo, hi, div, uint32_t(opt) | OFP_CL_QUANT),
                       "ofp_cl_create");
    if (box_nn != 1) openfpm_amd::check(ofp_cl_set_box_nn(h, uint32_t(box_nn)), "ofp_cl_set_box_nn");
  }
  CellList_gpu(const CellList_gpu&) = delete;
  CellList_gpu& operator=(const CellList_gpu&) = delete;
  CellList_gpu(CellList_gpu&& o) noexcept { *this = std::move(o); }
  CellList_gpu& operator=(CellList_gpu&& o) noexcept {
    std::swap(h, o.h);
    opt = o.opt;
    box_nn = o.box_nn;
    r_cut = o.r_cut;
    for (int k = 0; k < 3; ++k) {
      lo[k] = o.lo[k];
      hi[k] = o.hi[k];
      div[k] = o.div[k];
    }
    return *this;
  }
  ~CellList_gpu() {
    if (h) ofp_cl_destroy(h);
  }
  size_t getOpt() const { return opt; }
  int getBoxNN() const { return box_nn; }

  ker_type toKernel() const {
    ofp_cl_view v;
    openfpm_amd::check(ofp_cl_get_view(h, &v), "toKernel: cell list not constructed");
    ker_type k;
    if constexpr (is_sparse_) {
      ofp_cl_sparse_view sv;
      openfpm_amd::check(ofp_cl_get_sparse_view(h, &sv), "toKernel: sparse view");
      k.key = sv.cell_key;
      k.start = sv.cell_start;
      k.nn = sv.cell_nn;
      k.n_nz = uint32_t(sv.n_cells);
      k.K = int(sv.nn_per_cell);
    } else {
      k.cs = reinterpret_cast<const unsigned long long*>(v.cell_start);
      k.box_nn = box_nn;
    }
    k.s2n = v.sort_to_nonsort;
    k.dom = v.dom_sort_ids;
    for (int d = 0; d < 3; ++d) {
      k.lo[d] = St(v.lo[d]);
      k.sp[d] = St(v.spacing[d]);
      k.divc[d] = v.divc[d];
    }
    k.gm = uint32_t(v.ghost_marker);
    return k;
  }
};

// CELLLIST_GPU_SPARSE (vector_dist.hpp:49).
template <unsigned int dim, typename St>
using CELLLIST_GPU_SPARSE = CellList_gpu<dim, St, CudaMemory, shift_only<dim, St>, true>;

// Reduction of ghost_put<op, prp...>: add_ (the ghosts' values added to their
// sources').
template <typename A, typename B>
struct add_ {};

// Verlet list on the device (CSR over the domain rows of its cell list).
template <unsigned int dim, typename St>
struct VerletList_gpu_ker {
  const uint64_t* off;
  const uint32_t* nbr;
  const uint32_t* dom;
  uint64_t n_rows;
  __device__ uint32_t getNNPart(uint32_t row) const { return uint32_t(off[row + 1] - off[row]); }
  __device__ uint32_t get(uint32_t row, uint32_t k) const { return nbr[off[row] + k]; }
  __device__ uint32_t particle(uint32_t row) const { return dom[row]; }
};

template <unsigned int dim, typename St>
class VerletList_gpu {
 public:
  ofp_vl* h = nullptr;
  ofp_vl* full = nullptr;  // a symmetric list: the full list it was made from
  CellList_gpu<dim, St> cl;
  VerletList_gpu() { openfpm_amd::check(ofp_vl_create(&h), "ofp_vl_create"); }
  VerletList_gpu(const VerletList_gpu&) = delete;
  VerletList_gpu(VerletList_gpu&& o) noexcept : cl(std::move(o.cl)) {
    std::swap(h, o.h);
    std::swap(full, o.full);
  }
  ~VerletList_gpu() {
    if (h) ofp_vl_destroy(h);
    if (full) ofp_vl_destroy(full);
  }
  bool symmetric() const { return full != nullptr; }
  uint64_t size() const {
    ofp_vl_view v;
    openfpm_amd::check(ofp_vl_get_view(h, &v), "Verlet view");
    return v.n_rows;
  }
  uint64_t n_entries() const {
    ofp_vl_view v;
    openfpm_amd::check(ofp_vl_get_view(h, &v), "Verlet view");
    return v.n_entries;
  }
  VerletList_gpu_ker<dim, St> toKernel() const {
    ofp_vl_view v;
    openfpm_amd::check(ofp_vl_get_view(h, &v), "Verlet view");
    ofp_cl_view c;
    openfpm_amd::check(ofp_cl_get_view(cl.h, &c), "cell view");
    return VerletList_gpu_ker<dim, St>{v.offsets, v.nbr, c.dom_sort_ids, v.n_rows};
  }
};

// Host-side key of the domain/ghost iterators (vect_dist_key_dx).
struct vect_dist_key_dx {
  size_t key;
  size_t getKey() const { return key; }
};

struct vector_dist_iterator {
  size_t i, end;
  bool isNext() const { return i < end; }
  vect_dist_key_dx get() const { return vect_dist_key_dx{i}; }
  vector_dist_iterator& operator++() {
    ++i;
    return *this;
  }
};

template <unsigned int dim, typename St, typename prop>
class vector_dist_gpu {
  static_assert(dim == 3, "this build provides the 3-D path");
  using props = openfpm_amd::agg_list<prop>;
  static constexpr int n_prop = props::n;

  size_t cap_ = 0, gm_ = 0, size_ = 0;
  St dom_lo_[3], dom_hi_[3], g_lo_[3], g_hi_[3];
  int bc_[3];
  ofp_ghost* gh_ = nullptr;
  // device storage (current and reorder buffers)
  St* pos_ = nullptr;
  St* pos_alt_ = nullptr;
  void* prp_[n_prop > 0 ? n_prop : 1] = {};
  void* prp_alt_[n_prop > 0 ? n_prop : 1] = {};
  size_t prp_esz_[n_prop > 0 ? n_prop : 1] = {};
  int prp_planes_[n_prop > 0 ? n_prop : 1] = {};
  // host mirrors
  std::vector<St> hpos_;
  std::vector<std::vector<unsigned char>> hprp_;
  bool sorted_ = false;
  const ofp_cl* sorted_cl_ = nullptr;  // the cell list whose order the vector is in
  hipStream_t stream_ = nullptr;

  // checkpoint fields of property I (host mirror planes) at their natural
  // alignment inside the packed object (include/ofp_io.h)
  template <int I>
  void io_fields(std::vector<ofp_h5_field>& f, size_t& off, size_t& align) {
    if constexpr (I < n_prop) {
      using P = typename openfpm_amd::nth<I, typename props::type>::type;
      using E = typename openfpm_amd::prop_traits<P>::elem;
      off = (off + alignof(E) - 1) / alignof(E) * alignof(E);
      for (int k = 0; k < prp_planes_[I]; ++k)
        f.push_back(ofp_h5_field{hprp_[I].data() + size_t(k) * cap_ * sizeof(E), uint32_t(sizeof(E)),
                                 uint32_t(off + size_t(k) * sizeof(E))});
      off += size_t(prp_planes_[I]) * sizeof(E);
      align = alignof(E) > align ? alignof(E) : align;
      io_fields<I + 1>(f, off, align);
    }
  }
  uint32_t io_layout(std::vector<ofp_h5_field>& f) {
    size_t off = 0, align = 1;
    io_fields<0>(f, off, align);
    return uint32_t((off + align - 1) / align * align);
  }
  static void io_check(int rc, const char* what) {
    if (rc != OFP_OK) throw std::runtime_error(std::string(what) + ": " + ofp_io_last_error());
  }

  // ghost_put planes of property I: pointers and OFP_PUT_* value types
  template <int I>
  void put_planes(std::vector<void*>& ptr, std::vector<int>& type) {
    using P = typename openfpm_amd::nth<I, typename props::type>::type;
    using E = typename openfpm_amd::prop_traits<P>::elem;
    constexpr int t = std::is_same<E, float>::value    ? OFP_PUT_F32
                      : std::is_same<E, double>::value ? OFP_PUT_F64
                      : (std::is_integral<E>::value && sizeof(E) == 4)
                          ? OFP_PUT_U32
                          : (std::is_integral<E>::value && sizeof(E) == 8) ? OFP_PUT_I64 : -1;
    static_assert(t >= 0, "ghost_put<add_>: property elements must be float, double or 4/8-byte integers");
    for (int k = 0; k < prp_planes_[I]; ++k) {
      ptr.push_back(static_cast<char*>(prp_[I]) + size_t(k) * cap_ * prp_esz_[I]);
      type.push_back(t);
    }
  }

  template <int I>
  void init_props() {
    if constexpr (I < n_prop) {
      using P = typename openfpm_amd::nth<I, typename props::type>::type;
      using tr = openfpm_amd::prop_traits<P>;
      prp_esz_[I] = sizeof(typename tr::elem);
      prp_planes_[I] = tr::planes;
      static_assert(sizeof(typename tr::elem) == 4 || sizeof(typename tr::elem) == 8,
                    "property elements must be 4 or 8 bytes");
      init_props<I + 1>();
    }
  }

  void alloc(size_t cap, size_t keep) {
    St* np = nullptr;
    openfpm_amd::hip_check(hipMalloc(&np, 3 * cap * sizeof(St)), "hipMalloc pos");
    if (pos_ && keep)
      for (int k = 0; k < 3; ++k)
        openfpm_amd::hip_check(hipMemcpy(np + k * cap, pos_ + k * cap_, keep * sizeof(St), hipMemcpyDeviceToDevice),
                               "copy pos");
    if (pos_) (void)hipFree(pos_);
    if (pos_alt_) (void)hipFree(pos_alt_);
    pos_alt_ = nullptr;
    pos_ = np;
    for (int i = 0; i < n_prop; ++i) {
      void* q = nullptr;
      const size_t pl = size_t(prp_planes_[i]);
      openfpm_amd::hip_check(hipMalloc(&q, pl * cap * prp_esz_[i]), "hipMalloc prop");
      if (prp_[i] && keep)
        for (size_t k = 0; k < pl; ++k)
          openfpm_amd::hip_check(hipMemcpy(static_cast<char*>(q) + k * cap * prp_esz_[i],
                                           static_cast<char*>(prp_[i]) + k * cap_ * prp_esz_[i],
                                           keep * prp_esz_[i], hipMemcpyDeviceToDevice),
                                 "copy prop");
      if (prp_[i]) (void)hipFree(prp_[i]);
      if (prp_alt_[i]) (void)hipFree(prp_alt_[i]);
      prp_alt_[i] = nullptr;
      prp_[i] = q;
    }
    // host mirrors keep the same capacity layout
    std::vector<St> hp(3 * cap);
    for (int k = 0; k < 3; ++k)
      for (size_t i = 0; i < keep && i < cap_; ++i) hp[k * cap + i] = hpos_[k * cap_ + i];
    hpos_.swap(hp);
    hprp_.resize(n_prop);
    for (int i = 0; i < n_prop; ++i) {
      std::vector<unsigned char> b(size_t(prp_planes_[i]) * cap * prp_esz_[i]);
      for (int k = 0; k < prp_planes_[i]; ++k)
        if (keep && !hprp_[i].empty())
          std::memcpy(b.data() + k * cap * prp_esz_[i], hprp_[i].data() + k * cap_ * prp_esz_[i],
                      keep * prp_esz_[i]);
      hprp_[i].swap(b);
    }
    cap_ = cap;
  }

  void ensure_alt() {
    if (!pos_alt_) openfpm_amd::hip_check(hipMalloc(&pos_alt_, 3 * cap_ * sizeof(St)), "hipMalloc pos_alt");
    for (int i = 0; i < n_prop; ++i)
      if (!prp_alt_[i])
        openfpm_amd::hip_check(hipMalloc(&prp_alt_[i], size_t(prp_planes_[i]) * cap_ * prp_esz_[i]),
                               "hipMalloc prop_alt");
  }

  void planes_of(int i, void* const* base, std::vector<ofp_plane>& out) const {
    for (int k = 0; k < prp_planes_[i]; ++k)
      out.push_back(ofp_plane{static_cast<char*>(base[i]) + size_t(k) * cap_ * prp_esz_[i],
                              uint32_t(prp_esz_[i]), 0});
  }
  template <unsigned int... ids>
  std::vector<ofp_plane> planes(void* const* base) const {
    std::vector<ofp_plane> v;
    (planes_of(int(ids), base, v), ...);
    return v;
  }
  std::vector<ofp_plane> all_planes(void* const* base) const {
    std::vector<ofp_plane> v;
    for (int i = 0; i < n_prop; ++i) planes_of(i, base, v);
    return v;
  }

 public:
  typedef St stype;
  static constexpr unsigned int dims = dim;

  vector_dist_gpu(size_t np, const Box<dim, St>& domain, const size_t (&bc)[dim], const Ghost<dim, St>& g)
      : gm_(np), size_(np) {
    init_props<0>();
    for (int k = 0; k < 3; ++k) {
      dom_lo_[k] = domain.lo[k];
      dom_hi_[k] = domain.hi[k];
      g_lo_[k] = g.lo[k];
      g_hi_[k] = g.hi[k];
      bc_[k] = int(bc[k]);
    }
    alloc(np + np / 2 + 64, 0);
    // one sub-domain == the domain: local ghost boxes and shift vectors
    St blo[26 * 3], bhi[26 * 3], sh[27 * 3];
    uint32_t sv[26];
    int nb = 0;
    openfpm_amd::check(ofp_shift_vectors(openfpm_amd::dtype_of<St>(), dom_lo_, dom_hi_, sh), "shift vectors");
    openfpm_amd::check(ofp_local_ghost_boxes(openfpm_amd::dtype_of<St>(), dom_lo_, dom_hi_, g_lo_, g_hi_, bc_, blo,
                                             bhi, sv, &nb),
                       "ghost boxes");
    openfpm_amd::check(ofp_ghost_create(&gh_, openfpm_amd::dtype_of<St>(), nb, blo, bhi, sv, sh), "ghost create");
  }
  vector_dist_gpu(const vector_dist_gpu&) = delete;
  ~vector_dist_gpu() {
    if (gh_) ofp_ghost_destroy(gh_);
    if (pos_) (void)hipFree(pos_);
    if (pos_alt_) (void)hipFree(pos_alt_);
    for (int i = 0; i < n_prop; ++i) {
      if (prp_[i]) (void)hipFree(prp_[i]);
      if (prp_alt_[i]) (void)hipFree(prp_alt_[i]);
    }
  }

  size_t size_local() const { return gm_; }
  size_t size_local_with_ghost() const { return size_; }
  size_t capacity() const { return cap_; }

  // ---------------------------------------------------- host access
  vector_dist_iterator getDomainIterator() const { return vector_dist_iterator{0, gm_}; }
  vector_dist_iterator getDomainAndGhostIterator() const { return vector_dist_iterator{0, size_}; }
  openfpm_amd::pos_ref<St> getPos(vect_dist_key_dx k) { return openfpm_amd::pos_ref<St>{hpos_.data(), cap_, k.key}; }
  openfpm_amd::pos_ref<St> getPos(size_t k) { return openfpm_amd::pos_ref<St>{hpos_.data(), cap_, k}; }
  template <unsigned int id>
  decltype(auto) getProp(vect_dist_key_dx k) {
    return getProp<id>(k.key);
  }
  template <unsigned int id>
  decltype(auto) getProp(size_t k) {
    using P = typename openfpm_amd::nth<int(id), typename props::type>::type;
    using E = typename openfpm_amd::prop_traits<P>::elem;
    return openfpm_amd::prop_access<P>::get(reinterpret_cast<E*>(hprp_[id].data()), cap_, k);
  }

  void hostToDevicePos() {
    for (int k = 0; k < 3; ++k)
      openfpm_amd::hip_check(
          hipMemcpy(pos_ + k * cap_, hpos_.data() + k * cap_, size_ * sizeof(St), hipMemcpyHostToDevice), "h2d pos");
  }
  void deviceToHostPos() {
    for (int k = 0; k < 3; ++k)
      openfpm_amd::hip_check(
          hipMemcpy(hpos_.data() + k * cap_, pos_ + k * cap_, size_ * sizeof(St), hipMemcpyDeviceToHost), "d2h pos");
  }
  template <unsigned int... ids>
  void hostToDeviceProp() {
    (copy_prop(int(ids), true), ...);
  }
  template <unsigned int... ids>
  void deviceToHostProp() {
    (copy_prop(int(ids), false), ...);
  }
  void copy_prop(int i, bool h2d) {
    for (int k = 0; k < prp_planes_[i]; ++k) {
      char* d = static_cast<char*>(prp_[i]) + size_t(k) * cap_ * prp_esz_[i];
      unsigned char* hh = hprp_[i].data() + size_t(k) * cap_ * prp_esz_[i];
      openfpm_amd::hip_check(h2d ? hipMemcpy(d, hh, size_ * prp_esz_[i], hipMemcpyHostToDevice)
                                 : hipMemcpy(hh, d, size_ * prp_esz_[i], hipMemcpyDeviceToHost),
                             "prop copy");
    }
  }

  // ------------------------------------------------------ checkpoints
  // save / load (vector_dist.hpp:2459-2475): the domain particles and every
  // property in the reference's HDF5 checkpoint layout (libofp_io.so).  load
  // replaces the particles by the file's (every block, in file order), with no
  // ghosts.
  void save(const std::string& filename) {
    if (sorted_) throw std::runtime_error("save: call restoreOrder first");
    deviceToHostPos();
    for (int i = 0; i < n_prop; ++i) copy_prop(i, false);
    std::vector<ofp_h5_field> f;
    const uint32_t obj = io_layout(f);
    io_check(ofp_h5_save(filename.c_str(), openfpm_amd::dtype_of<St>(), hpos_.data(), cap_, gm_, f.data(),
                         int(f.size()), obj),
             "save");
  }
  void load(const std::string& filename) {
    if (sorted_) throw std::runtime_error("load: call restoreOrder first");
    std::vector<ofp_h5_field> f;
    uint32_t obj = io_layout(f);
    size_t n = 0;
    io_check(ofp_h5_load_count(filename.c_str(), openfpm_amd::dtype_of<St>(), obj, &n), "load");
    if (n > cap_) {
      alloc(n + n / 2 + 64, 0);
      f.clear();
      obj = io_layout(f);
    }
    io_check(ofp_h5_load(filename.c_str(), openfpm_amd::dtype_of<St>(), hpos_.data(), cap_, cap_, f.data(),
                         int(f.size()), obj, &n),
             "load");
    gm_ = size_ = n;
    hostToDevicePos();
    for (int i = 0; i < n_prop; ++i) copy_prop(i, true);
  }

  // --------------------------------------------------- distribution
  // map(): one rank, periodic wrap of the domain particles; ghosts dropped.
  void map(size_t opt = 0) {
    if (sorted_) throw std::runtime_error("map: call restoreOrder first");
    if (!(opt & RUN_ON_DEVICE)) hostToDevicePos();
    size_ = gm_;
    openfpm_amd::check(ofp_apply_bc(openfpm_amd::dtype_of<St>(), pos_, cap_, size_, dom_lo_, dom_hi_, bc_, stream_),
                       "map");
    if (!(opt & RUN_ON_DEVICE)) deviceToHostPos();
  }

  // ghost_get<prp...>(opt): local periodic ghosts after the ghost marker.
  template <unsigned int... ids>
  void ghost_get(size_t opt = WITH_POSITION) {
    if (sorted_) throw std::runtime_error("ghost_get: call restoreOrder first");
    const bool dev = opt & RUN_ON_DEVICE;
    if (!dev) {
      hostToDevicePos();
      (copy_prop(int(ids), true), ...);
    }
    std::vector<ofp_plane> pl = planes<ids...>(prp_);
    if (opt & SKIP_LABELLING) {
      openfpm_amd::check(ofp_ghost_refresh(gh_, pos_, cap_, gm_, (opt & NO_POSITION) ? 0 : 1, pl.data(),
                                           int(pl.size()), stream_),
                         "ghost refresh");
    } else {
      size_t ng = 0;
      size_ = gm_;
      openfpm_amd::check(ofp_ghost_label(gh_, pos_, cap_, gm_, &ng, stream_), "ghost label");
      if (gm_ + ng > cap_) {
        alloc(gm_ + ng + (gm_ + ng) / 4, gm_);
        pl = planes<ids...>(prp_);
      }
      openfpm_amd::check(ofp_ghost_emit(gh_, pos_, cap_, gm_, gm_, pl.data(), int(pl.size()), stream_),
                         "ghost emit");
      size_ = gm_ + ng;
    }
    if (!dev) {
      openfpm_amd::hip_check(hipStreamSynchronize(stream_), "sync");
      deviceToHostPos();
      (copy_prop(int(ids), false), ...);
    }
  }

  // ghost_put<op, prp...>(opt) (vector_dist.hpp:2304): each local ghost's
  // properties added to its source particle's (op = add_), in ghost order; in
  // storage order, or in the sorted order of the cell list the vector was
  // reordered by (the sums of a pass over a symmetric list).
  template <template <typename, typename> class op, int... ids>
  void ghost_put(size_t opt = 0) {
    static_assert(std::is_same<op<int, int>, add_<int, int>>::value, "ghost_put: this build provides add_");
    const bool dev = opt & RUN_ON_DEVICE;
    if (!dev) (copy_prop(ids, true), ...);
    std::vector<void*> ptr;
    std::vector<int> type;
    (put_planes<ids>(ptr, type), ...);
    if (sorted_)
      openfpm_amd::check(ofp_ghost_put_sorted(gh_, sorted_cl_, ptr.data(), type.data(), int(ptr.size()), stream_),
                         "ghost_put (sorted order)");
    else
      openfpm_amd::check(ofp_ghost_put(gh_, ptr.data(), type.data(), int(ptr.size()), stream_), "ghost_put");
    if (!dev) {
      openfpm_amd::hip_check(hipStreamSynchronize(stream_), "sync");
      (copy_prop(ids, false), ...);
    }
  }

  // ------------------------------------------------------ cell lists
  template <typename CellType = CellList_gpu<dim, St>>
  CellType getCellListGPU(St r_cut, size_t opt = CL_NON_SYMMETRIC, size_t NNIteratorBox = 1, bool no_se3 = false,
                          float ghostEnlargeFactor = 1.013) {
    (void)no_se3;
    St blo[3], bhi[3];
    uint32_t div[3];
    openfpm_amd::check(ofp_cl_param(openfpm_amd::dtype_of<St>(), dom_lo_, dom_hi_, double(r_cut), g_lo_, g_hi_,
                                    double(ghostEnlargeFactor), blo, bhi, div),
                       "cl_param");
    return CellType(blo, bhi, div, opt, int(NNIteratorBox), r_cut);
  }

  template <unsigned int... ids, typename CellList_type>
  void updateCellListGPU(CellList_type& cl, bool no_se3 = false) {
    (void)no_se3;
    const bool rpos = cl.getOpt() & CL_GPU_REORDER_POSITION;
    const bool rprp = cl.getOpt() & CL_GPU_REORDER_PROPERTY;
    if (rpos || rprp) ensure_alt();
    std::vector<ofp_plane> in, out;
    if (rprp) {
      in = planes<ids...>(prp_);
      out = planes<ids...>(prp_alt_);
    }
    openfpm_amd::check(ofp_cl_construct(cl.h, pos_, cap_, size_, gm_, rpos ? pos_alt_ : nullptr, cap_, in.data(),
                                        out.data(), int(in.size()), stream_),
                       "updateCellListGPU");
    if (rpos) std::swap(pos_, pos_alt_);
    if (rprp) (swap_prop(int(ids)), ...);
    if (rpos || rprp) {
      sorted_ = true;
      sorted_cl_ = cl.h;
    }
    if constexpr (CellList_type::is_sparse) openfpm_amd::check(ofp_cl_build_sparse(cl.h, stream_), "sparse list");
  }
  void swap_prop(int i) { std::swap(prp_[i], prp_alt_[i]); }

  template <unsigned int... ids, typename CellList_type>
  void restoreOrder(CellList_type& cl) {
    ensure_alt();
    if (cl.getOpt() & CL_GPU_REORDER_POSITION) {
      ofp_plane a[3], b[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = ofp_plane{pos_ + k * cap_, uint32_t(sizeof(St)), 0};
        b[k] = ofp_plane{pos_alt_ + k * cap_, uint32_t(sizeof(St)), 0};
      }
      openfpm_amd::check(ofp_cl_restore(cl.h, a, b, 3, stream_), "restoreOrder pos");
      std::swap(pos_, pos_alt_);
    }
    std::vector<ofp_plane> in = planes<ids...>(prp_), out = planes<ids...>(prp_alt_);
    if (!in.empty()) {
      openfpm_amd::check(ofp_cl_restore(cl.h, in.data(), out.data(), int(in.size()), stream_), "restoreOrder prop");
      (swap_prop(int(ids)), ...);
    }
    sorted_ = false;
    sorted_cl_ = nullptr;
  }

  // getVerlet on the GPU: rows follow the internal cell list's domain order;
  // the vector is left in that order, positions and the properties ids...
  // (restoreOrder<ids...>(vl.cl) undoes it).
  template <unsigned int... ids>
  VerletList_gpu<dim, St> getVerletGPU(St r_cut, int sub = 1) {
    VerletList_gpu<dim, St> vl;
    const size_t opt = CL_NON_SYMMETRIC | CL_GPU_REORDER_POSITION | (sizeof...(ids) ? CL_GPU_REORDER_PROPERTY : 0);
    vl.cl = getCellListGPU(r_cut / St(sub), opt, size_t(sub), false, float(1.013));
    updateCellListGPU<ids...>(vl.cl);
    openfpm_amd::check(ofp_vl_build(vl.h, vl.cl.h, pos_, cap_, double(r_cut), stream_), "getVerletGPU");
    return vl;
  }

  // getVerletSym (vector_dist.hpp:1412-1416) on the GPU: every interaction of
  // getVerletGPU's list once (domain and ghost neighbours), rows in the cell
  // list's domain order; the vector is left in that order (positions and the
  // properties ids..., as getVerletGPU<ids...>).  A pass that adds
  // each pair to both ends, then ghost_put<add_, prp...>, gives every particle
  // the full list's sums.
  template <unsigned int... ids>
  VerletList_gpu<dim, St> getVerletSym(St r_cut, int sub = 1) {
    VerletList_gpu<dim, St> vl = getVerletGPU<ids...>(r_cut, sub);
    std::swap(vl.h, vl.full);
    openfpm_amd::check(ofp_vl_create(&vl.h), "ofp_vl_create");
    openfpm_amd::check(ofp_vl_build_sym(vl.h, vl.full, gh_, stream_), "getVerletSym");
    return vl;
  }

  // ---------------------------------------------------- device views
  vector_dist_ker<dim, St, prop> toKernel() const {
    vector_dist_ker<dim, St, prop> k;
    k.pos = pos_;
    k.stride = cap_;
    k.gm = int(gm_);
    k.n_all = int(size_);
    for (int i = 0; i < n_prop; ++i) k.prp[i] = prp_[i];
    return k;
  }
  ite_gpu getDomainIteratorGPU(size_t n_thr = 1024) const {
    ite_gpu it;
    const size_t n = gm_ ? gm_ : 1;
    it.thr = dim3(unsigned(n_thr), 1, 1);
    it.wthr = dim3(unsigned((n + n_thr - 1) / n_thr), 1, 1);
    return it;
  }
  ofp_ghost* ghost_handle() const { return gh_; }
  hipStream_t stream() const { return stream_; }
};

template <unsigned int dim, typename St, typename prop>
using vector_dist = vector_dist_gpu<dim, St, prop>;

inline void openfpm_init(int*, char***) {}
inline void openfpm_finalize() {}
