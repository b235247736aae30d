// ofp_celllist.hip -- CellList_gpu construction, reorder and restore.
//
// Build (per updateCellListGPU):
//   1. k_cl_bin      cell id per particle; one 64-bit atomic per particle adds
//                    (domain << 32 | 1) or (1) to the cell's packed counter, so
//                    the returned old value gives the particle's provisional
//                    rank among the cell's domain or ghost particles.
//   2. scan (u64)    one exclusive scan of the packed counters yields both the
//                    cell starts (low word) and the domain-id starts (high word).
//   3. k_cl_scatter  provisional slot = start + (domain ? rank : n_dom_c + rank).
//   4. canonical     inside each cell the particles are put in ascending original
//                    index (domain ids < ghost ids, so domain particles come
//                    first): cells with <= kSmallCell particles by a per-thread
//                    rank count, denser cells by an LDS bitonic sort per chunk
//                    and a merge rank.  The final slot writes the permutations,
//                    the domain sort ids, the reordered positions and planes,
//                    and the quantised cell-relative stream.
// The result does not depend on the atomic order.
#include "ofp_internal.hpp"

#include <cmath>
#include <cstring>

struct ofp_cl {
  int dtype = OFP_F32;
  uint32_t opt = 0;
  ofp::CellGeom<float> gf;
  ofp::CellGeom<double> gd;
  double lo[3], sp[3];
  uint32_t div[3], divc[3];
  uint64_t n_cells = 0;
  uint64_t n_part = 0, gm = 0;
  bool built = false;
  uint32_t box_nn = 1;
  // target-item work queue (built on first use after each construct)
  bool items_valid = false;
  size_t items_cap = 0;
  ofp::DevBuf cnt64, scan64, cid, rank, tmp, s2n, n2s, dom, quant, scantmp, big, nbig;
  ofp::DevBuf items, nitems;  // nitems: [0] = item count, [1] = consumer queue head
};

namespace ofp {

constexpr int kSmallCell = 64;
constexpr int kBigChunk = 4096;
constexpr int kBigBlock = 256;
constexpr unsigned kBigGrid = 512;

template <class T>
__global__ __launch_bounds__(256) void k_cl_bin(CellGeom<T> g, const T* __restrict__ pos,
                                                size_t stride, uint32_t n, uint32_t gm,
                                                uint32_t* __restrict__ cid,
                                                uint32_t* __restrict__ rank,
                                                unsigned long long* __restrict__ cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t cx, cy, cz;
  uint32_t c = cell_of(g, pos[i], pos[stride + i], pos[2 * stride + i], &cx, &cy, &cz);
  cid[i] = c;
  const bool dom = i < gm;
  unsigned long long inc = dom ? ((1ull << 32) | 1ull) : 1ull;
  unsigned long long old = atomicAdd(&cnt[c], inc);
  uint32_t dom_old = uint32_t(old >> 32);
  uint32_t tot_old = uint32_t(old);
  rank[i] = dom ? dom_old : tot_old - dom_old;
}

__global__ __launch_bounds__(256) void k_cl_scatter(uint32_t n, uint32_t gm,
                                                    const uint32_t* __restrict__ cid,
                                                    const uint32_t* __restrict__ rank,
                                                    const unsigned long long* __restrict__ cnt,
                                                    const unsigned long long* __restrict__ scan,
                                                    uint32_t* __restrict__ tmp) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cid[i];
  uint32_t start = uint32_t(scan[c]);
  uint32_t ndom = uint32_t(cnt[c] >> 32);
  uint32_t slot = start + (i < gm ? rank[i] : ndom + rank[i]);
  tmp[slot] = i;
}

__global__ __launch_bounds__(256) void k_cl_find_big(uint64_t n_cells,
                                                     const unsigned long long* __restrict__ cnt,
                                                     uint32_t* __restrict__ big,
                                                     uint32_t* __restrict__ nbig) {
  uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  if (uint32_t(cnt[c]) > uint32_t(kSmallCell)) {
    uint32_t k = atomicAdd(nbig, 1u);
    big[k] = uint32_t(c);
  }
}

template <class T>
struct EmitArgs {
  const T* pos;
  size_t stride;
  T* pos_sorted;
  size_t sorted_stride;
  int4* quant;
  double qlo[3], qsp[3], qinv[3];  // cell origin, spacing, 2^29 / spacing
  uint32_t divc[3];
  uint32_t gm;
  uint32_t* s2n;
  uint32_t* n2s;
  uint32_t* dom;
  PlaneArgs pa;
};

template <class T>
__device__ __forceinline__ void cl_emit(const EmitArgs<T>& A, uint32_t o, uint32_t f, uint32_t c,
                                        uint32_t dstart, uint32_t rank) {
  A.s2n[f] = o;
  A.n2s[o] = f;
  if (o < A.gm) A.dom[dstart + rank] = f;
  if (A.pos_sorted || A.quant) {
    T x = A.pos[o], y = A.pos[A.stride + o], z = A.pos[2 * A.stride + o];
    if (A.pos_sorted) {
      A.pos_sorted[f] = x;
      A.pos_sorted[A.sorted_stride + f] = y;
      A.pos_sorted[2 * A.sorted_stride + f] = z;
    }
    if (A.quant) {
      uint32_t cx = c % A.divc[0];
      uint32_t cyz = c / A.divc[0];
      uint32_t cy = cyz % A.divc[1];
      uint32_t cz = cyz / A.divc[1];
      const double xs[3] = {double(x), double(y), double(z)};
      const uint32_t cc[3] = {cx, cy, cz};
      // g = u + c 2^28 (mod 2^32): the wrapped difference of two particles is
      // their exact integer separation whenever that fits 31 bits.
      uint32_t q[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double org = A.qlo[k] + double(int(cc[k]) - 1) * A.qsp[k];
        double u = rint((xs[k] - org) * A.qinv[k]);
        u = u < -536870912.0 ? -536870912.0 : (u > 536870912.0 ? 536870912.0 : u);
        q[k] = uint32_t(int(u)) + (cc[k] << kQuantBits);
      }
      A.quant[f] = make_int4(int(q[0]), int(q[1]), int(q[2]), pack_cell(cx, cy, cz));
    }
  }
  copy_planes(A.pa, o, f);
}

template <class T>
__global__ __launch_bounds__(256) void k_cl_canon_small(uint32_t n, EmitArgs<T> A,
                                                        const uint32_t* __restrict__ cid,
                                                        const unsigned long long* __restrict__ scan,
                                                        const uint32_t* __restrict__ tmp) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t o = tmp[j];
  uint32_t c = cid[o];
  unsigned long long s0 = scan[c];
  uint32_t start = uint32_t(s0);
  uint32_t end = uint32_t(scan[c + 1]);
  if (end - start > uint32_t(kSmallCell)) return;  // dense cell: k_cl_canon_big
  uint32_t r = 0;
  for (uint32_t k = start; k < end; ++k) r += tmp[k] < o ? 1u : 0u;
  cl_emit(A, o, start + r, c, uint32_t(s0 >> 32), r);
}

__device__ __forceinline__ void lds_bitonic_sort(uint32_t* s, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint32_t a = s[i], b = s[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Dense cells: one workgroup per cell (grid-stride over the dense-cell list,
// whose length is read from device memory; every workgroup exits when the list
// is exhausted).
template <class T>
__global__ __launch_bounds__(kBigBlock) void k_cl_canon_big(EmitArgs<T> A,
                                                            const unsigned long long* __restrict__ scan,
                                                            uint32_t* __restrict__ tmp,
                                                            const uint32_t* __restrict__ big,
                                                            const uint32_t* __restrict__ nbig_p) {
  __shared__ uint32_t s[kBigChunk];
  const uint32_t nbig = *nbig_p;
  for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
    const uint32_t c = big[bi];
    const unsigned long long s0 = scan[c];
    const uint32_t start = uint32_t(s0);
    const uint32_t end = uint32_t(scan[c + 1]);
    const uint32_t dstart = uint32_t(s0 >> 32);
    const uint32_t m = end - start;
    const uint32_t nch = (m + kBigChunk - 1) / kBigChunk;
    for (uint32_t ch = 0; ch < nch; ++ch) {
      const uint32_t a = start + ch * kBigChunk;
      const uint32_t len = min(uint32_t(kBigChunk), end - a);
      int P = 1;
      while (uint32_t(P) < len) P <<= 1;
      for (int i = threadIdx.x; i < P; i += blockDim.x)
        s[i] = uint32_t(i) < len ? tmp[a + i] : 0xFFFFFFFFu;
      __syncthreads();
      lds_bitonic_sort(s, P);
      if (nch == 1) {
        for (uint32_t r = threadIdx.x; r < len; r += blockDim.x) cl_emit(A, s[r], start + r, c, dstart, r);
      } else {
        for (uint32_t r = threadIdx.x; r < len; r += blockDim.x) tmp[a + r] = s[r];
      }
      __syncthreads();
    }
    if (nch > 1) {
      __threadfence_block();
      __syncthreads();
      // rank of o in the cell = sum over sorted chunks of #(elements < o)
      for (uint32_t e = start + threadIdx.x; e < end; e += blockDim.x) {
        const uint32_t o = tmp[e];
        uint32_t r = 0;
        for (uint32_t ch = 0; ch < nch; ++ch) {
          const uint32_t a = start + ch * kBigChunk;
          uint32_t lo = 0, hi = min(uint32_t(kBigChunk), end - a);
          while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (tmp[a + mid] < o)
              lo = mid + 1;
            else
              hi = mid;
          }
          r += lo;
        }
        cl_emit(A, o, start + r, c, dstart, r);
      }
      __syncthreads();
    }
  }
}

// out[j] = in[s2n[j]] (gather into sorted order)
__global__ __launch_bounds__(256) void k_cl_reorder(uint32_t n, const uint32_t* __restrict__ s2n,
                                                    PlaneArgs pa) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  copy_planes(pa, s2n[j], j);
}

// out[s2n[j]] = in[j] (scatter back to the original order)
__global__ __launch_bounds__(256) void k_cl_restore(uint32_t n, const uint32_t* __restrict__ s2n,
                                                    PlaneArgs pa) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  copy_planes(pa, j, s2n[j]);
}

template <class T>
static int cl_construct(ofp_cl* h, const CellGeom<T>& g, const void* pos_v, size_t stride,
                        size_t n_tot, size_t gm, void* pos_sorted_v, size_t sorted_stride,
                        const PlaneArgs& pa, hipStream_t s) {
  const T* pos = static_cast<const T*>(pos_v);
  const uint64_t C = h->n_cells;
  const size_t n = n_tot;
  OFP_TRY(h->cnt64.ensure(C * 8));
  OFP_TRY(h->scan64.ensure((C + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(C)));
  OFP_TRY(h->cid.ensure((n + 1) * 4));
  OFP_TRY(h->rank.ensure((n + 1) * 4));
  OFP_TRY(h->tmp.ensure((n + 1) * 4));
  OFP_TRY(h->s2n.ensure((n + 1) * 4));
  OFP_TRY(h->n2s.ensure((n + 1) * 4));
  OFP_TRY(h->dom.ensure((gm + 1) * 4));
  OFP_TRY(h->big.ensure(C * 4));
  OFP_TRY(h->nbig.ensure(4));
  const bool want_quant = (h->opt & OFP_CL_QUANT) != 0;
  if (want_quant) OFP_TRY(h->quant.ensure((n + 1) * 16));

  auto* cnt = h->cnt64.as<unsigned long long>();
  auto* scan = h->scan64.as<unsigned long long>();
  OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(cnt), 0, C, s));
  OFP_TRY(fill_u32(h->nbig.as<uint32_t>(), 0, 1, s));
  if (n > 0) {
    k_cl_bin<T><<<grid_for(n, 256), 256, 0, s>>>(g, pos, stride, uint32_t(n), uint32_t(gm),
                                                h->cid.as<uint32_t>(), h->rank.as<uint32_t>(), cnt);
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u64(reinterpret_cast<const uint64_t*>(cnt), reinterpret_cast<uint64_t*>(scan),
                        C, h->scantmp.p, s));
  if (n > 0) {
    k_cl_scatter<<<grid_for(n, 256), 256, 0, s>>>(uint32_t(n), uint32_t(gm), h->cid.as<uint32_t>(),
                                                  h->rank.as<uint32_t>(), cnt, scan,
                                                  h->tmp.as<uint32_t>());
    OFP_LAUNCHED();
    k_cl_find_big<<<grid_for(C, 256), 256, 0, s>>>(C, cnt, h->big.as<uint32_t>(),
                                                   h->nbig.as<uint32_t>());
    OFP_LAUNCHED();
    EmitArgs<T> A;
    std::memset(&A, 0, sizeof(A));
    A.pos = pos;
    A.stride = stride;
    A.pos_sorted = static_cast<T*>(pos_sorted_v);
    A.sorted_stride = sorted_stride;
    A.quant = want_quant ? h->quant.as<int4>() : nullptr;
    for (int k = 0; k < 3; ++k) {
      A.qlo[k] = double(g.lo[k]);
      A.qsp[k] = double(g.sp[k]);
      A.qinv[k] = double(kQuantOne) / double(g.sp[k]);
      A.divc[k] = g.divc[k];
    }
    A.gm = uint32_t(gm);
    A.s2n = h->s2n.as<uint32_t>();
    A.n2s = h->n2s.as<uint32_t>();
    A.dom = h->dom.as<uint32_t>();
    A.pa = pa;
    k_cl_canon_small<T><<<grid_for(n, 256), 256, 0, s>>>(uint32_t(n), A, h->cid.as<uint32_t>(), scan,
                                                        h->tmp.as<uint32_t>());
    OFP_LAUNCHED();
    k_cl_canon_big<T><<<kBigGrid, kBigBlock, 0, s>>>(A, scan, h->tmp.as<uint32_t>(),
                                                     h->big.as<uint32_t>(), h->nbig.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_part = n;
  h->gm = gm;
  h->built = true;
  h->items_valid = false;
  return OFP_OK;
}

// Target items for the Verlet build: one thread per x-row of cells groups
// consecutive cells greedily while they hold <= 64 domain particles in total
// (spanning at most kItemMaxCells cells); a cell with more than 64 is split
// into chunks of 64.  Every domain particle is in exactly one item.
constexpr uint32_t kItemTargets = 64;
constexpr uint32_t kItemMaxCells = 8;

__global__ __launch_bounds__(256) void k_cl_items(const unsigned long long* __restrict__ cs,
                                                  uint32_t divc0, uint32_t nrows,
                                                  TargetItem* __restrict__ items,
                                                  uint32_t* __restrict__ nitems) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const uint32_t base = r * divc0;
  auto emit = [&](uint32_t row0, uint32_t n, uint32_t cell0, uint32_t ncx) {
    uint32_t k = atomicAdd(nitems, 1u);
    items[k] = TargetItem{row0, n, cell0, ncx};
  };
  uint32_t g_first = 0, g_last = 0, g_n = 0, g_row0 = 0;
  unsigned long long prev = cs[base];
  for (uint32_t cx = 0; cx < divc0; ++cx) {
    unsigned long long next = cs[base + cx + 1];
    uint32_t d0 = uint32_t(prev >> 32);
    uint32_t nd = uint32_t(next >> 32) - d0;
    prev = next;
    if (nd == 0) continue;
    if (nd > kItemTargets) {
      if (g_n) emit(g_row0, g_n, base + g_first, g_last - g_first + 1);
      g_n = 0;
      for (uint32_t k = 0; k < nd; k += kItemTargets)
        emit(d0 + k, min(kItemTargets, nd - k), base + cx, 1);
      continue;
    }
    if (g_n && (g_n + nd > kItemTargets || cx - g_first >= kItemMaxCells)) {
      emit(g_row0, g_n, base + g_first, g_last - g_first + 1);
      g_n = 0;
    }
    if (g_n == 0) {
      g_first = cx;
      g_row0 = d0;
    }
    g_last = cx;
    g_n += nd;
  }
  if (g_n) emit(g_row0, g_n, base + g_first, g_last - g_first + 1);
}

}  // namespace ofp

using namespace ofp;

extern "C" {

int ofp_cl_create(ofp_cl** out, int dtype, const void* box_lo, const void* box_hi,
                  const uint32_t* div, uint32_t opt) {
  OFP_CHECK(out && box_lo && box_hi && div, OFP_E_ARG, "ofp_cl_create: null argument");
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  uint64_t C = 1;
  for (int k = 0; k < 3; ++k) {
    OFP_CHECK(div[k] >= 1, OFP_E_ARG, "div[%d] must be >= 1", k);
    C *= uint64_t(div[k]) + 2;
  }
  OFP_CHECK(C < (1ull << 31), OFP_E_LIMIT, "too many cells (%llu)", (unsigned long long)C);
  if (opt & OFP_CL_QUANT)
    OFP_CHECK(div[0] + 2 <= 2048 && div[1] + 2 <= 2048 && div[2] + 2 <= 1024, OFP_E_LIMIT,
              "OFP_CL_QUANT packs cell coordinates in 11/11/10 bits");
  ofp_cl* h = new ofp_cl();
  h->dtype = dtype;
  h->opt = opt;
  h->n_cells = C;
  for (int k = 0; k < 3; ++k) {
    h->div[k] = div[k];
    h->divc[k] = div[k] + 2;
  }
  if (dtype == OFP_F32) {
    const float* lo = (const float*)box_lo;
    const float* hi = (const float*)box_hi;
    for (int k = 0; k < 3; ++k) {
      h->gf.lo[k] = lo[k];
      h->gf.sp[k] = (hi[k] - lo[k]) / float(div[k]);
      h->gf.divc[k] = h->divc[k];
      h->lo[k] = h->gf.lo[k];
      h->sp[k] = h->gf.sp[k];
    }
  } else {
    const double* lo = (const double*)box_lo;
    const double* hi = (const double*)box_hi;
    for (int k = 0; k < 3; ++k) {
      h->gd.lo[k] = lo[k];
      h->gd.sp[k] = (hi[k] - lo[k]) / double(div[k]);
      h->gd.divc[k] = h->divc[k];
      h->lo[k] = h->gd.lo[k];
      h->sp[k] = h->gd.sp[k];
    }
  }
  *out = h;
  return OFP_OK;
}

int ofp_cl_destroy(ofp_cl* h) {
  if (!h) return OFP_OK;
  DevBuf* bufs[] = {&h->cnt64, &h->scan64, &h->cid, &h->rank, &h->tmp,     &h->s2n,
                    &h->n2s,   &h->dom,    &h->quant, &h->scantmp, &h->big, &h->nbig,
                    &h->items, &h->nitems};
  for (DevBuf* b : bufs) b->release();
  delete h;
  return OFP_OK;
}

int ofp_cl_construct(ofp_cl* h, const void* pos, size_t stride, size_t n_tot, size_t gm,
                     void* pos_sorted, size_t sorted_stride, const ofp_plane* planes_in,
                     const ofp_plane* planes_out, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_cl_construct: null handle");
  OFP_CHECK(n_tot == 0 || pos, OFP_E_ARG, "ofp_cl_construct: null positions");
  OFP_CHECK(gm <= n_tot, OFP_E_ARG, "ghost_marker %zu > n_tot %zu", gm, n_tot);
  OFP_CHECK(stride >= n_tot, OFP_E_ARG, "stride %zu < n_tot %zu", stride, n_tot);
  OFP_CHECK(!pos_sorted || sorted_stride >= n_tot, OFP_E_ARG, "sorted_stride < n_tot");
  OFP_CHECK(n_tot < 0xFFFFFFF0ull, OFP_E_LIMIT, "n_tot must fit 32 bits");
  OFP_CHECK(n_planes == 0 || planes_out, OFP_E_ARG, "reordered planes need outputs");
  PlaneArgs pa;
  OFP_TRY(make_plane_args(planes_in, planes_out, n_planes, &pa));
  if (h->dtype == OFP_F32)
    return cl_construct(h, h->gf, pos, stride, n_tot, gm, pos_sorted, sorted_stride, pa,
                        as_stream(stream));
  return cl_construct(h, h->gd, pos, stride, n_tot, gm, pos_sorted, sorted_stride, pa,
                      as_stream(stream));
}

int ofp_cl_set_box_nn(ofp_cl* h, uint32_t box_nn) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_cl_set_box_nn: null handle");
  OFP_CHECK(box_nn >= 1 && box_nn <= uint32_t(kMaxBoxNN), OFP_E_LIMIT,
            "box_nn must be in [1, %d] (got %u)", kMaxBoxNN, box_nn);
  h->box_nn = box_nn;
  return OFP_OK;
}

int ofp_cl_dev_(const ofp_cl* h, ClDev* d, ofp_cl_view* v) {
  OFP_TRY(ofp_cl_get_view(h, v));
  d->cs = h->scan64.as<unsigned long long>();
  d->dom = h->dom.as<uint32_t>();
  d->quant = (h->opt & OFP_CL_QUANT) ? h->quant.as<int4>() : nullptr;
  for (int k = 0; k < 3; ++k) d->divc[k] = h->divc[k];
  d->n_dom = uint32_t(h->gm);
  d->box_nn = h->box_nn;
  d->n_cells = h->n_cells;
  return OFP_OK;
}

int ofp_cl_items_(const ofp_cl* hc, const TargetItem** items, const uint32_t** n_items,
                  uint32_t** counter, size_t* cap, hipStream_t s) {
  ofp_cl* h = const_cast<ofp_cl*>(hc);
  OFP_CHECK(h && h->built, OFP_E_STATE, "cell list not constructed");
  if (!h->items_valid) {
    size_t need = size_t(h->gm) + 1;
    OFP_TRY(h->items.ensure(need * sizeof(TargetItem)));
    OFP_TRY(h->nitems.ensure(2 * sizeof(uint32_t)));
    OFP_TRY(fill_u32(h->nitems.as<uint32_t>(), 0, 2, s));
    uint32_t nrows = h->divc[1] * h->divc[2];
    k_cl_items<<<grid_for(nrows, 256), 256, 0, s>>>(h->scan64.as<unsigned long long>(), h->divc[0],
                                                    nrows, h->items.as<TargetItem>(),
                                                    h->nitems.as<uint32_t>());
    OFP_LAUNCHED();
    h->items_cap = need;
    h->items_valid = true;
  }
  *items = h->items.as<TargetItem>();
  *n_items = h->nitems.as<uint32_t>();
  *counter = h->nitems.as<uint32_t>() + 1;
  *cap = h->items_cap;
  return OFP_OK;
}

int ofp_cl_get_view(const ofp_cl* h, ofp_cl_view* v) {
  OFP_CHECK(h && v, OFP_E_ARG, "ofp_cl_get_view: null argument");
  OFP_CHECK(h->built, OFP_E_STATE, "cell list not constructed");
  std::memset(v, 0, sizeof(*v));
  v->cell_start = h->scan64.as<uint64_t>();
  v->sort_to_nonsort = h->s2n.as<uint32_t>();
  v->nonsort_to_sort = h->n2s.as<uint32_t>();
  v->dom_sort_ids = h->dom.as<uint32_t>();
  v->quant = (h->opt & OFP_CL_QUANT) ? h->quant.as<int32_t>() : nullptr;
  for (int k = 0; k < 3; ++k) {
    v->lo[k] = h->lo[k];
    v->spacing[k] = h->sp[k];
    v->div[k] = h->div[k];
    v->divc[k] = h->divc[k];
  }
  v->n_cells = h->n_cells;
  v->n_part = h->n_part;
  v->ghost_marker = h->gm;
  v->dtype = h->dtype;
  return OFP_OK;
}

int ofp_cl_restore(const ofp_cl* h, const ofp_plane* sorted_in, const ofp_plane* orig_out,
                   int n_planes, void* stream) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "ofp_cl_restore: cell list not constructed");
  OFP_CHECK(n_planes == 0 || orig_out, OFP_E_ARG, "ofp_cl_restore: null outputs");
  PlaneArgs pa;
  OFP_TRY(make_plane_args(sorted_in, orig_out, n_planes, &pa));
  if (h->n_part == 0 || n_planes == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  k_cl_restore<<<grid_for(h->n_part, 256), 256, 0, s>>>(uint32_t(h->n_part), h->s2n.as<uint32_t>(), pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_cl_reorder(const ofp_cl* h, const ofp_plane* orig_in, const ofp_plane* sorted_out,
                   int n_planes, void* stream) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "ofp_cl_reorder: cell list not constructed");
  OFP_CHECK(n_planes == 0 || sorted_out, OFP_E_ARG, "ofp_cl_reorder: null outputs");
  PlaneArgs pa;
  OFP_TRY(make_plane_args(orig_in, sorted_out, n_planes, &pa));
  if (h->n_part == 0 || n_planes == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  k_cl_reorder<<<grid_for(h->n_part, 256), 256, 0, s>>>(uint32_t(h->n_part), h->s2n.as<uint32_t>(), pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

}  // extern "C"
