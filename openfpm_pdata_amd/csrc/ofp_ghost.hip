// ofp_ghost.hip -- one-rank periodic ghosts: label (count + scan), emit,
// SKIP_LABELLING refresh, sorted-space property sync, and the map() wrap.
//
// Order of the emitted ghosts is the reference GPU order: particle-major, and
// inside a particle ascending shift id (the boxes are sorted by shift id), with
// a box skipped when its shift id equals the last one counted for the particle
// (Vector/cuda/vector_dist_cuda_funcs.cuh:157-233).
#include "ofp_internal.hpp"

#include <cstring>
#include <vector>

struct ofp_ghost {
  int dtype = OFP_F32;
  ofp::GhostBoxes<float> bf;
  ofp::GhostBoxes<double> bd;
  ofp::DevBuf cnt, starts, tmp, opid, osid;
  size_t n_dom = 0;    // ghost marker at the last label
  size_t n_ghost = 0;  // ghosts of the last label
  bool labelled = false;
  bool emitted = false;
};

namespace ofp {

template <class T>
__global__ __launch_bounds__(256) void k_ghost_count(GhostBoxes<T> B, const T* __restrict__ pos,
                                                     size_t stride, uint32_t n_dom,
                                                     uint32_t* __restrict__ cnt) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_dom) return;
  T x = pos[p], y = pos[stride + p], z = pos[2 * stride + p];
  uint32_t old = 0xFFFFFFFFu, n = 0;
  if (!no_box(B, x, y, z)) {
    for (int b = 0; b < B.nb; ++b) {
      if (inside_np(B, b, x, y, z) && B.sv[b] != old) {
        old = B.sv[b];
        ++n;
      }
    }
  }
  cnt[p] = n;
}

template <class T>
__global__ __launch_bounds__(256) void k_ghost_emit(GhostBoxes<T> B, T* __restrict__ pos,
                                                    size_t stride, uint32_t n_dom,
                                                    size_t old_size,
                                                    const uint32_t* __restrict__ starts,
                                                    uint32_t* __restrict__ opid,
                                                    uint32_t* __restrict__ osid, PlaneArgs pa) {
#pragma clang fp contract(off)
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_dom) return;
  T x = pos[p], y = pos[stride + p], z = pos[2 * stride + p];
  if (no_box(B, x, y, z)) return;
  uint32_t base_o = starts[p];
  uint32_t old = 0xFFFFFFFFu, n = 0;
  for (int b = 0; b < B.nb; ++b) {
    if (inside_np(B, b, x, y, z) && B.sv[b] != old) {
      uint32_t s = B.sv[b];
      size_t g = old_size + base_o + n;
      pos[g] = x - B.shift[s][0];
      pos[stride + g] = y - B.shift[s][1];
      pos[2 * stride + g] = z - B.shift[s][2];
      opid[base_o + n] = p;
      osid[base_o + n] = s;
      copy_planes(pa, p, g);
      old = s;
      ++n;
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_ghost_refresh(GhostBoxes<T> B, T* __restrict__ pos,
                                                       size_t stride, size_t n_ghost,
                                                       size_t old_size, int with_pos,
                                                       const uint32_t* __restrict__ opid,
                                                       const uint32_t* __restrict__ osid,
                                                       PlaneArgs pa) {
#pragma clang fp contract(off)
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_ghost) return;
  uint32_t p = opid[i];
  size_t g = old_size + i;
  if (with_pos) {
    uint32_t s = osid[i];
    pos[g] = pos[p] - B.shift[s][0];
    pos[stride + g] = pos[stride + p] - B.shift[s][1];
    pos[2 * stride + g] = pos[2 * stride + p] - B.shift[s][2];
  }
  copy_planes(pa, p, g);
}

// ghost_get of properties in the cell list's sorted space: the ghost with original
// index gm + i takes the values of its source particle opid[i].
__global__ __launch_bounds__(256) void k_ghost_sync_sorted(size_t n_ghost, size_t gm,
                                                           const uint32_t* __restrict__ opid,
                                                           const uint32_t* __restrict__ n2s,
                                                           PlaneArgs pa) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_ghost) return;
  uint32_t dst = n2s[gm + i];
  uint32_t src = n2s[opid[i]];
  copy_planes(pa, src, dst);
}

// ghost_put<add_>: each run of ghosts of one source particle (the o_part list is
// source-major, as the emit writes it) is added to the source in ghost order by
// the thread of the run's first ghost.  n2s = the cell list's original-to-sorted
// map while the vector is in sorted order, null in storage order (ghost k at
// gm + k, its source at opid[k]).
struct PutArgs {
  uint64_t p[kMaxPlanes];
  int type[kMaxPlanes];  // OFP_PUT_*
  int n;
};

template <class V>
__device__ __forceinline__ void put_run(uint64_t base, uint32_t dst, const uint32_t* __restrict__ opid,
                                        const uint32_t* __restrict__ n2s, size_t gm, size_t k, size_t n_ghost) {
  V* P = reinterpret_cast<V*>(base);
  V v = P[dst];
  const uint32_t src = opid[k];
  for (size_t m = k; m < n_ghost && opid[m] == src; ++m) v = v + P[n2s ? size_t(n2s[gm + m]) : gm + m];
  P[dst] = v;
}

__global__ __launch_bounds__(256) void k_ghost_put_sorted(size_t n_ghost, size_t gm,
                                                          const uint32_t* __restrict__ opid,
                                                          const uint32_t* __restrict__ n2s, PutArgs pa) {
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_ghost) return;
  if (k > 0 && opid[k - 1] == opid[k]) return;  // not the first ghost of its source
  const uint32_t dst = n2s ? n2s[opid[k]] : opid[k];
  for (int q = 0; q < pa.n; ++q) {
    switch (pa.type[q]) {
      case OFP_PUT_F32: put_run<float>(pa.p[q], dst, opid, n2s, gm, k, n_ghost); break;
      case OFP_PUT_F64: put_run<double>(pa.p[q], dst, opid, n2s, gm, k, n_ghost); break;
      case OFP_PUT_U32: put_run<uint32_t>(pa.p[q], dst, opid, n2s, gm, k, n_ghost); break;
      default: put_run<long long>(pa.p[q], dst, opid, n2s, gm, k, n_ghost); break;
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_apply_bc(T* __restrict__ pos, size_t stride, size_t n,
                                                  T lo0, T lo1, T lo2, T hi0, T hi1, T hi2,
                                                  int bc0, int bc1, int bc2) {
#pragma clang fp contract(off)
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T lo[3] = {lo0, lo1, lo2};
  const T hi[3] = {hi0, hi1, hi2};
  const int bc[3] = {bc0, bc1, bc2};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (!bc[k]) continue;
    T L = hi[k] - lo[k];
    T x = pos[k * stride + i];
    if (x >= hi[k])
      x = x - L;
    else if (x < lo[k])
      x = x + L;
    pos[k * stride + i] = x;
  }
}

template <class T>
static int ghost_label(ofp_ghost* h, const GhostBoxes<T>& B, const void* pos, size_t stride,
                       size_t n_dom, size_t* n_ghost, hipStream_t s) {
  OFP_TRY(h->cnt.ensure((n_dom + 1) * sizeof(uint32_t)));
  OFP_TRY(h->starts.ensure((n_dom + 1) * sizeof(uint32_t)));
  OFP_TRY(h->tmp.ensure(scan_tmp_bytes(n_dom)));
  if (n_dom > 0) {
    k_ghost_count<T><<<grid_for(n_dom, 256), 256, 0, s>>>(B, static_cast<const T*>(pos), stride,
                                                         uint32_t(n_dom), h->cnt.as<uint32_t>());
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32(h->cnt.as<uint32_t>(), h->starts.as<uint32_t>(), n_dom, h->tmp.p, s));
  uint32_t tot = 0;
  OFP_HIP(hipMemcpyAsync(&tot, h->starts.as<uint32_t>() + n_dom, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  h->n_dom = n_dom;
  h->n_ghost = tot;
  h->labelled = true;
  h->emitted = false;
  *n_ghost = tot;
  return OFP_OK;
}

template <class T>
static int ghost_emit(ofp_ghost* h, const GhostBoxes<T>& B, void* pos, size_t stride,
                      size_t n_dom, size_t old_size, const PlaneArgs& pa, hipStream_t s) {
  OFP_TRY(h->opid.ensure((h->n_ghost + 1) * sizeof(uint32_t)));
  OFP_TRY(h->osid.ensure((h->n_ghost + 1) * sizeof(uint32_t)));
  if (n_dom > 0) {
    k_ghost_emit<T><<<grid_for(n_dom, 256), 256, 0, s>>>(
        B, static_cast<T*>(pos), stride, uint32_t(n_dom), old_size, h->starts.as<uint32_t>(),
        h->opid.as<uint32_t>(), h->osid.as<uint32_t>(), pa);
    OFP_LAUNCHED();
  }
  h->emitted = true;
  return OFP_OK;
}

}  // namespace ofp

using namespace ofp;

extern "C" {

int ofp_ghost_create(ofp_ghost** out, int dtype, int nb, const void* blo, const void* bhi,
                     const uint32_t* sv, const void* shifts27) {
  OFP_CHECK(out, OFP_E_ARG, "ofp_ghost_create: null handle pointer");
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  OFP_CHECK(nb >= 0 && nb <= kMaxBoxes, OFP_E_LIMIT, "at most %d ghost boxes (got %d)",
            kMaxBoxes, nb);
  OFP_CHECK(nb == 0 || (blo && bhi && sv), OFP_E_ARG, "ofp_ghost_create: null box arrays");
  OFP_CHECK(shifts27, OFP_E_ARG, "ofp_ghost_create: null shift vectors");
  for (int b = 0; b < nb; ++b)
    OFP_CHECK(sv[b] < 27, OFP_E_ARG, "shift id %u out of range", sv[b]);
  ofp_ghost* h = new ofp_ghost();
  h->dtype = dtype;
  if (dtype == OFP_F32)
    fill_boxes(h->bf, nb, blo, bhi, sv, shifts27);
  else
    fill_boxes(h->bd, nb, blo, bhi, sv, shifts27);
  *out = h;
  return OFP_OK;
}

int ofp_ghost_destroy(ofp_ghost* h) {
  if (!h) return OFP_OK;
  h->cnt.release();
  h->starts.release();
  h->tmp.release();
  h->opid.release();
  h->osid.release();
  delete h;
  return OFP_OK;
}

int ofp_ghost_label(ofp_ghost* h, const void* pos, size_t stride, size_t n_dom,
                    size_t* n_ghost, void* stream) {
  OFP_CHECK(h && n_ghost, OFP_E_ARG, "ofp_ghost_label: null argument");
  OFP_CHECK(n_dom == 0 || pos, OFP_E_ARG, "ofp_ghost_label: null positions");
  OFP_CHECK(stride >= n_dom, OFP_E_ARG, "stride %zu < n_dom %zu", stride, n_dom);
  OFP_CHECK(n_dom < 0xFFFFFFF0ull, OFP_E_LIMIT, "n_dom must fit 32 bits");
  if (h->dtype == OFP_F32) return ghost_label(h, h->bf, pos, stride, n_dom, n_ghost, as_stream(stream));
  return ghost_label(h, h->bd, pos, stride, n_dom, n_ghost, as_stream(stream));
}

int ofp_ghost_emit(ofp_ghost* h, void* pos, size_t stride, size_t n_dom, size_t old_size,
                   const ofp_plane* planes, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_ghost_emit: null handle");
  OFP_CHECK(h->labelled && h->n_dom == n_dom, OFP_E_STATE,
            "ofp_ghost_emit: call ofp_ghost_label with the same n_dom first");
  OFP_CHECK(n_dom == 0 || pos, OFP_E_ARG, "ofp_ghost_emit: null positions");
  OFP_CHECK(old_size >= n_dom, OFP_E_ARG, "old_size < n_dom");
  OFP_CHECK(stride >= old_size + h->n_ghost, OFP_E_CAPACITY,
            "positions need room for %zu particles (stride %zu)", old_size + h->n_ghost, stride);
  PlaneArgs pa;
  OFP_TRY(make_plane_args(planes, nullptr, n_planes, &pa));
  if (h->dtype == OFP_F32)
    return ghost_emit(h, h->bf, pos, stride, n_dom, old_size, pa, as_stream(stream));
  return ghost_emit(h, h->bd, pos, stride, n_dom, old_size, pa, as_stream(stream));
}

int ofp_ghost_refresh(ofp_ghost* h, void* pos, size_t stride, size_t old_size, int with_pos,
                      const ofp_plane* planes, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_ghost_refresh: null handle");
  OFP_CHECK(h->emitted, OFP_E_STATE, "ofp_ghost_refresh: no cached ghost list (emit first)");
  OFP_CHECK(!with_pos || pos, OFP_E_ARG, "ofp_ghost_refresh: null positions");
  OFP_CHECK(!with_pos || stride >= old_size + h->n_ghost, OFP_E_CAPACITY,
            "positions need room for %zu particles", old_size + h->n_ghost);
  PlaneArgs pa;
  OFP_TRY(make_plane_args(planes, nullptr, n_planes, &pa));
  if (h->n_ghost == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  if (h->dtype == OFP_F32)
    k_ghost_refresh<float><<<grid_for(h->n_ghost, 256), 256, 0, s>>>(
        h->bf, static_cast<float*>(pos), stride, h->n_ghost, old_size, with_pos,
        h->opid.as<uint32_t>(), h->osid.as<uint32_t>(), pa);
  else
    k_ghost_refresh<double><<<grid_for(h->n_ghost, 256), 256, 0, s>>>(
        h->bd, static_cast<double*>(pos), stride, h->n_ghost, old_size, with_pos,
        h->opid.as<uint32_t>(), h->osid.as<uint32_t>(), pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_ghost_opart(const ofp_ghost* h, const uint32_t** pid, const uint32_t** sid,
                    size_t* n) {
  OFP_CHECK(h && pid && sid && n, OFP_E_ARG, "ofp_ghost_opart: null argument");
  OFP_CHECK(h->emitted, OFP_E_STATE, "ofp_ghost_opart: no ghost list yet");
  *pid = h->opid.as<uint32_t>();
  *sid = h->osid.as<uint32_t>();
  *n = h->n_ghost;
  return OFP_OK;
}

int ofp_ghost_sync_sorted(const ofp_ghost* h, const ofp_cl* cl, const ofp_plane* planes,
                          int n_planes, void* stream) {
  OFP_CHECK(h && cl, OFP_E_ARG, "ofp_ghost_sync_sorted: null handle");
  OFP_CHECK(h->emitted, OFP_E_STATE, "ofp_ghost_sync_sorted: no ghost list yet");
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(cl, &v));
  OFP_CHECK(v.ghost_marker == h->n_dom && v.n_part == h->n_dom + h->n_ghost, OFP_E_STATE,
            "cell list was not built on this ghost layout (gm %llu vs %zu, n %llu vs %zu)",
            (unsigned long long)v.ghost_marker, h->n_dom, (unsigned long long)v.n_part,
            h->n_dom + h->n_ghost);
  PlaneArgs pa;
  OFP_TRY(make_plane_args(planes, nullptr, n_planes, &pa));
  if (h->n_ghost == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  k_ghost_sync_sorted<<<grid_for(h->n_ghost, 256), 256, 0, s>>>(
      h->n_ghost, h->n_dom, h->opid.as<uint32_t>(), v.nonsort_to_sort, pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

static int ghost_put(const ofp_ghost* h, const uint32_t* n2s, void* const* planes, const int* types, int n_planes,
                     void* stream) {
  OFP_CHECK(n_planes >= 0 && n_planes <= kMaxPlanes, OFP_E_LIMIT, "at most %d planes per call (got %d)",
            kMaxPlanes, n_planes);
  OFP_CHECK(n_planes == 0 || (planes && types), OFP_E_ARG, "ghost_put: null plane list");
  PutArgs pa{};
  pa.n = n_planes;
  for (int q = 0; q < n_planes; ++q) {
    OFP_CHECK(planes[q], OFP_E_ARG, "plane %d: null pointer", q);
    OFP_CHECK(types[q] >= OFP_PUT_F32 && types[q] <= OFP_PUT_I64, OFP_E_ARG, "plane %d: bad value type %d", q,
              types[q]);
    pa.p[q] = reinterpret_cast<uint64_t>(planes[q]);
    pa.type[q] = types[q];
  }
  if (h->n_ghost == 0 || n_planes == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  k_ghost_put_sorted<<<grid_for(h->n_ghost, 256), 256, 0, s>>>(h->n_ghost, h->n_dom, h->opid.as<uint32_t>(), n2s,
                                                                pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_ghost_put_sorted(const ofp_ghost* h, const ofp_cl* cl, void* const* planes, const int* types,
                         int n_planes, void* stream) {
  OFP_CHECK(h && cl, OFP_E_ARG, "ofp_ghost_put_sorted: null handle");
  OFP_CHECK(h->emitted, OFP_E_STATE, "ofp_ghost_put_sorted: no ghost list yet");
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(cl, &v));
  OFP_CHECK(v.ghost_marker == h->n_dom && v.n_part == h->n_dom + h->n_ghost, OFP_E_STATE,
            "cell list was not built on this ghost layout (gm %llu vs %zu, n %llu vs %zu)",
            (unsigned long long)v.ghost_marker, h->n_dom, (unsigned long long)v.n_part,
            h->n_dom + h->n_ghost);
  return ghost_put(h, v.nonsort_to_sort, planes, types, n_planes, stream);
}

int ofp_ghost_put(const ofp_ghost* h, void* const* planes, const int* types, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_ghost_put: null handle");
  OFP_CHECK(h->emitted, OFP_E_STATE, "ofp_ghost_put: no ghost list yet");
  return ghost_put(h, nullptr, planes, types, n_planes, stream);
}

int ofp_apply_bc(int dtype, void* pos, size_t stride, size_t n, const void* lo, const void* hi,
                 const int* bc, void* stream) {
  OFP_CHECK(lo && hi && bc, OFP_E_ARG, "ofp_apply_bc: null argument");
  OFP_CHECK(n == 0 || pos, OFP_E_ARG, "ofp_apply_bc: null positions");
  OFP_CHECK(stride >= n, OFP_E_ARG, "stride < n");
  if (n == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == OFP_F32) {
    const float* l = (const float*)lo;
    const float* u = (const float*)hi;
    k_apply_bc<float><<<grid_for(n, 256), 256, 0, s>>>((float*)pos, stride, n, l[0], l[1], l[2],
                                                       u[0], u[1], u[2], bc[0], bc[1], bc[2]);
  } else if (dtype == OFP_F64) {
    const double* l = (const double*)lo;
    const double* u = (const double*)hi;
    k_apply_bc<double><<<grid_for(n, 256), 256, 0, s>>>((double*)pos, stride, n, l[0], l[1],
                                                        l[2], u[0], u[1], u[2], bc[0], bc[1],
                                                        bc[2]);
  } else {
    return set_error(OFP_E_ARG, "bad dtype %d", dtype);
  }
  OFP_LAUNCHED();
  return OFP_OK;
}

}  // extern "C"
