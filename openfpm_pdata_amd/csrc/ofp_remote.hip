// ofp_remote.hip -- one rank's side of the multi-rank exchanges: the remote
// part of ghost_get (label every particle against the internal ghost boxes that
// other ranks need, order the (processor, particle, shift) outputs by processor,
// pack the send buffer) and map (find every particle's processor, order by
// destination, split into the particles that stay and the send buffers).
//
//  * label (labelParticlesGhost_impl, Vector/cuda/vector_dist_comm_util_funcs.cuh:
//    53-177): per particle the number of outputs (num_proc_ghost_each_part,
//    Vector/cuda/vector_dist_cuda_funcs.cuh:33-43), an exclusive scan, and the
//    outputs (proc_label_id_ghost, :19-31), both by the walk of
//    ghost_processorID_general_impl (Decomposition/cuda/ie_ghost_gpu.cuh:40-96):
//    the boxes of the particle's geometric cell in list order; a box that
//    contains the particle (half-open) emits (its processor, p | shift << 32)
//    unless the previous box had the same processor and shift and an output was
//    made since that pair began.
//  * order by processor: the reference sorts the outputs by processor id and
//    finds the per-processor offsets (find_buffer_offsets); here a stable
//    counting sort over the destination processors (slots in ascending
//    processor id): one wave per tile of 1024 outputs counts per slot into a
//    slot-major table, one exclusive scan of the table gives every (slot, tile)
//    its first output, and the same walk scatters (ranks inside a round of 64
//    from ballots over the slots present).  Output order: processor ascending,
//    then the label order (particle-major).
//  * pack (process_ghost_particles_pos / _prp, vector_dist_cuda_funcs.cuh:90-131):
//    send entry i = position of particle id minus shift sid, and its planes.
//
// map (labelParticleProcessor + fill_send_map_buf, Vector/vector_dist_comm.hpp:
// 1398-1456 and 1240-1310): periodic wrap in place (ofp_apply_bc), then the
// processor of the first sub-domain of the particle's geometric cell that
// contains it (processorID_impl, Decomposition/cuda/CartDecomposition_gpu.cuh:
// 13-52; half-open, the upper face closed where it is the domain's upper face in
// a non-periodic dimension), the particles ordered by destination with the same
// stable counting sort (processor ascending, own rank included; the reference's
// order inside a destination comes from atomics, here it is ascending particle
// index), and process_map_particles (vector_dist_cuda_funcs.cuh:76-85): the own
// rank's range to the new local vector, the others to the send buffer.  A
// particle inside no sub-domain (beyond a non-periodic face) is dropped.
//
// The geometric cell list (ie_ghost_gpu's geo_cell) is built on the host:
// a grid over the boxes' bounding box, at most 64 cells per dimension and no
// cell narrower than the narrowest box; a cell lists, in ascending box id, the
// boxes whose closed extent meets the cell widened by 1e-9 of the grid's
// extent (so a point's box is always in its cell's list, whatever the rounding
// of its cell index).  The lists fix the walk's order, and with it which
// outputs the same-pair rule merges.
#include "ofp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace ofp {

constexpr int kRemoteTile = 1024;     // outputs per wave in the ordering passes
constexpr int kRemoteMaxSlots = 256;  // destination processors per rank
constexpr uint32_t kGeoMaxDiv = 64;   // geometric cells per dimension

struct GeoGrid {
  double lo[3], cs[3];
  uint32_t div[3];
};

template <class T>
struct RemoteBoxes {
  GeoGrid g;
  const uint32_t* cell_start;  // n_cells + 1
  const uint32_t* cell_box;    // box ids, ascending per cell
  const T* lo;                 // 3 per box
  const T* hi;
  const uint32_t* proc;  // destination processor per box
  const uint32_t* slot;  // its rank among the distinct processors (ascending)
  const uint32_t* shift;
};

}  // namespace ofp

struct ofp_map {
  int dtype = OFP_F32;
  uint32_t rank = 0;
  ofp::GeoGrid g{};
  double dom_lo[3] = {0, 0, 0}, dom_hi[3] = {1, 1, 1};
  int bc[3] = {0, 0, 0};
  ofp::DevBuf cell_start, cell_box, blo, bhi, bproc, bslot, bshift;
  std::vector<uint32_t> procs;  // processors of the sub-domains, ascending (own rank included)
  ofp::DevBuf oslot, oproc, oid, table, toff, sproc, sid, pcount, tmp;
  size_t n = 0, n_stay = 0, stay_begin = 0, n_lost = 0;
  std::vector<uint64_t> counts;  // particles per processor slot of the last label
  bool labelled = false;
};

struct ofp_ghost_remote {
  int dtype = OFP_F32;
  int n_boxes = 0;
  ofp::GeoGrid g{};
  ofp::DevBuf cell_start, cell_box, blo, bhi, bproc, bslot, bshift, shifts;
  std::vector<uint32_t> procs;  // distinct destination processors, ascending
  ofp::DevBuf cnt, starts, tmp, oslot, oproc, oid, table, toff, sproc, sid, pcount;
  size_t n_out = 0;
  std::vector<uint64_t> counts;  // outputs per destination processor of the last label
  bool labelled = false;
};

namespace ofp {

// The grid and the per-cell box lists (the rule of the file comment), in double.
static void geo_build(int n, const double* lo, const double* hi, GeoGrid* g, std::vector<uint32_t>* start,
                      std::vector<uint32_t>* boxes) {
  double glo[3], ghi[3], wmin[3];
  for (int k = 0; k < 3; ++k) {
    glo[k] = INFINITY;
    ghi[k] = -INFINITY;
    wmin[k] = INFINITY;
  }
  for (int b = 0; b < n; ++b)
    for (int k = 0; k < 3; ++k) {
      glo[k] = std::min(glo[k], lo[3 * b + k]);
      ghi[k] = std::max(ghi[k], hi[3 * b + k]);
      const double w = hi[3 * b + k] - lo[3 * b + k];
      if (w > 0.0) wmin[k] = std::min(wmin[k], w);
    }
  double eps[3];
  for (int k = 0; k < 3; ++k) {
    if (n == 0) {
      glo[k] = 0.0;
      ghi[k] = 1.0;
    }
    const double ext = ghi[k] > glo[k] ? ghi[k] - glo[k] : 1.0;
    double d = std::isfinite(wmin[k]) ? std::floor(ext / wmin[k]) : 1.0;
    d = std::max(1.0, std::min(d, double(kGeoMaxDiv)));
    g->lo[k] = glo[k];
    g->div[k] = uint32_t(d);
    g->cs[k] = ext / d;
    eps[k] = 1e-9 * ext;
  }
  const size_t nc = size_t(g->div[0]) * g->div[1] * g->div[2];
  start->assign(nc + 1, 0u);
  boxes->clear();
  for (uint32_t cz = 0; cz < g->div[2]; ++cz)
    for (uint32_t cy = 0; cy < g->div[1]; ++cy)
      for (uint32_t cx = 0; cx < g->div[0]; ++cx) {
        const size_t c = cx + size_t(g->div[0]) * (cy + size_t(g->div[1]) * cz);
        const uint32_t ci[3] = {cx, cy, cz};
        (*start)[c] = uint32_t(boxes->size());
        for (int b = 0; b < n; ++b) {
          bool meet = true;
          for (int k = 0; k < 3; ++k) {
            const double clo = g->lo[k] + double(ci[k]) * g->cs[k] - eps[k];
            const double chi = g->lo[k] + double(ci[k] + 1) * g->cs[k] + eps[k];
            meet = meet && lo[3 * b + k] <= chi && hi[3 * b + k] >= clo;
          }
          if (meet) boxes->push_back(uint32_t(b));
        }
      }
  (*start)[nc] = uint32_t(boxes->size());
}

// Geometric cell of a point: floor((x - lo) / cs) per dimension in double,
// clamped to the grid.
__device__ __forceinline__ uint32_t geo_cell(const GeoGrid& g, double x, double y, double z) {
  const double v[3] = {x, y, z};
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double f = floor((v[k] - g.lo[k]) / g.cs[k]);
    c[k] = f < 0.0 ? 0u : (f >= double(g.div[k]) ? g.div[k] - 1u : uint32_t(f));
  }
  return c[0] + g.div[0] * (c[1] + g.div[1] * c[2]);
}

// The walk of ghost_processorID_general_impl; emit(n, box) for each output.
template <class T, class F>
__device__ __forceinline__ uint32_t remote_walk(const RemoteBoxes<T>& B, T x, T y, T z, F&& emit) {
  const uint32_t cell = geo_cell(B.g, double(x), double(y), double(z));
  const uint32_t b0 = B.cell_start[cell], b1 = B.cell_start[cell + 1];
  uint32_t n = 0;
  bool merged = false;  // an output was made since the current (processor, shift) pair began
  uint32_t pprev = 0, sprev = 0;
  for (uint32_t i = b0; i < b1; ++i) {
    const uint32_t b = B.cell_box[i];
    const uint32_t pa = B.proc[b], sa = B.shift[b];
    if (i > b0) merged = merged && pa == pprev && sa == sprev;
    const bool in = B.lo[3 * b] <= x && x < B.hi[3 * b] && B.lo[3 * b + 1] <= y && y < B.hi[3 * b + 1] &&
                    B.lo[3 * b + 2] <= z && z < B.hi[3 * b + 2];
    if (in && !merged) {
      emit(n, b);
      merged = true;
      ++n;
    }
    pprev = pa;
    sprev = sa;
  }
  return n;
}

template <class T>
__global__ __launch_bounds__(256) void k_remote_count(RemoteBoxes<T> B, const T* __restrict__ pos, size_t stride,
                                                      uint32_t n, uint32_t* __restrict__ cnt) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  cnt[p] = remote_walk(B, pos[p], pos[stride + p], pos[2 * stride + p], [](uint32_t, uint32_t) {});
}

template <class T>
__global__ __launch_bounds__(256) void k_remote_emit(RemoteBoxes<T> B, const T* __restrict__ pos, size_t stride,
                                                     uint32_t n, const uint32_t* __restrict__ starts,
                                                     uint32_t* __restrict__ oslot, uint32_t* __restrict__ oproc,
                                                     unsigned long long* __restrict__ oid) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t base = starts[p];
  remote_walk(B, pos[p], pos[stride + p], pos[2 * stride + p], [&](uint32_t k, uint32_t b) {
    oslot[base + k] = B.slot[b];
    oproc[base + k] = B.proc[b];
    oid[base + k] = static_cast<unsigned long long>(p) | (static_cast<unsigned long long>(B.shift[b]) << 32);
  });
}

// One wave per tile of kRemoteTile outputs, in rounds of 64 in output order.
// kScatter = false: table[slot * n_tiles + tile] = outputs of the tile in slot.
// kScatter = true: every output goes to toff[slot * n_tiles + tile] + (earlier
// outputs of the tile in the same slot).
template <bool kScatter>
__global__ __launch_bounds__(256) void k_remote_order(const uint32_t* __restrict__ oslot,
                                                      const uint32_t* __restrict__ oproc,
                                                      const unsigned long long* __restrict__ oid, uint32_t m,
                                                      uint32_t n_slots, uint32_t n_tiles,
                                                      uint32_t* __restrict__ table,
                                                      const uint32_t* __restrict__ toff,
                                                      uint32_t* __restrict__ sproc,
                                                      unsigned long long* __restrict__ sid) {
  __shared__ uint32_t run[256 / kWave][kRemoteMaxSlots];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint32_t tile = blockIdx.x * (256 / kWave) + uint32_t(w);
  if (tile >= n_tiles) return;  // wave-uniform
  uint32_t* const r = run[w];
  for (uint32_t s = uint32_t(lane); s < n_slots; s += kWave) r[s] = kScatter ? toff[size_t(s) * n_tiles + tile] : 0u;
  __builtin_amdgcn_wave_barrier();
  const uint32_t t0 = tile * uint32_t(kRemoteTile);
  for (uint32_t i0 = t0; i0 < t0 + uint32_t(kRemoteTile) && i0 < m; i0 += kWave) {
    const uint32_t i = i0 + uint32_t(lane);
    const bool valid = i < m;
    const uint32_t key = valid ? oslot[i] : 0xFFFFFFFFu;
    unsigned long long todo = __ballot(valid);
    while (todo) {
      const int leader = __builtin_ctzll(todo);
      const uint32_t k = __builtin_amdgcn_readlane(key, leader);
      const unsigned long long mk = __ballot(key == k);
      const uint32_t base = r[k];
      if (kScatter && key == k) {
        const uint32_t dst = base + __builtin_amdgcn_mbcnt_hi(uint32_t(mk >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mk), 0u));
        sproc[dst] = oproc[i];
        sid[dst] = oid[i];
      }
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) r[k] = base + uint32_t(__popcll(mk));
      __builtin_amdgcn_wave_barrier();
      todo &= ~mk;
    }
  }
  if (!kScatter)
    for (uint32_t s = uint32_t(lane); s < n_slots; s += kWave) table[size_t(s) * n_tiles + tile] = r[s];
}

// counts[s] = first output of slot s (the scanned table's entry of tile 0), s <= n_slots
__global__ void k_remote_slot_starts(const uint32_t* __restrict__ toff, uint32_t n_slots, uint32_t n_tiles,
                                     uint32_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= n_slots) out[s] = toff[size_t(s) * n_tiles];
}

template <class T>
__global__ __launch_bounds__(256) void k_remote_pack(const T* __restrict__ pos, size_t stride,
                                                     const unsigned long long* __restrict__ sid, uint32_t m,
                                                     const T* __restrict__ shifts, T* __restrict__ out,
                                                     size_t out_stride, PlaneArgs pa) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const unsigned long long v = sid[i];
  const uint32_t id = uint32_t(v), sh = uint32_t(v >> 32);
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k * out_stride + i] = pos[k * stride + id] - shifts[3 * sh + k];
  copy_planes(pa, id, i);
}

// map: slot of the first sub-domain that contains the particle (lost_slot if
// none); closed_hi[k]: the domain's upper face belongs to the sub-domains that
// end on it (non-periodic dimension).
template <class T>
__global__ __launch_bounds__(256) void k_map_label(RemoteBoxes<T> B, const T* __restrict__ pos, size_t stride,
                                                   uint32_t n, T dhi0, T dhi1, T dhi2, int ch0, int ch1, int ch2,
                                                   uint32_t lost_slot, uint32_t* __restrict__ oslot,
                                                   uint32_t* __restrict__ oproc, unsigned long long* __restrict__ oid) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const T x[3] = {pos[p], pos[stride + p], pos[2 * stride + p]};
  const T dhi[3] = {dhi0, dhi1, dhi2};
  const int ch[3] = {ch0, ch1, ch2};
  const uint32_t cell = geo_cell(B.g, double(x[0]), double(x[1]), double(x[2]));
  uint32_t slot = lost_slot, proc = 0xFFFFFFFFu;
  for (uint32_t i = B.cell_start[cell]; i < B.cell_start[cell + 1]; ++i) {
    const uint32_t b = B.cell_box[i];
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T lo = B.lo[3 * b + k], hi = B.hi[3 * b + k];
      in = in && lo <= x[k] && (x[k] < hi || (ch[k] && hi == dhi[k] && x[k] == hi));
    }
    if (in) {
      slot = B.slot[b];
      proc = B.proc[b];
      break;
    }
  }
  oslot[p] = slot;
  oproc[p] = proc;
  oid[p] = p;
}

// Entry i < n_kept of the destination order: the own rank's range
// [stay_begin, stay_begin + n_stay) goes to the new local vector, the rest to
// the send buffer (ranks ascending).
template <class T>
__global__ __launch_bounds__(256) void k_map_pack(const T* __restrict__ pos, size_t stride,
                                                  const unsigned long long* __restrict__ sid, uint32_t n_kept,
                                                  uint32_t stay_begin, uint32_t n_stay, T* __restrict__ stay,
                                                  size_t stay_stride, T* __restrict__ send, size_t send_stride,
                                                  PlaneArgs stay_pa, PlaneArgs send_pa) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_kept) return;
  const uint32_t id = uint32_t(sid[i]);
  if (i >= stay_begin && i < stay_begin + n_stay) {
    const uint32_t j = i - stay_begin;
#pragma unroll
    for (int k = 0; k < 3; ++k) stay[k * stay_stride + j] = pos[k * stride + id];
    copy_planes(stay_pa, id, j);
  } else {
    const uint32_t j = i < stay_begin ? i : i - n_stay;
#pragma unroll
    for (int k = 0; k < 3; ++k) send[k * send_stride + j] = pos[k * stride + id];
    copy_planes(send_pa, id, j);
  }
}

// Stable order of m keyed entries by slot (the two passes and the scan between):
// sorted proc / id arrays and, per slot, its first entry (n_slots + 1 values on
// the host).
static int order_by_slot(DevBuf& table, DevBuf& toff, DevBuf& tmp, DevBuf& pcount, const uint32_t* oslot,
                         const uint32_t* oproc, const unsigned long long* oid, uint32_t m, uint32_t n_slots,
                         uint32_t* sproc, unsigned long long* sid, std::vector<uint32_t>* starts, hipStream_t s) {
  starts->assign(n_slots + 1, 0u);
  if (m == 0) return OFP_OK;
  const uint32_t n_tiles = (m + uint32_t(kRemoteTile) - 1) / uint32_t(kRemoteTile);
  const size_t tn = size_t(n_slots) * n_tiles;
  OFP_TRY(table.ensure((tn + 1) * sizeof(uint32_t)));
  OFP_TRY(toff.ensure((tn + 1) * sizeof(uint32_t)));
  OFP_TRY(tmp.ensure(scan_tmp_bytes(tn)));
  OFP_TRY(pcount.ensure((size_t(n_slots) + 1) * sizeof(uint32_t)));
  const unsigned grid = grid_for(n_tiles, 256 / kWave);
  k_remote_order<false><<<grid, 256, 0, s>>>(oslot, nullptr, nullptr, m, n_slots, n_tiles, table.as<uint32_t>(),
                                             nullptr, nullptr, nullptr);
  OFP_LAUNCHED();
  OFP_TRY(scan_excl_u32(table.as<uint32_t>(), toff.as<uint32_t>(), tn, tmp.p, s));
  k_remote_order<true><<<grid, 256, 0, s>>>(oslot, oproc, oid, m, n_slots, n_tiles, nullptr, toff.as<uint32_t>(),
                                            sproc, sid);
  OFP_LAUNCHED();
  k_remote_slot_starts<<<grid_for(n_slots + 1, 256), 256, 0, s>>>(toff.as<uint32_t>(), n_slots, n_tiles,
                                                                 pcount.as<uint32_t>());
  OFP_LAUNCHED();
  OFP_HIP(hipMemcpyAsync(starts->data(), pcount.as<uint32_t>(), (n_slots + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  return OFP_OK;
}

template <class T>
static RemoteBoxes<T> boxes_of(const ofp_ghost_remote* h) {
  RemoteBoxes<T> B;
  B.g = h->g;
  B.cell_start = h->cell_start.as<uint32_t>();
  B.cell_box = h->cell_box.as<uint32_t>();
  B.lo = h->blo.as<T>();
  B.hi = h->bhi.as<T>();
  B.proc = h->bproc.as<uint32_t>();
  B.slot = h->bslot.as<uint32_t>();
  B.shift = h->bshift.as<uint32_t>();
  return B;
}

template <class T>
static int remote_label(ofp_ghost_remote* h, const void* pos_v, size_t stride, size_t n, hipStream_t s) {
  const T* pos = static_cast<const T*>(pos_v);
  const RemoteBoxes<T> B = boxes_of<T>(h);
  OFP_TRY(h->cnt.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->starts.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->tmp.ensure(scan_tmp_bytes(n)));
  if (n > 0) {
    k_remote_count<T><<<grid_for(n, 256), 256, 0, s>>>(B, pos, stride, uint32_t(n), h->cnt.as<uint32_t>());
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32(h->cnt.as<uint32_t>(), h->starts.as<uint32_t>(), n, h->tmp.p, s));
  uint32_t m = 0;
  OFP_HIP(hipMemcpyAsync(&m, h->starts.as<uint32_t>() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  OFP_TRY(h->oslot.ensure((size_t(m) + 1) * sizeof(uint32_t)));
  OFP_TRY(h->oproc.ensure((size_t(m) + 1) * sizeof(uint32_t)));
  OFP_TRY(h->oid.ensure((size_t(m) + 1) * sizeof(unsigned long long)));
  OFP_TRY(h->sproc.ensure((size_t(m) + 1) * sizeof(uint32_t)));
  OFP_TRY(h->sid.ensure((size_t(m) + 1) * sizeof(unsigned long long)));
  const uint32_t n_slots = uint32_t(h->procs.size());
  if (m > 0) {
    k_remote_emit<T><<<grid_for(n, 256), 256, 0, s>>>(B, pos, stride, uint32_t(n), h->starts.as<uint32_t>(),
                                                       h->oslot.as<uint32_t>(), h->oproc.as<uint32_t>(),
                                                       h->oid.as<unsigned long long>());
    OFP_LAUNCHED();
  }
  std::vector<uint32_t> st;
  OFP_TRY(order_by_slot(h->table, h->toff, h->tmp, h->pcount, h->oslot.as<uint32_t>(), h->oproc.as<uint32_t>(),
                        h->oid.as<unsigned long long>(), m, n_slots, h->sproc.as<uint32_t>(),
                        h->sid.as<unsigned long long>(), &st, s));
  h->counts.assign(n_slots, 0);
  if (m > 0)
    for (uint32_t k = 0; k < n_slots; ++k) h->counts[k] = st[k + 1] - st[k];
  h->n_out = m;
  h->labelled = true;
  return OFP_OK;
}

template <class T>
static int remote_create(ofp_ghost_remote* h, int nb, const void* lo_v, const void* hi_v, const uint32_t* proc,
                         const uint32_t* shift, int n_shifts, const void* shifts_v) {
  const T* lo = static_cast<const T*>(lo_v);
  const T* hi = static_cast<const T*>(hi_v);
  std::vector<double> dlo(size_t(3) * nb), dhi(size_t(3) * nb);
  for (size_t i = 0; i < dlo.size(); ++i) {
    dlo[i] = double(lo[i]);
    dhi[i] = double(hi[i]);
  }
  std::vector<uint32_t> start, boxes;
  geo_build(nb, dlo.data(), dhi.data(), &h->g, &start, &boxes);
  std::vector<uint32_t> ps(proc, proc + nb);
  std::sort(ps.begin(), ps.end());
  ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
  OFP_CHECK(ps.size() <= size_t(kRemoteMaxSlots), OFP_E_LIMIT,
            "at most %d destination processors per rank (got %zu)", kRemoteMaxSlots, ps.size());
  std::vector<uint32_t> slot(nb);
  for (int b = 0; b < nb; ++b) {
    OFP_CHECK(shift[b] < uint32_t(n_shifts), OFP_E_ARG, "box %d: shift id %u >= %d", b, shift[b], n_shifts);
    slot[b] = uint32_t(std::lower_bound(ps.begin(), ps.end(), proc[b]) - ps.begin());
  }
  h->procs = ps;
  auto up = [](ofp::DevBuf& d, const void* src, size_t bytes) -> int {
    OFP_TRY(d.ensure(bytes ? bytes : 4));
    if (bytes) OFP_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
    return OFP_OK;
  };
  OFP_TRY(up(h->cell_start, start.data(), start.size() * sizeof(uint32_t)));
  OFP_TRY(up(h->cell_box, boxes.data(), boxes.size() * sizeof(uint32_t)));
  OFP_TRY(up(h->blo, lo, size_t(3) * nb * sizeof(T)));
  OFP_TRY(up(h->bhi, hi, size_t(3) * nb * sizeof(T)));
  OFP_TRY(up(h->bproc, proc, size_t(nb) * sizeof(uint32_t)));
  OFP_TRY(up(h->bslot, slot.data(), size_t(nb) * sizeof(uint32_t)));
  OFP_TRY(up(h->bshift, shift, size_t(nb) * sizeof(uint32_t)));
  OFP_TRY(up(h->shifts, shifts_v, size_t(3) * n_shifts * sizeof(T)));
  return OFP_OK;
}

template <class T>
static int map_create(ofp_map* h, int n_sub, const void* lo_v, const void* hi_v, const uint32_t* proc) {
  const T* lo = static_cast<const T*>(lo_v);
  const T* hi = static_cast<const T*>(hi_v);
  std::vector<double> dlo(size_t(3) * n_sub), dhi(size_t(3) * n_sub);
  for (size_t i = 0; i < dlo.size(); ++i) {
    dlo[i] = double(lo[i]);
    dhi[i] = double(hi[i]);
  }
  std::vector<uint32_t> start, boxes;
  geo_build(n_sub, dlo.data(), dhi.data(), &h->g, &start, &boxes);
  std::vector<uint32_t> ps(proc, proc + n_sub);
  std::sort(ps.begin(), ps.end());
  ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
  OFP_CHECK(ps.size() < size_t(kRemoteMaxSlots), OFP_E_LIMIT, "at most %d processors (got %zu)",
            kRemoteMaxSlots - 1, ps.size());
  std::vector<uint32_t> slot(n_sub), zero(n_sub, 0u);
  for (int b = 0; b < n_sub; ++b)
    slot[b] = uint32_t(std::lower_bound(ps.begin(), ps.end(), proc[b]) - ps.begin());
  h->procs = ps;
  auto up = [](DevBuf& d, const void* src, size_t bytes) -> int {
    OFP_TRY(d.ensure(bytes ? bytes : 4));
    if (bytes) OFP_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
    return OFP_OK;
  };
  OFP_TRY(up(h->cell_start, start.data(), start.size() * sizeof(uint32_t)));
  OFP_TRY(up(h->cell_box, boxes.data(), boxes.size() * sizeof(uint32_t)));
  OFP_TRY(up(h->blo, lo, size_t(3) * n_sub * sizeof(T)));
  OFP_TRY(up(h->bhi, hi, size_t(3) * n_sub * sizeof(T)));
  OFP_TRY(up(h->bproc, proc, size_t(n_sub) * sizeof(uint32_t)));
  OFP_TRY(up(h->bslot, slot.data(), size_t(n_sub) * sizeof(uint32_t)));
  OFP_TRY(up(h->bshift, zero.data(), size_t(n_sub) * sizeof(uint32_t)));
  return OFP_OK;
}

template <class T>
static int map_label(ofp_map* h, const void* pos_v, size_t stride, size_t n, hipStream_t s) {
  const T* pos = static_cast<const T*>(pos_v);
  RemoteBoxes<T> B;
  B.g = h->g;
  B.cell_start = h->cell_start.as<uint32_t>();
  B.cell_box = h->cell_box.as<uint32_t>();
  B.lo = h->blo.as<T>();
  B.hi = h->bhi.as<T>();
  B.proc = h->bproc.as<uint32_t>();
  B.slot = h->bslot.as<uint32_t>();
  B.shift = h->bshift.as<uint32_t>();
  const uint32_t n_slots = uint32_t(h->procs.size()) + 1;  // the last slot: lost particles
  OFP_TRY(h->oslot.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->oproc.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->oid.ensure((n + 1) * sizeof(unsigned long long)));
  OFP_TRY(h->sproc.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->sid.ensure((n + 1) * sizeof(unsigned long long)));
  if (n > 0) {
    const T dh[3] = {T(h->dom_hi[0]), T(h->dom_hi[1]), T(h->dom_hi[2])};
    k_map_label<T><<<grid_for(n, 256), 256, 0, s>>>(B, pos, stride, uint32_t(n), dh[0], dh[1], dh[2], !h->bc[0],
                                                    !h->bc[1], !h->bc[2], n_slots - 1, h->oslot.as<uint32_t>(),
                                                    h->oproc.as<uint32_t>(), h->oid.as<unsigned long long>());
    OFP_LAUNCHED();
  }
  std::vector<uint32_t> st;
  OFP_TRY(order_by_slot(h->table, h->toff, h->tmp, h->pcount, h->oslot.as<uint32_t>(), h->oproc.as<uint32_t>(),
                        h->oid.as<unsigned long long>(), uint32_t(n), n_slots, h->sproc.as<uint32_t>(),
                        h->sid.as<unsigned long long>(), &st, s));
  h->counts.assign(n_slots, 0);
  if (n > 0)
    for (uint32_t k = 0; k < n_slots; ++k) h->counts[k] = st[k + 1] - st[k];
  const size_t own = size_t(std::lower_bound(h->procs.begin(), h->procs.end(), h->rank) - h->procs.begin());
  const bool has_own = own < h->procs.size() && h->procs[own] == h->rank;
  h->n = n;
  h->n_lost = n > 0 ? h->counts[n_slots - 1] : 0;
  h->n_stay = has_own && n > 0 ? h->counts[own] : 0;
  h->stay_begin = has_own && n > 0 ? st[own] : 0;
  h->labelled = true;
  return OFP_OK;
}

}  // namespace ofp

using namespace ofp;

extern "C" {

int ofp_map_create(ofp_map** out, int dtype, int n_sub, const void* sub_lo, const void* sub_hi,
                   const uint32_t* sub_proc, const void* dom_lo, const void* dom_hi, const int* bc, uint32_t rank) {
  OFP_CHECK(out, OFP_E_ARG, "ofp_map_create: null handle pointer");
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  OFP_CHECK(n_sub >= 1 && sub_lo && sub_hi && sub_proc && dom_lo && dom_hi && bc, OFP_E_ARG,
            "ofp_map_create: null argument or no sub-domain");
  ofp_map* h = new ofp_map();
  h->dtype = dtype;
  h->rank = rank;
  for (int k = 0; k < 3; ++k) {
    h->dom_lo[k] = dtype == OFP_F32 ? double(static_cast<const float*>(dom_lo)[k]) : static_cast<const double*>(dom_lo)[k];
    h->dom_hi[k] = dtype == OFP_F32 ? double(static_cast<const float*>(dom_hi)[k]) : static_cast<const double*>(dom_hi)[k];
    h->bc[k] = bc[k] ? 1 : 0;
  }
  const int rc = dtype == OFP_F32 ? map_create<float>(h, n_sub, sub_lo, sub_hi, sub_proc)
                                  : map_create<double>(h, n_sub, sub_lo, sub_hi, sub_proc);
  if (rc != OFP_OK) {
    ofp_map_destroy(h);
    return rc;
  }
  *out = h;
  return OFP_OK;
}

int ofp_map_destroy(ofp_map* h) {
  if (!h) return OFP_OK;
  DevBuf* bufs[] = {&h->cell_start, &h->cell_box, &h->blo, &h->bhi, &h->bproc, &h->bslot, &h->bshift,
                    &h->oslot, &h->oproc, &h->oid, &h->table, &h->toff, &h->sproc, &h->sid, &h->pcount, &h->tmp};
  for (DevBuf* b : bufs) b->release();
  delete h;
  return OFP_OK;
}

int ofp_map_label(ofp_map* h, void* pos, size_t stride, size_t n, size_t* n_stay, int* n_dest, size_t* n_lost,
                  void* stream) {
  OFP_CHECK(h && n_stay && n_dest && n_lost && (n == 0 || pos), OFP_E_ARG, "ofp_map_label: null argument");
  OFP_CHECK(n < (size_t(1) << 32) && stride >= n, OFP_E_ARG, "ofp_map_label: bad size or stride");
  h->labelled = false;  // a failed label leaves nothing for procs / pack
  hipStream_t s = as_stream(stream);
  if (n > 0) {
    if (h->dtype == OFP_F32) {
      const float lo[3] = {float(h->dom_lo[0]), float(h->dom_lo[1]), float(h->dom_lo[2])};
      const float hi[3] = {float(h->dom_hi[0]), float(h->dom_hi[1]), float(h->dom_hi[2])};
      OFP_TRY(ofp_apply_bc(OFP_F32, pos, stride, n, lo, hi, h->bc, stream));
    } else {
      OFP_TRY(ofp_apply_bc(OFP_F64, pos, stride, n, h->dom_lo, h->dom_hi, h->bc, stream));
    }
  }
  OFP_TRY(h->dtype == OFP_F32 ? map_label<float>(h, pos, stride, n, s) : map_label<double>(h, pos, stride, n, s));
  int nd = 0;
  for (size_t k = 0; k < h->procs.size(); ++k) nd += h->procs[k] != h->rank && h->counts[k] > 0 ? 1 : 0;
  *n_stay = h->n_stay;
  *n_dest = nd;
  *n_lost = h->n_lost;
  return OFP_OK;
}

int ofp_map_procs(const ofp_map* h, uint32_t* procs, uint64_t* counts) {
  OFP_CHECK(h && procs && counts, OFP_E_ARG, "ofp_map_procs: null argument");
  OFP_CHECK(h->labelled, OFP_E_STATE, "ofp_map_procs: no label yet");
  int k = 0;
  for (size_t s = 0; s < h->procs.size(); ++s)
    if (h->procs[s] != h->rank && h->counts[s] > 0) {
      procs[k] = h->procs[s];
      counts[k] = h->counts[s];
      ++k;
    }
  return OFP_OK;
}

int ofp_map_pack(const ofp_map* h, const void* pos, size_t stride, void* stay_pos, size_t stay_stride,
                 void* send_pos, size_t send_stride, const ofp_plane* planes_in, const ofp_plane* stay_planes,
                 const ofp_plane* send_planes, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_map_pack: null handle");
  OFP_CHECK(h->labelled, OFP_E_STATE, "ofp_map_pack: no label yet");
  const size_t n_send = h->n - h->n_stay - h->n_lost;
  OFP_CHECK(h->n == 0 || pos, OFP_E_ARG, "ofp_map_pack: null positions");
  OFP_CHECK(h->n_stay == 0 || (stay_pos && stay_stride >= h->n_stay), OFP_E_ARG,
            "ofp_map_pack: stay buffer null or its stride below the staying count");
  OFP_CHECK(n_send == 0 || (send_pos && send_stride >= n_send), OFP_E_ARG,
            "ofp_map_pack: send buffer null or its stride below the leaving count");
  PlaneArgs spa, dpa;
  OFP_TRY(make_plane_args(planes_in, stay_planes, n_planes, &spa));
  OFP_TRY(make_plane_args(planes_in, send_planes, n_planes, &dpa));
  OFP_CHECK(n_planes == 0 || (stay_planes && send_planes), OFP_E_ARG, "ofp_map_pack: null output planes");
  const uint32_t n_kept = uint32_t(h->n - h->n_lost);
  if (n_kept == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  const unsigned long long* sid = h->sid.as<unsigned long long>();
  if (h->dtype == OFP_F32)
    k_map_pack<float><<<grid_for(n_kept, 256), 256, 0, s>>>(
        static_cast<const float*>(pos), stride, sid, n_kept, uint32_t(h->stay_begin), uint32_t(h->n_stay),
        static_cast<float*>(stay_pos), stay_stride, static_cast<float*>(send_pos), send_stride, spa, dpa);
  else
    k_map_pack<double><<<grid_for(n_kept, 256), 256, 0, s>>>(
        static_cast<const double*>(pos), stride, sid, n_kept, uint32_t(h->stay_begin), uint32_t(h->n_stay),
        static_cast<double*>(stay_pos), stay_stride, static_cast<double*>(send_pos), send_stride, spa, dpa);
  OFP_LAUNCHED();
  return OFP_OK;
}

int ofp_ghost_remote_geo(int n_boxes, const double* box_lo, const double* box_hi, double grid_lo[3],
                         double cell_size[3], uint32_t div[3], uint32_t* cell_start, uint32_t* cell_boxes,
                         size_t* n_entries) {
  OFP_CHECK(n_boxes >= 0 && (n_boxes == 0 || (box_lo && box_hi)) && grid_lo && cell_size && div && n_entries,
            OFP_E_ARG, "ofp_ghost_remote_geo: null argument");
  GeoGrid g;
  std::vector<uint32_t> start, boxes;
  geo_build(n_boxes, box_lo, box_hi, &g, &start, &boxes);
  for (int k = 0; k < 3; ++k) {
    grid_lo[k] = g.lo[k];
    cell_size[k] = g.cs[k];
    div[k] = g.div[k];
  }
  *n_entries = boxes.size();
  if (cell_start) std::memcpy(cell_start, start.data(), start.size() * sizeof(uint32_t));
  if (cell_boxes) std::memcpy(cell_boxes, boxes.data(), boxes.size() * sizeof(uint32_t));
  return OFP_OK;
}

int ofp_ghost_remote_create(ofp_ghost_remote** out, int dtype, int n_boxes, const void* box_lo, const void* box_hi,
                            const uint32_t* proc, const uint32_t* shift_id, int n_shifts, const void* shifts) {
  OFP_CHECK(out, OFP_E_ARG, "ofp_ghost_remote_create: null handle pointer");
  OFP_CHECK(dtype == OFP_F32 || dtype == OFP_F64, OFP_E_ARG, "bad dtype %d", dtype);
  OFP_CHECK(n_boxes >= 0 && n_shifts >= 1 && shifts, OFP_E_ARG, "ofp_ghost_remote_create: bad box or shift count");
  OFP_CHECK(n_boxes == 0 || (box_lo && box_hi && proc && shift_id), OFP_E_ARG,
            "ofp_ghost_remote_create: null box array");
  ofp_ghost_remote* h = new ofp_ghost_remote();
  h->dtype = dtype;
  h->n_boxes = n_boxes;
  const int rc = dtype == OFP_F32
                     ? remote_create<float>(h, n_boxes, box_lo, box_hi, proc, shift_id, n_shifts, shifts)
                     : remote_create<double>(h, n_boxes, box_lo, box_hi, proc, shift_id, n_shifts, shifts);
  if (rc != OFP_OK) {
    ofp_ghost_remote_destroy(h);
    return rc;
  }
  *out = h;
  return OFP_OK;
}

int ofp_ghost_remote_destroy(ofp_ghost_remote* h) {
  if (!h) return OFP_OK;
  DevBuf* bufs[] = {&h->cell_start, &h->cell_box, &h->blo, &h->bhi, &h->bproc, &h->bslot, &h->bshift,
                    &h->shifts, &h->cnt, &h->starts, &h->tmp, &h->oslot, &h->oproc, &h->oid,
                    &h->table, &h->toff, &h->sproc, &h->sid, &h->pcount};
  for (DevBuf* b : bufs) b->release();
  delete h;
  return OFP_OK;
}

int ofp_ghost_remote_label(ofp_ghost_remote* h, const void* pos, size_t stride, size_t n, size_t* n_out,
                           int* n_procs, void* stream) {
  OFP_CHECK(h && n_out && n_procs && (n == 0 || pos), OFP_E_ARG, "ofp_ghost_remote_label: null argument");
  OFP_CHECK(n < (size_t(1) << 32) && stride >= n, OFP_E_ARG, "ofp_ghost_remote_label: bad size or stride");
  h->labelled = false;  // a failed label leaves nothing for procs / pack
  hipStream_t s = as_stream(stream);
  OFP_TRY(h->dtype == OFP_F32 ? remote_label<float>(h, pos, stride, n, s) : remote_label<double>(h, pos, stride, n, s));
  *n_out = h->n_out;
  int np = 0;
  for (uint64_t c : h->counts) np += c > 0 ? 1 : 0;
  *n_procs = np;
  return OFP_OK;
}

int ofp_ghost_remote_procs(const ofp_ghost_remote* h, uint32_t* procs, uint64_t* counts) {
  OFP_CHECK(h && procs && counts, OFP_E_ARG, "ofp_ghost_remote_procs: null argument");
  OFP_CHECK(h->labelled, OFP_E_STATE, "ofp_ghost_remote_procs: no label yet");
  int k = 0;
  for (size_t s = 0; s < h->procs.size(); ++s)
    if (h->counts[s] > 0) {
      procs[k] = h->procs[s];
      counts[k] = h->counts[s];
      ++k;
    }
  return OFP_OK;
}

int ofp_ghost_remote_opart(const ofp_ghost_remote* h, const uint32_t** proc, const uint64_t** id, size_t* n) {
  OFP_CHECK(h && proc && id && n, OFP_E_ARG, "ofp_ghost_remote_opart: null argument");
  OFP_CHECK(h->labelled, OFP_E_STATE, "ofp_ghost_remote_opart: no label yet");
  *proc = h->sproc.as<uint32_t>();
  *id = reinterpret_cast<const uint64_t*>(h->sid.p);
  *n = h->n_out;
  return OFP_OK;
}

int ofp_ghost_remote_pack(const ofp_ghost_remote* h, const void* pos, size_t stride, void* send_pos,
                          size_t send_stride, const ofp_plane* planes_in, const ofp_plane* planes_out, int n_planes,
                          void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_ghost_remote_pack: null handle");
  OFP_CHECK(h->labelled, OFP_E_STATE, "ofp_ghost_remote_pack: no label yet");
  OFP_CHECK(h->n_out == 0 || (pos && send_pos), OFP_E_ARG, "ofp_ghost_remote_pack: null positions");
  OFP_CHECK(send_stride >= h->n_out, OFP_E_ARG, "ofp_ghost_remote_pack: send stride below the output count");
  PlaneArgs pa;
  OFP_TRY(make_plane_args(planes_in, planes_out, n_planes, &pa));
  OFP_CHECK(n_planes == 0 || planes_out, OFP_E_ARG, "ofp_ghost_remote_pack: null output planes");
  if (h->n_out == 0) return OFP_OK;
  hipStream_t s = as_stream(stream);
  const uint32_t m = uint32_t(h->n_out);
  const unsigned long long* sid = h->sid.as<unsigned long long>();
  if (h->dtype == OFP_F32)
    k_remote_pack<float><<<grid_for(m, 256), 256, 0, s>>>(static_cast<const float*>(pos), stride, sid, m,
                                                          h->shifts.as<float>(), static_cast<float*>(send_pos),
                                                          send_stride, pa);
  else
    k_remote_pack<double><<<grid_for(m, 256), 256, 0, s>>>(static_cast<const double*>(pos), stride, sid, m,
                                                           h->shifts.as<double>(), static_cast<double*>(send_pos),
                                                           send_stride, pa);
  OFP_LAUNCHED();
  return OFP_OK;
}

}  // extern "C"
