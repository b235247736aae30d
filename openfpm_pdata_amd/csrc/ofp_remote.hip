// ofp_remote.hip -- the remote part of ghost_get on one rank: label every
// particle against the internal ghost boxes that other ranks need, order the
// (processor, particle, shift) outputs by processor, and pack the send buffer.
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
  h->counts.assign(n_slots, 0);
  if (m > 0) {
    k_remote_emit<T><<<grid_for(n, 256), 256, 0, s>>>(B, pos, stride, uint32_t(n), h->starts.as<uint32_t>(),
                                                       h->oslot.as<uint32_t>(), h->oproc.as<uint32_t>(),
                                                       h->oid.as<unsigned long long>());
    OFP_LAUNCHED();
    const uint32_t n_tiles = (m + uint32_t(kRemoteTile) - 1) / uint32_t(kRemoteTile);
    const size_t tn = size_t(n_slots) * n_tiles;
    OFP_TRY(h->table.ensure((tn + 1) * sizeof(uint32_t)));
    OFP_TRY(h->toff.ensure((tn + 1) * sizeof(uint32_t)));
    OFP_TRY(h->tmp.ensure(scan_tmp_bytes(tn)));
    OFP_TRY(h->pcount.ensure((size_t(n_slots) + 1) * sizeof(uint32_t)));
    const unsigned grid = grid_for(n_tiles, 256 / kWave);
    k_remote_order<false><<<grid, 256, 0, s>>>(h->oslot.as<uint32_t>(), nullptr, nullptr, m, n_slots, n_tiles,
                                               h->table.as<uint32_t>(), nullptr, nullptr, nullptr);
    OFP_LAUNCHED();
    OFP_TRY(scan_excl_u32(h->table.as<uint32_t>(), h->toff.as<uint32_t>(), tn, h->tmp.p, s));
    k_remote_order<true><<<grid, 256, 0, s>>>(h->oslot.as<uint32_t>(), h->oproc.as<uint32_t>(),
                                              h->oid.as<unsigned long long>(), m, n_slots, n_tiles, nullptr,
                                              h->toff.as<uint32_t>(), h->sproc.as<uint32_t>(),
                                              h->sid.as<unsigned long long>());
    OFP_LAUNCHED();
    k_remote_slot_starts<<<grid_for(n_slots + 1, 256), 256, 0, s>>>(h->toff.as<uint32_t>(), n_slots, n_tiles,
                                                                   h->pcount.as<uint32_t>());
    OFP_LAUNCHED();
    std::vector<uint32_t> st(n_slots + 1);
    OFP_HIP(hipMemcpyAsync(st.data(), h->pcount.as<uint32_t>(), (n_slots + 1) * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < n_slots; ++k) h->counts[k] = st[k + 1] - st[k];
  }
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

}  // namespace ofp

using namespace ofp;

extern "C" {

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
