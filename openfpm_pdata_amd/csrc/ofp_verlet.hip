// ofp_verlet.hip -- getVerlet on the GPU: non-symmetric Verlet lists (self
// included) from a built cell list, in two passes (count, fill) over target
// items pulled from a device-side work queue.
//
// A target item is up to 64 domain particles (one per lane) whose cells lie on
// one x-row of the grid; its candidates are the (2B+1)^2 x-runs of the stencil
// (B = box_nn), each a contiguous range of sorted indices.  Candidates are
// loaded 64 at a time (coalesced) and broadcast lane by lane with v_readlane,
// so every lane tests its own target against every candidate with the
// reference's rounding (distance2 unfused, in the position dtype).  Entries of
// a row therefore come out in ascending sorted index, which is the reference
// traversal order (cells in ascending linear id, ascending index inside a cell).
// Persistent waves: each wave takes the next item from an atomic counter and
// leaves when the queue is empty, so every wave reaches its exit.
#include "ofp_internal.hpp"

struct ofp_vl {
  ofp::DevBuf cnt, off, nbr, scantmp;
  uint64_t n_rows = 0, n_entries = 0;
  bool built = false;
  const ofp_cl* cl = nullptr;
};

namespace ofp {

constexpr int kVlBlock = 256;

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
  if (sys) {
    const uint32_t s = ti.ncx >> 8;
    g.d0 = sys[s].divc[0];
    g.d1 = sys[s].divc[1];
    g.d2 = sys[s].divc[2];
    g.cbase = sys[s].cbase;
    g.r2 = r2_sys[s];
  } else {
    g.d0 = c.divc[0];
    g.d1 = c.divc[1];
    g.d2 = c.divc[2];
    g.cbase = 0;
    g.r2 = r2;
  }
  const uint32_t cl = ti.cell0 - g.cbase;
  g.cx0 = int(cl % g.d0);
  g.cy = int((cl / g.d0) % g.d1);
  g.cz = int(cl / (g.d0 * g.d1));
  return g;
}

// 64 consecutive candidates (sorted indices j0 + lane, lanes at or past jb
// hold 0), one per lane.
template <class T>
struct Chunk {
  T x, y, z;
};

template <class T>
__device__ __forceinline__ Chunk<T> load_chunk(const T* __restrict__ pos, size_t stride, uint32_t j0, uint32_t jb,
                                               int lane) {
  const uint32_t j = j0 + uint32_t(lane);
  const bool jv = j < jb;
  Chunk<T> q;
  q.x = jv ? pos[j] : T(0);
  q.y = jv ? pos[stride + j] : T(0);
  q.z = jv ? pos[2 * stride + j] : T(0);
  return q;
}

// Walks the item's (2B+1)^2 x-runs chunk by chunk, the next chunk of a run
// loaded while the current one is tested; fn(chunk, j0, m) with m <= 64 valid
// candidates.  Run bounds are wave-uniform.
template <class T, class F>
__device__ __forceinline__ void walk_item_chunks(const ClDev& c, const ItemGeom<T>& G, const T* __restrict__ pos,
                                                 size_t stride, int lane, F&& fn) {
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
      if (ja >= jb) continue;
      Chunk<T> cur = load_chunk(pos, stride, ja, jb, lane);
      for (uint32_t j0 = ja; j0 < jb; j0 += kWave) {
        const uint32_t jn = j0 + kWave;
        Chunk<T> nxt = cur;
        if (jn < jb) nxt = load_chunk(pos, stride, jn, jb, lane);
        const int m = int(jb - j0 < uint32_t(kWave) ? jb - j0 : uint32_t(kWave));
        fn(cur, j0, m);
        cur = nxt;
      }
    }
  }
}

// Count pass with the targets in the lanes: every candidate is broadcast with
// v_readlane and tested by all targets of the item.
template <class T>
__global__ __launch_bounds__(kVlBlock) void k_vl_count(ClDev c, const T* __restrict__ pos, size_t stride,
                                                       T r2_one, const SysGrid<T>* __restrict__ sys,
                                                       const T* __restrict__ r2_sys,
                                                       const TargetItem* __restrict__ items,
                                                       const uint32_t* __restrict__ n_items_p,
                                                       uint32_t* __restrict__ counter,
                                                       uint32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t n_items = *n_items_p;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const TargetItem ti = items[it];
    const ItemGeom<T> G = item_geom(c, ti, r2_one, sys, r2_sys);
    const T r2 = G.r2;
    const bool valid = uint32_t(lane) < ti.n;
    const uint32_t row = ti.row0 + (valid ? uint32_t(lane) : 0u);
    const uint32_t p = c.dom[row];
    const T xp = pos[p], yp = pos[stride + p], zp = pos[2 * stride + p];
    uint32_t n = 0;
    auto test = [&](const Chunk<T>& q, int k) {
      const T xk = bcast(q.x, k), yk = bcast(q.y, k), zk = bcast(q.z, k);
      const T dd = dist2_unfused<T>(xk - xp, yk - yp, zk - zp);
      n += (dd < r2) ? 1u : 0u;
    };
    walk_item_chunks(c, G, pos, stride, lane, [&](const Chunk<T>& q, uint32_t, int m) {
      // unrolled by hand: v_readlane is convergent, so the compiler does not
      // unroll a loop with a run-time trip count around it
      int k = 0;
      for (; k + 4 <= m; k += 4) {
        test(q, k);
        test(q, k + 1);
        test(q, k + 2);
        test(q, k + 3);
      }
      for (; k < m; ++k) test(q, k);
    });
    if (valid) cnt[row] = n;
  }
}

// Fill pass with the candidates in the lanes: for each target t of the item
// (uniform loop) every lane tests its candidate against t; the hits are written
// to row t at base + (number of hitting lanes below), one contiguous store per
// (chunk, target).  Lane t keeps row t's write cursor.  Same test, same order as
// the count pass, so each row ends exactly where the count said.
// e_total / e_cap: the list's entry count (device) and the capacity of nbr; a
// list that does not fit is not written (the caller sees e_total > e_cap).
template <class T>
__global__ __launch_bounds__(kVlBlock) void k_vl_fill(ClDev c, const T* __restrict__ pos, size_t stride, T r2_one,
                                                      const SysGrid<T>* __restrict__ sys,
                                                      const T* __restrict__ r2_sys,
                                                      const TargetItem* __restrict__ items,
                                                      const uint32_t* __restrict__ n_items_p,
                                                      uint32_t* __restrict__ counter,
                                                      const unsigned long long* __restrict__ off,
                                                      const unsigned long long* __restrict__ e_total,
                                                      unsigned long long e_cap,
                                                      uint32_t* __restrict__ nbr) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t n_items = *n_items_p;
  if (*e_total > e_cap) return;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const TargetItem ti = items[it];
    const ItemGeom<T> G = item_geom(c, ti, r2_one, sys, r2_sys);
    const T r2 = G.r2;
    const int nt = int(ti.n);
    const bool valid = lane < nt;
    const uint32_t row = ti.row0 + (valid ? uint32_t(lane) : 0u);
    const uint32_t p = c.dom[row];
    const T xp = pos[p], yp = pos[stride + p], zp = pos[2 * stride + p];
    unsigned long long wpos = off[row];
    walk_item_chunks(c, G, pos, stride, lane, [&](const Chunk<T>& q, uint32_t j0, int m) {
      const uint32_t j = j0 + uint32_t(lane);
      const bool jv = lane < m;
      for (int t = 0; t < nt; ++t) {
        const T xt = bcast(xp, t), yt = bcast(yp, t), zt = bcast(zp, t);
        const T dd = dist2_unfused<T>(q.x - xt, q.y - yt, q.z - zt);
        const bool hit = jv && dd < r2;
        const unsigned long long mask = __ballot(hit);
        if (mask) {
          const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(wpos), t);
          const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(wpos >> 32), t);
          const unsigned long long base = (static_cast<unsigned long long>(hi) << 32) | lo;
          const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                           __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
          if (hit) nbr[base + below] = j;
          if (lane == t) wpos += uint32_t(__popcll(mask));
        }
      }
    });
  }
}

static unsigned persistent_grid() {
  static unsigned g = 0;
  if (!g) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    g = unsigned(ncu) * 8;  // 8 workgroups of 4 waves per CU
  }
  return g;
}

template <class T>
static int vl_build(ofp_vl* h, const ofp_cl* cl, const void* pos, size_t stride, double r_cut,
                    hipStream_t s) {
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(ofp_cl_dev_(cl, &c, &v));
  for (int k = 0; k < 3; ++k)
    OFP_CHECK(v.spacing[k] * c.box_nn >= r_cut * (1.0 - 1e-6), OFP_E_ARG,
              "stencil %u x spacing %g is smaller than the Verlet radius %g", c.box_nn,
              v.spacing[k], r_cut);
  const size_t n = c.n_dom;
  OFP_TRY(h->cnt.ensure((n + 1) * 4));
  OFP_TRY(h->off.ensure((n + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(n)));
  const TargetItem* items;
  const uint32_t* n_items;
  uint32_t* counter;
  size_t cap;
  OFP_TRY(ofp_cl_items_(cl, &items, &n_items, &counter, &cap, s));
  const T rc = T(r_cut);
  const T r2 = rc * rc;
  const unsigned grid = persistent_grid();
  if (n > 0) {
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_count<T><<<grid, kVlBlock, 0, s>>>(c, (const T*)pos, stride, r2, nullptr, nullptr, items, n_items, counter,
                                            h->cnt.as<uint32_t>());
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
  uint64_t E = 0;
  OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  OFP_TRY(h->nbr.ensure((E + 1) * 4));
  if (n > 0) {
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_fill<T><<<grid, kVlBlock, 0, s>>>(c, (const T*)pos, stride, r2, nullptr, nullptr, items, n_items, counter,
                                           h->off.as<unsigned long long>(), h->off.as<unsigned long long>() + n, E,
                                           h->nbr.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_rows = n;
  h->n_entries = E;
  h->built = true;
  h->cl = cl;
  return OFP_OK;
}

// Ensemble build, with no host synchronisation (capturable): count, scan, and a
// fill that writes only when the entry count fits e_cap.  nbr must already hold
// e_cap entries.
template <class T>
static int vl_build_many_t(ofp_vl* h, const ofp_cl* cl, const void* pos, size_t stride, const void* sys,
                           const void* r2_sys, size_t* e_cap_io, double headroom, int sizing, unsigned grid,
                           uint64_t* e_count, hipStream_t s) {
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(ofp_cl_dev_(cl, &c, &v));
  const size_t n = c.n_dom;
  OFP_TRY(h->cnt.ensure((n + 1) * 4));
  OFP_TRY(h->off.ensure((n + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(n)));
  const TargetItem* items;
  const uint32_t* n_items;
  uint32_t* counter;
  size_t cap;
  OFP_TRY(ofp_cl_items_(cl, &items, &n_items, &counter, &cap, s));
  if (grid == 0) grid = persistent_grid();
  const SysGrid<T>* sg = static_cast<const SysGrid<T>*>(sys);
  const T* r2s = static_cast<const T*>(r2_sys);
  if (n > 0) {
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_count<T><<<grid, kVlBlock, 0, s>>>(c, (const T*)pos, stride, T(0), sg, r2s, items, n_items, counter,
                                            h->cnt.as<uint32_t>());
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
  if (sizing) {
    uint64_t E = 0;
    OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    const size_t want = size_t(double(E) * (headroom > 1.0 ? headroom : 1.0)) + 1024;
    if (want > *e_cap_io) *e_cap_io = want;
    if (e_count) *e_count = E;
  }
  const size_t e_cap = *e_cap_io;
  OFP_TRY(h->nbr.ensure((e_cap + 1) * 4));
  if (n > 0) {
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_fill<T><<<grid, kVlBlock, 0, s>>>(c, (const T*)pos, stride, T(0), sg, r2s, items, n_items, counter,
                                           h->off.as<unsigned long long>(), h->off.as<unsigned long long>() + n,
                                           (unsigned long long)e_cap, h->nbr.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_rows = n;
  h->n_entries = 0;  // on the device: off[n_rows]
  h->built = true;
  h->cl = cl;
  return OFP_OK;
}

}  // namespace ofp

using namespace ofp;

int ofp::vl_build_many(ofp_vl* h, const ofp_cl* cl, int dtype, const void* pos_sorted, size_t stride,
                       const void* sys_dev, const void* r2_sys_dev, size_t* e_cap, double headroom, int sizing,
                       unsigned grid, uint64_t* e_count, hipStream_t s) {
  OFP_CHECK(h && cl && sys_dev && r2_sys_dev && e_cap, OFP_E_ARG, "vl_build_many: null argument");
  if (dtype == OFP_F32)
    return vl_build_many_t<float>(h, cl, pos_sorted, stride, sys_dev, r2_sys_dev, e_cap, headroom, sizing, grid,
                                  e_count, s);
  return vl_build_many_t<double>(h, cl, pos_sorted, stride, sys_dev, r2_sys_dev, e_cap, headroom, sizing, grid,
                                 e_count, s);
}

int ofp::vl_offsets(const ofp_vl* h, const unsigned long long** off) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "Verlet list not built");
  *off = h->off.as<unsigned long long>();
  return OFP_OK;
}

extern "C" {

// accessor for the pair kernels (same library)
int ofp_vl_dev_(const ofp_vl* h, const ofp_cl* cl, const unsigned long long** off,
                const uint32_t** nbr) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "Verlet list not built");
  OFP_CHECK(h->cl == cl, OFP_E_STATE, "Verlet list not built from this cell list");
  *off = h->off.as<unsigned long long>();
  *nbr = h->nbr.as<uint32_t>();
  return OFP_OK;
}

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

}  // extern "C"
