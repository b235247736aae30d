// ofp_verlet.hip -- getVerlet on the GPU: non-symmetric Verlet lists (self
// included) from a built cell list, in two passes over target items pulled
// from a device-side work queue.
//
// A target item is up to 64 domain particles (one per lane) whose cells lie on
// one x-row of the grid; its candidates are the (2B+1)^2 x-runs of the stencil
// (B = box_nn), each a contiguous range of sorted indices.
//  * count: the candidate index is wave-uniform, so the candidates' coordinates
//    come in through scalar loads (8 candidates at a time, from an interleaved
//    xyz copy of the sorted positions) and every lane tests its own
//    target against every candidate with the reference's rounding (distance2
//    unfused, in the position dtype).  Each compare appends one bit to the
//    lane's mask of the chunk (64 candidates): one 64-bit mask per (target,
//    chunk), kept in a region of a mask buffer taken by one atomic per item.
//  * fill: the candidates are in the lanes; a target's mask of a chunk is the
//    exec mask of the chunk's store into its row, and the lanes below each hit
//    give its place, so no distance is tested twice.  A group of up to 8 chunks
//    is read at once (lane t: target t's masks); a target's hits of the whole
//    group are written back to back at its row's cursor, so the row grows by
//    whole cache lines while the wave is on it.
// Entries of a row come out in ascending sorted index, which is the reference
// traversal order (cells in ascending linear id, ascending index inside a cell).
// Persistent waves: each wave takes the next item from an atomic counter and
// leaves when the queue is empty, so every wave reaches its exit.
// A mask buffer too small for the items is flagged (mstate[1]); the fill then
// writes nothing and the host grows the buffer and counts again.
#include "ofp_internal.hpp"

#include <cmath>
#include <type_traits>

struct ofp_vl {
  ofp::DevBuf cnt, off, nbr, scantmp, masks, moff, mstate, xyz;
  ofp::DevBuf rmask;  // thread-per-row build: the rows' hit masks, count pass to fill
  ofp::DevBuf aux;  // pair kernels' per-particle gather stream over the rows
  size_t mask_cap = 0;  // 8-byte masks the buffer holds
  bool rows_path = false;  // the last build ran the thread-per-row kernels
  uint64_t n_rows = 0, n_entries = 0;
  bool built = false;
  bool symmetric = false;  // built by ofp_vl_build_sym (each interaction once)
  const ofp_cl* cl = nullptr;
};

namespace ofp {

constexpr int kVlBlock = 256;
constexpr int kFillGroup = 8;  // chunks of 64 candidates per fill group

// Lane `lane` (uniform) of the returned VGPR takes the uniform value sval; the
// other lanes keep vdst (v_writelane_b32, a VALU op; the lane select goes
// through M0 since gfx9 VOP3 reads one SGPR besides M0).
__device__ __forceinline__ uint32_t write_lane(uint32_t vdst, uint32_t sval, uint32_t lane) {
  asm("s_mov_b32 m0, %2\n\tv_writelane_b32 %0, %1, m0" : "+v"(vdst) : "s"(sval), "s"(lane) : "m0");
  return vdst;
}

__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// The coordinates of 8 consecutive particles, x y z interleaved (read with
// scalar loads; the element alignment is all a scalar load needs).
template <class T>
struct alignas(sizeof(T)) Run8 {
  T v[24];
};

// Sorted positions (three planes of stride `stride`) to one interleaved xyz
// stream: the count pass reads the candidates of a run as one contiguous range,
// whatever the planes' stride (planes a large power of two apart would share
// cache sets).
template <class T>
__global__ void k_vl_pack(const T* __restrict__ pos, size_t stride, uint32_t n, T* __restrict__ xyz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xyz[3 * size_t(i)] = pos[i];
  xyz[3 * size_t(i) + 1] = pos[stride + i];
  xyz[3 * size_t(i) + 2] = pos[2 * stride + i];
}

// One target's masks of a full fill group (64 contiguous bytes).
struct alignas(8) MaskRun {
  uint2 v[kFillGroup];
};

// Count pass with the targets in the lanes.  Lane t builds, per chunk of 64
// candidates, the 64-bit mask of the candidates within r of target t (bit k:
// candidate k of the chunk).  The masks of a fill group (up to kFillGroup
// chunks of one run) are stored target-major, ng per target, so the fill reads
// a group with one contiguous load per lane.  mstate[0]: mask cursor,
// mstate[1]: overflow flag; item_moff[it] = the item's first mask.
template <class T>
__global__ __launch_bounds__(kVlBlock) void k_vl_count(ClDev c, const T* __restrict__ xyz,
                                                       T r2_one, const SysGrid<T>* __restrict__ sys,
                                                       const T* __restrict__ r2_sys,
                                                       const TargetItem* __restrict__ items,
                                                       const uint32_t* __restrict__ n_items_p,
                                                       uint32_t* __restrict__ counter, uint32_t* __restrict__ cnt,
                                                       unsigned long long* __restrict__ mstate,
                                                       unsigned long long* __restrict__ item_moff,
                                                       uint2* __restrict__ masks, unsigned long long mask_cap) {
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
    const T xp = xyz[3 * size_t(p)], yp = xyz[3 * size_t(p) + 1], zp = xyz[3 * size_t(p) + 2];
    // the item's mask region: ti.n masks per chunk of 64 candidates
    uint32_t nch = 0;
    for_each_run(c, G, [&](uint32_t ja, uint32_t jb) { nch += (jb - ja + kWave - 1) / kWave; });
    const unsigned long long need = static_cast<unsigned long long>(nch) * ti.n;
    unsigned long long moff = 0;
    if (lane == 0) moff = atomicAdd(&mstate[0], need);
    moff = uniform_u64(moff);
    const bool keep = moff + need <= mask_cap;
    if (lane == 0) {
      item_moff[it] = moff;
      if (!keep) atomicOr(&mstate[1], 1ull);
    }
    // Hit bits of candidates j0 .. j0 + cnt - 1 (1 <= cnt <= 32, uniform), bit k
    // for candidate j0 + k.  The candidate index is wave-uniform: coordinates come
    // in through scalar loads (read-only data), 8 candidates at a time, and enter
    // the VALU as SGPR operands.  Lanes past the item's targets compute bits that
    // are never stored.
    auto hit_word = [&](uint32_t j0, uint32_t cnt) -> uint32_t {
      uint32_t w = 0;  // candidate j0 + k at bit (k' - 1 - k) after k' shifts
      uint32_t k = 0;
      for (; k + 8 <= cnt; k += 8) {
        const Run8<T> R = *reinterpret_cast<const Run8<T>*>(xyz + 3 * size_t(j0 + k));
#pragma unroll
        for (int i = 0; i < 8; ++i)
          w = shift_in(w, dist2_unfused<T>(R.v[3 * i] - xp, R.v[3 * i + 1] - yp, R.v[3 * i + 2] - zp) < r2);
      }
      for (; k < cnt; ++k) {
        const T* q = xyz + 3 * size_t(j0 + k);
        w = shift_in(w, dist2_unfused<T>(q[0] - xp, q[1] - yp, q[2] - zp) < r2);
      }
      return __builtin_bitreverse32(w << (32u - cnt));
    };
    uint32_t n = 0;
    unsigned long long goff = moff;  // the group's first mask
    for_each_run(c, G, [&](uint32_t ja, uint32_t jb) {
      for (uint32_t g0 = ja; g0 < jb; g0 += kFillGroup * kWave) {
        const uint32_t rem = jb - g0 < uint32_t(kFillGroup * kWave) ? jb - g0 : uint32_t(kFillGroup * kWave);
        const uint32_t ng = (rem + kWave - 1) / kWave;  // chunks in the group
        for (uint32_t i = 0; i < ng; ++i) {
          const uint32_t m = rem - i * kWave < uint32_t(kWave) ? rem - i * kWave : uint32_t(kWave);
          const uint32_t lo = hit_word(g0 + i * kWave, m < 32u ? m : 32u);
          const uint32_t hi = m > 32u ? hit_word(g0 + i * kWave + 32u, m - 32u) : 0u;
          n += uint32_t(__builtin_popcount(lo)) + uint32_t(__builtin_popcount(hi));
          if (keep && valid) masks[goff + static_cast<unsigned long long>(lane) * ng + i] = make_uint2(lo, hi);
        }
        goff += static_cast<unsigned long long>(ng) * ti.n;
      }
    });
    if (valid) cnt[row] = n;
  }
}

// Fill pass from the masks, candidates in the lanes: for each target t of the
// item (uniform loop) and chunk, the lanes set in t's mask hit it; they are
// written to row t at base + (number of mask bits below the lane), one
// contiguous store per (chunk, target); lane t keeps row t's write cursor.  Same
// runs, same order as the count pass, so each row ends exactly where the count
// said.  e_total / e_cap: the list's entry count (device) and the capacity of
// nbr; a list that does not fit, or whose masks did not, is not written.
template <class T>
__global__ __launch_bounds__(kVlBlock) void k_vl_fill(ClDev c, const SysGrid<T>* __restrict__ sys,
                                                      const TargetItem* __restrict__ items,
                                                      const uint32_t* __restrict__ n_items_p,
                                                      uint32_t* __restrict__ counter,
                                                      const unsigned long long* __restrict__ off,
                                                      const unsigned long long* __restrict__ e_total,
                                                      unsigned long long e_cap,
                                                      const unsigned long long* __restrict__ mstate,
                                                      const unsigned long long* __restrict__ item_moff,
                                                      const uint2* __restrict__ masks, uint32_t* __restrict__ nbr) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t n_items = *n_items_p;
  if (*e_total > e_cap || mstate[1] != 0ull) return;
  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const TargetItem ti = items[it];
    const ItemGeom<T> G = item_geom(c, ti, T(0), sys, static_cast<const T*>(nullptr));
    const int nt = int(ti.n);
    const bool valid = lane < nt;
    const uint32_t row = ti.row0 + (valid ? uint32_t(lane) : 0u);
    // lane t: row t's write cursor (low, high word)
    const unsigned long long w0 = off[row];
    uint32_t wlo = uint32_t(w0), whi = uint32_t(w0 >> 32);
    unsigned long long goff = uniform_u64(item_moff[it]);  // the group's first mask
    // groups of up to kFillGroup chunks of one run: each target's hits in the
    // group are written back to back (its row grows by whole cache lines while
    // the wave is on it), one cursor update per target and group
    for_each_run(c, G, [&](uint32_t ja, uint32_t jb) {
      for (uint32_t g0 = ja; g0 < jb; g0 += kFillGroup * kWave) {
        const uint32_t rem = jb - g0 < uint32_t(kFillGroup * kWave) ? jb - g0 : uint32_t(kFillGroup * kWave);
        const int ng = int((rem + kWave - 1) / kWave);  // chunks in the group (uniform)
        // Lane t holds target t's masks of the group (ng of them, stored back to
        // back); chunk i's mask of target t, read out of lane t, is the exec mask
        // of its store: the lanes (the candidates) below each hit give the
        // entry's place.  Two targets per step, t and t + 32, whose chains are
        // independent (their instructions interleave).  full: the group has all
        // kFillGroup chunks.  A row's write position is a uniform pointer plus a
        // byte offset below 4 KiB (at most kFillGroup * 64 entries per target and
        // group).
        uint2 mk[kFillGroup];
        if (ng == kFillGroup) {
          const MaskRun M = valid ? *reinterpret_cast<const MaskRun*>(masks + goff + uint32_t(lane) * uint32_t(kFillGroup))
                                  : MaskRun{};
#pragma unroll
          for (int i = 0; i < kFillGroup; ++i) mk[i] = M.v[i];
        } else {
#pragma unroll
          for (int i = 0; i < kFillGroup; ++i)
            mk[i] = valid && i < ng ? masks[goff + uint32_t(lane) * uint32_t(ng) + uint32_t(i)] : make_uint2(0u, 0u);
        }
        auto targets = [&](auto full, auto two, int t) {
          constexpr bool kTwo = decltype(two)::value;
          const int u = t + 32;
          const uint32_t lo0 = __builtin_amdgcn_readlane(wlo, t), hi0 = __builtin_amdgcn_readlane(whi, t);
          const unsigned long long base0 = (static_cast<unsigned long long>(hi0) << 32) | lo0;
          unsigned long long base1 = 0;
          if constexpr (kTwo) {
            const uint32_t lo1 = __builtin_amdgcn_readlane(wlo, u), hi1 = __builtin_amdgcn_readlane(whi, u);
            base1 = (static_cast<unsigned long long>(hi1) << 32) | lo1;
          }
          char* const row0 = reinterpret_cast<char*>(nbr + base0);
          char* const row1 = reinterpret_cast<char*>(nbr + base1);
          uint32_t run0 = 0, run1 = 0;  // entries written for the targets in the group
#pragma unroll
          for (int i = 0; i < kFillGroup; ++i) {
            if constexpr (!decltype(full)::value) {
              if (i >= ng) break;
            }
            const uint32_t j = g0 + uint32_t(i) * kWave + uint32_t(lane);
            const uint32_t m0l = __builtin_amdgcn_readlane(mk[i].x, t), m0h = __builtin_amdgcn_readlane(mk[i].y, t);
            const unsigned long long m0 = (static_cast<unsigned long long>(m0h) << 32) | m0l;
            const uint32_t b0 = __builtin_amdgcn_mbcnt_hi(m0h, __builtin_amdgcn_mbcnt_lo(m0l, 0u));
            if (__builtin_amdgcn_inverse_ballot_w64(m0)) *reinterpret_cast<uint32_t*>(row0 + (run0 + b0) * 4u) = j;
            run0 += uint32_t(__popcll(m0));
            if constexpr (kTwo) {
              const uint32_t m1l = __builtin_amdgcn_readlane(mk[i].x, u), m1h = __builtin_amdgcn_readlane(mk[i].y, u);
              const unsigned long long m1 = (static_cast<unsigned long long>(m1h) << 32) | m1l;
              const uint32_t b1 = __builtin_amdgcn_mbcnt_hi(m1h, __builtin_amdgcn_mbcnt_lo(m1l, 0u));
              if (__builtin_amdgcn_inverse_ballot_w64(m1)) *reinterpret_cast<uint32_t*>(row1 + (run1 + b1) * 4u) = j;
              run1 += uint32_t(__popcll(m1));
            }
          }
          if (run0) {
            const unsigned long long nb = base0 + run0;
            wlo = write_lane(wlo, uint32_t(nb), uint32_t(t));
            whi = write_lane(whi, uint32_t(nb >> 32), uint32_t(t));
          }
          if (kTwo && run1) {
            const unsigned long long nb = base1 + run1;
            wlo = write_lane(wlo, uint32_t(nb), uint32_t(u));
            whi = write_lane(whi, uint32_t(nb >> 32), uint32_t(u));
          }
        };
        auto all_targets = [&](auto full) {
          const int pairs = nt > 32 ? nt - 32 : 0;  // t with a partner t + 32
          for (int t = 0; t < pairs; ++t) targets(full, std::true_type{}, t);
          for (int t = pairs; t < (nt < 32 ? nt : 32); ++t) targets(full, std::false_type{}, t);
        };
        if (ng == kFillGroup)
          all_targets(std::true_type{});
        else
          all_targets(std::false_type{});
        goff += static_cast<unsigned long long>(ng) * uint32_t(nt);
      }
    });
  }
}

// Thread-per-row build for dilute lists (few candidates per target: the item
// kernels' 64 lanes would mostly idle).  Thread k takes row k (domain particle
// dom[k]), finds its cell with the build's key formula, and walks the
// (2B+1)^2 x-runs of its own stencil in traversal order.
//  * count (kFill = false): every candidate is tested with the same rounding
//    as the item kernels (distance2 unfused, in the position dtype); each test
//    appends one bit to the row's hit mask (candidate c of the row is bit
//    31 - c % 32 of word c / 32), staged in LDS and stored for the row's first
//    kRowWords words (word w of the wave's rows at one 256-byte line), and the
//    hits are counted.
//  * fill (kFill = true): the row's entries come from its mask words (the
//    tests of candidates past kRowWords words are repeated), in ascending
//    sorted index as the item kernels write them, gathered in LDS up to
//    kRowSeg of them per wave and written out with coalesced stores; past
//    kRowSeg they go straight to the list.
// Ensembles: the row's system from sid_sorted.
constexpr int kRowAhead = 4;  // candidates in flight per thread
constexpr uint32_t kRowWords = 16;  // mask words kept per row (32 candidates each)
constexpr uint32_t kRowSeg = 1024;  // entries per wave staged in LDS
constexpr int kRowWaves = 256 / kWave;

// kB: the stencil half-width B (c.box_nn) as a constant, so the row's
// (2B+1)^2 run bounds are all loaded before its first candidate is read.
// masks: kRowWords x 64 words per wave of rows.
template <class T, bool kFill, int kB>
__global__ __launch_bounds__(256) void k_vl_rows(ClDev c, CellGeom<T> g1, const T* __restrict__ xyz, T r2_one,
                                                 const SysGrid<T>* __restrict__ sys, const T* __restrict__ r2_sys,
                                                 const uint32_t* __restrict__ sid_sorted, uint32_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ masks,
                                                 const unsigned long long* __restrict__ off,
                                                 const unsigned long long* __restrict__ e_total,
                                                 unsigned long long e_cap, uint32_t* __restrict__ nbr) {
  __shared__ uint32_t lds[kRowWaves][kFill ? kRowSeg : kRowWords * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k0 = k - uint32_t(lane);  // the wave's first row
  if (k0 >= c.n_dom) return;               // wave-uniform
  if constexpr (kFill) {
    if (*e_total > e_cap) return;  // the list does not fit: nothing is written (grid-uniform)
  }
  const bool valid = k < c.n_dom;
  uint32_t* const wl = lds[threadIdx.x / kWave];
  uint32_t* const wmask = masks + size_t(k0 / uint32_t(kWave)) * (kRowWords * kWave);
  T xp = T(0), yp = T(0), zp = T(0), r2 = r2_one;
  CellGeom<T> g = g1;
  uint32_t cbase = 0;
  int cxi = 0, cyi = 0, czi = 0;
  if (valid) {
    const uint32_t p = c.dom[k];
    xp = xyz[3 * size_t(p)];
    yp = xyz[3 * size_t(p) + 1];
    zp = xyz[3 * size_t(p) + 2];
    if (sys) {
      const uint32_t s = sid_sorted[p];
      g = sys[s].geom();
      cbase = sys[s].cbase;
      if (r2_sys) r2 = r2_sys[s];
    }
    uint32_t cx, cy, cz;
    (void)cell_of(g, xp, yp, zp, &cx, &cy, &cz);
    cxi = int(cx);
    cyi = int(cy);
    czi = int(cz);
  }
  const int d0 = int(g.divc[0]), d1 = int(g.divc[1]), d2 = int(g.divc[2]);
  const int xa = cxi - kB > 0 ? cxi - kB : 0;
  const int xb = cxi + kB < d0 - 1 ? cxi + kB : d0 - 1;
  // the stencil's x-runs [a, b) of sorted indices in traversal order (dz, then
  // dy); runs off the grid (and every run of an invalid lane) are empty
  auto run_bounds = [&](int r, uint32_t* a, uint32_t* b) {
    const int zz = czi + r / (2 * kB + 1) - kB, yy = cyi + r % (2 * kB + 1) - kB;
    const bool in = valid && zz >= 0 && zz < d2 && yy >= 0 && yy < d1;
    const uint32_t rowc = cbase + uint32_t(d0) * (uint32_t(in ? yy : 0) + uint32_t(d1) * uint32_t(in ? zz : 0));
    *a = in ? uint32_t(c.cs[rowc + uint32_t(xa)]) : 0u;
    *b = in ? uint32_t(c.cs[rowc + uint32_t(xb) + 1]) : 0u;
  };
  constexpr int kRuns = (2 * kB + 1) * (2 * kB + 1);
  constexpr bool kKeep = kB == 1;  // the 9 runs' bounds loaded up front, kept in registers
  uint32_t run_a[kKeep ? kRuns : 1], run_b[kKeep ? kRuns : 1];
  if constexpr (kKeep) {
#pragma unroll
    for (int r = 0; r < kRuns; ++r) run_bounds(r, &run_a[r], &run_b[r]);
  }
  auto for_runs = [&](auto&& fn) {
    if constexpr (kKeep) {
#pragma unroll
      for (int r = 0; r < kRuns; ++r) fn(run_a[r], run_b[r]);
    } else {
      for (int r = 0; r < kRuns; ++r) {
        uint32_t a, b;
        run_bounds(r, &a, &b);
        fn(a, b);
      }
    }
  };
  if constexpr (!kFill) {
    // the candidates in groups of kRowAhead from each run's start, all of a
    // group's coordinates loaded before any is tested (up to kRowAhead - 1
    // candidates past the run's end are read, from the next run or the
    // coordinate stream's padding; their tests are dropped).  A group appends its m tests to the
    // pending bits (acc, the newest lowest; fill of them not yet stored), and
    // every 32 of them go to the next mask word.
    uint32_t n = 0, nwd = 0, fill = 0;  // hits, mask words completed, bits pending
    unsigned long long acc = 0;
    for_runs([&](uint32_t ja, uint32_t jb) {
      for (uint32_t j0 = ja; j0 < jb; j0 += kRowAhead) {
        const T* __restrict__ q = xyz + 3 * size_t(j0);
        T c[kRowAhead][3];
#pragma unroll
        for (int u = 0; u < kRowAhead; ++u) {
          c[u][0] = q[3 * u];
          c[u][1] = q[3 * u + 1];
          c[u][2] = q[3 * u + 2];
        }
        const uint32_t m = jb - j0 < uint32_t(kRowAhead) ? jb - j0 : uint32_t(kRowAhead);
        uint32_t h = 0;  // the group's tests, its first candidate highest
#pragma unroll
        for (int u = 0; u < kRowAhead; ++u) {
          const bool hit = dist2_unfused<T>(c[u][0] - xp, c[u][1] - yp, c[u][2] - zp) < r2;
          h |= (hit && uint32_t(u) < m ? 1u : 0u) << (kRowAhead - 1 - u);
        }
        h >>= uint32_t(kRowAhead) - m;
        n += uint32_t(__popc(h));
        acc = (acc << m) | h;
        fill += m;
        if (fill >= 32u) {
          fill -= 32u;
          if (nwd < kRowWords) wl[nwd * kWave + uint32_t(lane)] = uint32_t(acc >> fill);
          ++nwd;
        }
      }
    });
    // the last, partial word, its first candidate at bit 31
    if (fill != 0u && nwd < kRowWords) wl[nwd * kWave + uint32_t(lane)] = uint32_t(acc << (32u - fill));
    const uint32_t cand = nwd * 32u + fill;
    if (valid) cnt[k] = n;
    // the mask words out, word w of the wave's rows on one line (a lane past
    // its own words stores what is there; nothing reads it)
    uint32_t nw = (cand + 31u) >> 5;
    nw = nw < kRowWords ? nw : kRowWords;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      const uint32_t o = __shfl_xor(nw, d, kWave);
      nw = o > nw ? o : nw;
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t w = 0; w < nw; ++w) wmask[w * kWave + uint32_t(lane)] = wl[w * kWave + uint32_t(lane)];
  } else {
    const unsigned long long base = off[k0];
    const unsigned long long lo = valid ? off[k] - base : 0;
    uint32_t i = 0;  // entries of the row written
    auto emit = [&](uint32_t j) {
      const unsigned long long at = lo + i;
      if (at < kRowSeg)
        wl[at] = j;
      else
        nbr[base + at] = j;
      ++i;
    };
    uint32_t cb = 0;              // candidate index of the run's first candidate
    uint32_t cw = ~0u, cbits = 0;  // the mask word last read
    for_runs([&](uint32_t ja, uint32_t jb) {
      const uint32_t end = cb + (jb - ja);
      const uint32_t lim = end < 32u * kRowWords ? end : 32u * kRowWords;
      uint32_t cc = cb;
      while (cc < lim) {
        const uint32_t w = cc >> 5, sh = cc & 31u;
        const uint32_t wend = (w + 1u) * 32u < lim ? (w + 1u) * 32u : lim;
        const uint32_t m = wend - cc;  // the run's candidates in word w (1-32)
        if (w != cw) {
          cw = w;
          cbits = wmask[w * kWave + uint32_t(lane)];
        }
        uint32_t bm = cbits << sh;
        if (m < 32u) bm &= ~(0xFFFFFFFFu >> m);
        while (bm) {
          const uint32_t lead = uint32_t(__builtin_clz(bm));
          bm ^= 0x80000000u >> lead;
          emit(ja + (cc - cb) + lead);
        }
        cc = wend;
      }
      for (uint32_t j = ja + (cc - cb); j < jb; ++j)
        if (dist2_unfused<T>(xyz[3 * size_t(j)] - xp, xyz[3 * size_t(j) + 1] - yp, xyz[3 * size_t(j) + 2] - zp) < r2)
          emit(j);
      cb = end;
    });
    // the wave's staged entries, written out by all its lanes (the wave's LDS
    // writes all precede these reads in its program order)
    __builtin_amdgcn_wave_barrier();
    const uint32_t k1 = k0 + uint32_t(kWave) < c.n_dom ? k0 + uint32_t(kWave) : c.n_dom;
    const unsigned long long len = off[k1] - base;
    const uint32_t ms = len < kRowSeg ? uint32_t(len) : kRowSeg;
    for (uint32_t e = uint32_t(lane); e < ms; e += kWave) nbr[base + e] = wl[e];
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

// One build.  sys/r2_sys: ensemble tables (null: one grid, radius r2_one).
// sync: the counts are read back (entries, mask total), nbr and the mask
// buffer grow to fit (the count pass is repeated if the masks did not fit), and
// e_cap_io becomes max(e_cap_io, E x headroom); without sync nothing is read
// back and the capacities stay (a list that would not fit is flagged).
// rows_mode: 1 = the item kernels (targets in the lanes), 2 = thread per row,
// 0 = thread per row when the stencil holds fewer than a wave of candidates per
// target on average (n_part (2B+1)^3 < 64 n_cells), the item kernels otherwise.
template <class T>
static int vl_build_t(ofp_vl* h, const ofp_cl* cl, const void* pos_v, size_t stride, T r2_one, const void* sys_v,
                      const void* r2_sys_v, bool sync, size_t* e_cap_io, double headroom, unsigned grid,
                      uint64_t* e_count, hipStream_t s, int rows_mode = 0) {
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(ofp_cl_dev_(cl, &c, &v));
  const T* pos = static_cast<const T*>(pos_v);
  const SysGrid<T>* sys = static_cast<const SysGrid<T>*>(sys_v);
  const T* r2_sys = static_cast<const T*>(r2_sys_v);
  const size_t n = c.n_dom;
  OFP_TRY(h->cnt.ensure((n + 1) * 4));
  OFP_TRY(h->off.ensure((n + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(n)));
  OFP_TRY(h->moff.ensure((n + 1) * 8));
  OFP_TRY(h->mstate.ensure(16));
  OFP_TRY(h->xyz.ensure((v.n_part + kRowAhead) * 3 * sizeof(T)));  // the row count reads whole groups
  const double stencil = std::pow(2.0 * double(c.box_nn) + 1.0, 3.0);
  const bool rows = rows_mode == 2 || (rows_mode == 0 && double(v.n_part) * stencil < 64.0 * double(v.n_cells));
  if (rows) {
    auto* mstate = h->mstate.as<unsigned long long>();
    OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(mstate), 0, 2, s));  // no masks: nothing to overflow
    if (v.n_part > 0) {
      k_vl_pack<T><<<grid_for(v.n_part, 256), 256, 0, s>>>(pos, stride, uint32_t(v.n_part), h->xyz.as<T>());
      OFP_LAUNCHED();
    }
    CellGeom<T> g1;
    for (int k = 0; k < 3; ++k) {
      g1.lo[k] = T(v.lo[k]);
      g1.sp[k] = T(v.spacing[k]);
      g1.divc[k] = v.divc[k];
    }
    const uint32_t* sid = nullptr;
    if (sys) OFP_TRY(cl_sid_sorted(cl, &sid));
    const unsigned g = grid_for(n, 256);
    const size_t n_waves = (n + kWave - 1) / kWave;
    OFP_TRY(h->rmask.ensure(n_waves * kRowWords * kWave * 4 + 4));
    auto launch = [&](auto fill, auto kb, size_t e_cap) {
      k_vl_rows<T, decltype(fill)::value, decltype(kb)::value><<<g, 256, 0, s>>>(
          c, g1, h->xyz.as<T>(), r2_one, sys, r2_sys, sid, h->cnt.as<uint32_t>(), h->rmask.as<uint32_t>(),
          h->off.as<unsigned long long>(), h->off.as<unsigned long long>() + n, (unsigned long long)e_cap,
          h->nbr.as<uint32_t>());
    };
    auto pass = [&](auto fill, size_t e_cap) {
      if (c.box_nn == 1)
        launch(fill, std::integral_constant<int, 1>{}, e_cap);
      else if (c.box_nn == 2)
        launch(fill, std::integral_constant<int, 2>{}, e_cap);
      else
        launch(fill, std::integral_constant<int, 3>{}, e_cap);
    };
    if (n > 0) {
      pass(std::false_type{}, 0);
      OFP_LAUNCHED();
    }
    OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
    uint64_t E = 0;
    if (sync) {
      OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
      OFP_HIP(hipStreamSynchronize(s));
      const size_t want = size_t(double(E) * (headroom > 1.0 ? headroom : 1.0)) + 1024;
      if (want > *e_cap_io) *e_cap_io = want;
      if (e_count) *e_count = E;
    }
    const size_t e_cap = *e_cap_io;
    OFP_TRY(h->nbr.ensure((e_cap + 4) * 4));  // + 3: the one-lane row kernels read whole groups of four entries
    if (n > 0) {
      pass(std::true_type{}, e_cap);
      OFP_LAUNCHED();
    }
    h->n_rows = n;
    h->n_entries = sync ? E : 0;
    h->built = true;
    h->symmetric = false;
    h->cl = cl;
    h->rows_path = true;
    return OFP_OK;
  }
  h->rows_path = false;
  if (h->mask_cap == 0) h->mask_cap = 16 * n + 4096;  // grown after the first count if short
  OFP_TRY(h->masks.ensure(h->mask_cap * 8));
  const TargetItem* items;
  const uint32_t* n_items;
  uint32_t* counter;
  size_t cap;
  OFP_TRY(ofp_cl_items_(cl, &items, &n_items, &counter, &cap, s));
  if (grid == 0) grid = persistent_grid();
  auto* mstate = h->mstate.as<unsigned long long>();
  if (v.n_part > 0) {
    k_vl_pack<T><<<grid_for(v.n_part, 256), 256, 0, s>>>(pos, stride, uint32_t(v.n_part), h->xyz.as<T>());
    OFP_LAUNCHED();
  }
  auto count = [&]() -> int {
    OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(mstate), 0, 2, s));
    if (n == 0) return OFP_OK;
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_count<T><<<grid, kVlBlock, 0, s>>>(c, h->xyz.as<T>(), r2_one, sys, r2_sys, items, n_items, counter,
                                            h->cnt.as<uint32_t>(), mstate, h->moff.as<unsigned long long>(),
                                            h->masks.as<uint2>(), (unsigned long long)h->mask_cap);
    OFP_LAUNCHED();
    return OFP_OK;
  };
  OFP_TRY(count());
  OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
  uint64_t E = 0;
  if (sync) {
    unsigned long long ms[2] = {0, 0};
    OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipMemcpyAsync(ms, mstate, 16, hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    if (ms[1] != 0) {  // masks did not fit: grow to the exact total (+ headroom) and count again
      h->mask_cap = size_t(double(ms[0]) * (headroom > 1.0 ? headroom : 1.0)) + 4096;
      OFP_TRY(h->masks.ensure(h->mask_cap * 8));
      OFP_TRY(count());
    }
    const size_t want = size_t(double(E) * (headroom > 1.0 ? headroom : 1.0)) + 1024;
    if (want > *e_cap_io) *e_cap_io = want;
    if (e_count) *e_count = E;
  }
  const size_t e_cap = *e_cap_io;
  OFP_TRY(h->nbr.ensure((e_cap + 4) * 4));  // + 3: the one-lane row kernels read whole groups of four entries
  if (n > 0) {
    OFP_TRY(fill_u32(counter, 0, 1, s));
    k_vl_fill<T><<<grid, kVlBlock, 0, s>>>(c, sys, items, n_items, counter, h->off.as<unsigned long long>(),
                                           h->off.as<unsigned long long>() + n, (unsigned long long)e_cap, mstate,
                                           h->moff.as<unsigned long long>(), h->masks.as<uint2>(),
                                           h->nbr.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_rows = n;
  h->n_entries = sync ? E : 0;  // otherwise on the device: off[n_rows]
  h->built = true;
  h->symmetric = false;
  h->cl = cl;
  return OFP_OK;
}

template <class T>
static int vl_build(ofp_vl* h, const ofp_cl* cl, const void* pos, size_t stride, double r_cut, hipStream_t s) {
  ClDev c;
  ofp_cl_view v;
  OFP_TRY(ofp_cl_dev_(cl, &c, &v));
  for (int k = 0; k < 3; ++k)
    OFP_CHECK(v.spacing[k] * c.box_nn >= r_cut * (1.0 - 1e-6), OFP_E_ARG,
              "stencil %u x spacing %g is smaller than the Verlet radius %g", c.box_nn, v.spacing[k], r_cut);
  const T rc = T(r_cut);
  size_t e_cap = 0;
  return vl_build_t<T>(h, cl, pos, stride, rc * rc, nullptr, nullptr, true, &e_cap, 1.0, 0, nullptr, s);
}

// Symmetric list from a full one: row r keeps the entries whose particle comes
// after the row's own in one total order, so every unordered interaction is kept
// once.  Domain neighbour j: kept iff j > i (sorted indices).  Ghost neighbour
// (ghost k of source particle q, shift id sid): kept iff sorted(q) > i, or
// q is the row's own particle and sid > 13 (shift ids s and -s are id and
// 26 - id, 13 the zero shift, so one image of each pair is kept).  One wave
// per row; kFill = false counts, true writes the kept entries in row order.
template <bool kFill>
__global__ __launch_bounds__(256) void k_vl_sym(const uint64_t* __restrict__ off, const uint32_t* __restrict__ nbr,
                                                uint64_t n_rows, const uint32_t* __restrict__ dom,
                                                const uint32_t* __restrict__ s2n, const uint32_t* __restrict__ n2s,
                                                uint32_t gm, const uint32_t* __restrict__ opid,
                                                const uint32_t* __restrict__ osid, uint32_t* __restrict__ cnt,
                                                const uint64_t* __restrict__ off_sym,
                                                uint32_t* __restrict__ nbr_sym) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t r = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (r >= n_rows) return;  // wave-uniform
  const uint32_t i = dom[r];
  const uint64_t e1 = off[r + 1];
  uint64_t out = kFill ? off_sym[r] : 0;
  uint32_t kept = 0;
  for (uint64_t e0 = off[r]; e0 < e1; e0 += kWave) {
    const uint64_t e = e0 + uint64_t(lane);
    bool keep = false;
    uint32_t j = 0;
    if (e < e1) {
      j = nbr[e];
      const uint32_t o = s2n[j];
      if (o < gm) {
        keep = j > i;
      } else {
        const uint32_t k = o - gm;
        const uint32_t q = n2s[opid[k]];
        keep = q > i || (q == i && osid[k] > 13u);
      }
    }
    const unsigned long long m = __ballot(keep);
    if (kFill && keep)
      nbr_sym[out + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u))] = j;
    out += uint64_t(__popcll(m));
    kept += uint32_t(__popcll(m));
  }
  if (!kFill && lane == 0) cnt[r] = kept;
}

}  // namespace ofp

using namespace ofp;

int ofp::vl_build_many(ofp_vl* h, const ofp_cl* cl, int dtype, const void* pos_sorted, size_t stride,
                       const void* sys_dev, const void* r2_sys_dev, size_t* e_cap, double headroom, int sizing,
                       unsigned grid, uint64_t* e_count, hipStream_t s, int rows_mode) {
  OFP_CHECK(h && cl && sys_dev && r2_sys_dev && e_cap, OFP_E_ARG, "vl_build_many: null argument");
  if (dtype == OFP_F32)
    return vl_build_t<float>(h, cl, pos_sorted, stride, 0.f, sys_dev, r2_sys_dev, sizing != 0, e_cap, headroom,
                             grid, e_count, s, rows_mode);
  return vl_build_t<double>(h, cl, pos_sorted, stride, 0.0, sys_dev, r2_sys_dev, sizing != 0, e_cap, headroom, grid,
                            e_count, s, rows_mode);
}

int ofp::vl_offsets(const ofp_vl* h, const unsigned long long** off) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "Verlet list not built");
  *off = h->off.as<unsigned long long>();
  return OFP_OK;
}

int ofp::vl_mask_state(const ofp_vl* h, const unsigned long long** mstate) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "Verlet list not built");
  *mstate = h->mstate.as<unsigned long long>();
  return OFP_OK;
}

int ofp::vl_set_mask_cap(ofp_vl* h, size_t mask_cap, size_t* old_cap) {
  OFP_CHECK(h && mask_cap > 0, OFP_E_ARG, "vl_set_mask_cap: null handle or zero capacity");
  if (old_cap) *old_cap = h->mask_cap;
  h->masks.release();
  h->mask_cap = mask_cap;
  return h->masks.ensure(mask_cap * 8);
}

extern "C" {

// accessor for the pair kernels (same library)
int ofp_vl_dev_(const ofp_vl* h, const ofp_cl* cl, int symmetric, const unsigned long long** off,
                const uint32_t** nbr) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "Verlet list not built");
  OFP_CHECK(h->cl == cl, OFP_E_STATE, "Verlet list not built from this cell list");
  OFP_CHECK(h->symmetric == (symmetric != 0), OFP_E_ARG,
            symmetric ? "a symmetric pass needs a list from ofp_vl_build_sym"
                      : "a symmetric list holds each interaction once: use the symmetric pair passes");
  *off = h->off.as<unsigned long long>();
  *nbr = h->nbr.as<uint32_t>();
  return OFP_OK;
}

// A device buffer of at least `bytes` owned by the list, for a pair kernel's
// per-particle stream over the rows' entries (rebuilt by each pass that uses it).
int ofp_vl_aux_(const ofp_vl* h, size_t bytes, void** p) {
  OFP_CHECK(h && p, OFP_E_ARG, "ofp_vl_aux_: null argument");
  ofp_vl* m = const_cast<ofp_vl*>(h);
  OFP_TRY(m->aux.ensure(bytes));
  *p = m->aux.p;
  return OFP_OK;
}

int ofp_vl_create(ofp_vl** h) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_vl_create: null handle pointer");
  *h = new ofp_vl();
  return OFP_OK;
}

int ofp_vl_destroy(ofp_vl* h) {
  if (!h) return OFP_OK;
  DevBuf* bufs[] = {&h->cnt, &h->off, &h->nbr, &h->scantmp, &h->masks, &h->moff, &h->mstate, &h->xyz, &h->aux, &h->rmask};
  for (DevBuf* b : bufs) b->release();
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

int ofp_vl_build_sym(ofp_vl* h, const ofp_vl* full, const ofp_ghost* g, void* stream) {
  OFP_CHECK(h && full && g && h != full, OFP_E_ARG, "ofp_vl_build_sym: null or aliased handle");
  OFP_CHECK(full->built && full->cl, OFP_E_STATE, "ofp_vl_build_sym: full list not built");
  ofp_cl_view v;
  OFP_TRY(ofp_cl_get_view(full->cl, &v));
  const uint32_t *pid = nullptr, *sid = nullptr;
  size_t n_ghost = 0;
  OFP_TRY(ofp_ghost_opart(g, &pid, &sid, &n_ghost));
  OFP_CHECK(v.n_part == v.ghost_marker + n_ghost, OFP_E_STATE,
            "ofp_vl_build_sym: the cell list holds %llu particles, the ghost list %zu after %llu",
            (unsigned long long)v.n_part, n_ghost, (unsigned long long)v.ghost_marker);
  OFP_CHECK(full->n_rows == v.ghost_marker, OFP_E_STATE, "ofp_vl_build_sym: rows %llu != domain particles %llu",
            (unsigned long long)full->n_rows, (unsigned long long)v.ghost_marker);
  hipStream_t s = as_stream(stream);
  const size_t n = full->n_rows;
  OFP_TRY(h->cnt.ensure((n + 1) * sizeof(uint32_t)));
  OFP_TRY(h->off.ensure((n + 1) * sizeof(uint64_t)));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(n)));
  const unsigned grid = grid_for(n * kWave, 256);
  if (n > 0) {
    k_vl_sym<false><<<grid, 256, 0, s>>>(full->off.as<uint64_t>(), full->nbr.as<uint32_t>(), n, v.dom_sort_ids,
                                          v.sort_to_nonsort, v.nonsort_to_sort, uint32_t(v.ghost_marker), pid, sid,
                                          h->cnt.as<uint32_t>(), nullptr, nullptr);
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u32_to_u64(h->cnt.as<uint32_t>(), h->off.as<uint64_t>(), n, h->scantmp.p, s));
  uint64_t E = 0;
  OFP_HIP(hipMemcpyAsync(&E, h->off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  OFP_TRY(h->nbr.ensure((E + 1) * sizeof(uint32_t)));
  if (n > 0) {
    k_vl_sym<true><<<grid, 256, 0, s>>>(full->off.as<uint64_t>(), full->nbr.as<uint32_t>(), n, v.dom_sort_ids,
                                         v.sort_to_nonsort, v.nonsort_to_sort, uint32_t(v.ghost_marker), pid, sid,
                                         nullptr, h->off.as<uint64_t>(), h->nbr.as<uint32_t>());
    OFP_LAUNCHED();
  }
  h->n_rows = n;
  h->n_entries = E;
  h->cl = full->cl;
  h->built = true;
  h->symmetric = true;
  return OFP_OK;
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
