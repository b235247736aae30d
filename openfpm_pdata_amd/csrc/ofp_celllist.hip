// ofp_celllist.hip -- CellList_gpu construction, reorder and restore.
//
// Build (per updateCellListGPU): a stable LSD radix sort of the particles by
// cell id, started from the identity order, so inside every cell the particles
// come out in ascending original index (domain ids < ghost ids: domain first).
// That is the canonical order, reached without atomics on cell counters.
//   1. k_rx_hist     pass 0: cell id per particle (the key), and per tile of
//                    4096 particles the histogram of the pass's digit (<= 8 bits),
//                    counted per wave with ballots, one LDS add per digit run.
//   2. column scan   of the [tile][digit] histogram table gives every (tile,
//                    digit) its first output slot.
//   3. k_rx_scatter  per tile: stable rank of each key inside the tile (ballot
//                    peer masks, per-wave running counts in LDS), staged in LDS in
//                    digit order, then written out as contiguous runs.
//   Passes 1.. repeat 1-3 on the next digit (k_rx_hist reads the keys).
//   Positions to be reordered ride along as a payload when pass 0 counts many
//   descents (input far from cell order); otherwise step 6 gathers them.
//   4. k_cl_runs     per run of equal sorted keys (one occupied cell): its
//                    length and its number of domain particles, from ballots of
//                    run starts and domain flags over 64 consecutive slots.
//   5. scan (u64)    packed (domain << 32 | count) per cell -> cell starts (low
//                    word) and domain-id starts (high word).
//   6. k_cl_emit     per sorted slot: nonsort_to_sort, domain sort ids, reordered
//                    positions and planes, and the quantised cell-relative stream.
// The particle count may be read from device memory (ensembles whose ghost count
// stays on the device): every kernel is launched for the capacity and returns
// past the count.
#include "ofp_internal.hpp"

#include <cmath>
#include <cstddef>
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
  // ensemble mode: per-system grids in device memory (owned by the caller)
  uint32_t n_sys = 0;               // 0: one grid
  const void* sys = nullptr;        // SysGrid<T>[n_sys]
  const uint32_t* sid = nullptr;    // system of each particle, input order
  const uint32_t* rbase = nullptr;  // first global x-row of each system (n_sys + 1)
  uint32_t n_rows = 0;              // x-rows of all grids
  const uint32_t* n_dev = nullptr;  // particle count in device memory (or null)
  // target-item work queue (built on first use after each construct)
  bool items_valid = false;
  size_t items_cap = 0;
  ofp::DevBuf packed, scan64, keys, keys2, s2n, vals2, n2s, dom, quant, scantmp, hist, sid_sorted, payload;
  ofp::DevBuf ctl;  // [0] descents counted by pass 0 (the payload decision)
  ofp::DevBuf colpart;  // chunk sums of the histogram table
  ofp::DevBuf items, nitems;  // nitems: [0] = item count, [1] = consumer queue head
  ofp::DevBuf aux;            // pair kernels' per-particle record stream (cell tiles)
  // sparse view (ofp_cl_build_sparse): non-empty cells only
  ofp::DevBuf nz_flag, nz_key, nz_start, nz_nn, nz_inv;
  uint64_t n_nz = 0;
  bool sparse_valid = false;
  uint32_t nz_box_nn = 1;  // stencil half-width of the sparse view's neighbour table
};

namespace ofp {

constexpr int kRxBlock = 256;
constexpr int kRxWaves = kRxBlock / kWave;
constexpr int kRxIters = 16;                        // 64-key rounds per wave
constexpr uint32_t kRxTile = kRxBlock * kRxIters;  // keys per tile
constexpr int kRxMaxBits = 8;
constexpr int kRxBins = 1 << kRxMaxBits;

__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Lanes of `act` whose digit equals this lane's (digit < 2^bits).
__device__ __forceinline__ unsigned long long rx_peers(uint32_t d, int bits, unsigned long long act) {
  unsigned long long m = act;
  for (int b = 0; b < bits; ++b) {
    const bool bit = ((d >> b) & 1u) != 0u;
    const unsigned long long v = __ballot(bit);
    m &= bit ? v : ~v;
  }
  return m;
}

__device__ __forceinline__ uint32_t count_of(uint32_t n_host, const uint32_t* n_dev) {
  return n_dev ? *n_dev : n_host;
}


// Tile histogram of one digit.  Pass 0 (kin == null) evaluates the key, stores
// it to kout and adds the tile's descents (k[e] > k[e+1] inside each 64-key
// round) to *disorder, which then tells how far the input is from cell order.
template <class Key>
__global__ __launch_bounds__(kRxBlock) void k_rx_hist(Key key, const uint32_t* __restrict__ kin,
                                                      uint32_t* __restrict__ kout, uint32_t n_host,
                                                      const uint32_t* __restrict__ n_dev, int shift, int bits,
                                                      uint32_t n_tiles, uint32_t* __restrict__ hist,
                                                      unsigned long long* __restrict__ disorder) {
  __shared__ uint32_t h[kRxBins];
  __shared__ uint32_t wdesc[kRxWaves];
  const uint32_t n = count_of(n_host, n_dev);
  const uint32_t t = blockIdx.x;
  for (int d = threadIdx.x; d < kRxBins; d += kRxBlock) h[d] = 0;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint32_t mask = (1u << bits) - 1u;
  const uint64_t base = uint64_t(t) * kRxTile + uint64_t(w) * (kRxIters * kWave);
  uint32_t desc = 0;
  // batches of rounds: the batch's loads first, then its stores, then the
  // counting; all 16 rounds at once for stored keys, 4 when the keys are
  // evaluated from the positions (their registers)
  auto rounds = [&](auto batch) {
    constexpr int kB = decltype(batch)::value;
#pragma unroll 1
    for (int b = 0; b < kRxIters; b += kB) {
      uint32_t kr[kB];
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const uint64_t e = base + uint64_t(b + i) * kWave + uint64_t(lane);
        kr[i] = 0u;
        if (e < n) kr[i] = kin ? kin[e] : key(uint32_t(e));
      }
      if (!kin) {
#pragma unroll
        for (int i = 0; i < kB; ++i) {
          const uint64_t e = base + uint64_t(b + i) * kWave + uint64_t(lane);
          if (e < n) kout[e] = kr[i];
        }
      }
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const uint64_t e = base + uint64_t(b + i) * kWave + uint64_t(lane);
        const bool v = e < n;
        const unsigned long long act = __ballot(v);
        if (act == 0ull) continue;  // wave-uniform: a round past n
        const uint32_t k = kr[i];
        if (!kin) {
          const uint32_t next = __shfl_down(k, 1, kWave);
          const bool nv = lane + 1 < kWave && ((act >> (lane + 1)) & 1ull);
          desc += (v && nv && k > next) ? 1u : 0u;
        }
        const uint32_t d = (k >> shift) & mask;
        const unsigned long long peers = rx_peers(d, bits, act);
        if (v && lanes_below(peers) == 0u) atomicAdd(&h[d], uint32_t(__popcll(peers)));
      }
    }
  };
  if (kin)
    rounds(std::integral_constant<int, kRxIters>{});
  else
    rounds(std::integral_constant<int, 4>{});
  if (!kin) {
    desc = wave_sum(desc);
    if (lane == 0) wdesc[w] = desc;
  }
  __syncthreads();
  for (int d = threadIdx.x; d < (1 << bits); d += kRxBlock) hist[size_t(t) * kRxBins + d] = h[d];
  if (!kin && threadIdx.x == 0) {
    uint32_t s = 0;
    for (int ww = 0; ww < kRxWaves; ++ww) s += wdesc[ww];
    if (s) atomicAdd(disorder, static_cast<unsigned long long>(s));
  }
}

// Carry the positions through the radix passes (instead of one gather at the
// end) when at least one key in kDisorderDen is followed by a smaller one: a
// gather from input that far from cell order reads a cache line per coordinate.
constexpr uint64_t kDisorderDen = 8;
__device__ __forceinline__ bool carry_positions(const unsigned long long* disorder, uint32_t n) {
  return disorder && *disorder * kDisorderDen > uint64_t(n);
}

// Offsets of the [tile][digit] table (rows of kRxBins, written and read as
// coalesced rows): offs[t][d] = sum of all counts of digits < d + sum of the
// counts of digit d in tiles < t.  The tiles form at most kColChunks chunks;
// every loop below loads kColBatch rows before it uses them.
constexpr uint32_t kColChunks = 256;
constexpr uint32_t kColBatch = 16;

// part[c][d] = counts of digit d in the tiles of chunk c
__global__ __launch_bounds__(kRxBins) void k_rx_colsum(const uint32_t* __restrict__ tab, uint32_t n_tiles,
                                                       uint32_t ch, int nb, uint32_t* __restrict__ part) {
  const uint32_t c = blockIdx.x;
  const int d = threadIdx.x;
  const uint32_t t0 = c * ch;
  const uint32_t t1 = t0 + ch < n_tiles ? t0 + ch : n_tiles;
  uint32_t s = 0;
  if (d < nb) {
    for (uint32_t t = t0; t < t1; t += kColBatch) {
      uint32_t v[kColBatch];
#pragma unroll
      for (uint32_t i = 0; i < kColBatch; ++i) v[i] = t + i < t1 ? tab[size_t(t + i) * kRxBins + d] : 0u;
#pragma unroll
      for (uint32_t i = 0; i < kColBatch; ++i) s += v[i];
    }
  }
  part[size_t(c) * kRxBins + d] = s;
}

// One block: part[c][d] becomes the count of digit d in the chunks before c, and
// row kColChunks the digit bases (counts of all smaller digits).
__global__ __launch_bounds__(kRxBins) void k_rx_colscan(uint32_t* __restrict__ part, uint32_t n_chunks) {
  __shared__ uint32_t ws[kRxBins / kWave];
  const int d = threadIdx.x;
  uint32_t run = 0;
  for (uint32_t c = 0; c < n_chunks; c += kColBatch) {
    uint32_t v[kColBatch];
#pragma unroll
    for (uint32_t i = 0; i < kColBatch; ++i) v[i] = c + i < n_chunks ? part[size_t(c + i) * kRxBins + d] : 0u;
#pragma unroll
    for (uint32_t i = 0; i < kColBatch; ++i) {
      if (c + i < n_chunks) part[size_t(c + i) * kRxBins + d] = run;
      run += v[i];
    }
  }
  // digit bases: exclusive scan of the digit totals over the block
  const uint32_t inc = wave_incl_scan(run);
  const int lane = d & (kWave - 1), w = d / kWave;
  if (lane == kWave - 1) ws[w] = inc;
  __syncthreads();
  uint32_t base = inc - run;
  for (int i = 0; i < w; ++i) base += ws[i];
  part[size_t(kColChunks) * kRxBins + d] = base;
}

// tab[t][d] becomes offs[t][d]
__global__ __launch_bounds__(kRxBins) void k_rx_colapply(uint32_t* __restrict__ tab, uint32_t n_tiles,
                                                         uint32_t ch, int nb, const uint32_t* __restrict__ part) {
  const uint32_t c = blockIdx.x;
  const int d = threadIdx.x;
  if (d >= nb) return;
  const uint32_t t0 = c * ch;
  const uint32_t t1 = t0 + ch < n_tiles ? t0 + ch : n_tiles;
  uint32_t run = part[size_t(c) * kRxBins + d] + part[size_t(kColChunks) * kRxBins + d];
  for (uint32_t t = t0; t < t1; t += kColBatch) {
    uint32_t v[kColBatch];
#pragma unroll
    for (uint32_t i = 0; i < kColBatch; ++i) v[i] = t + i < t1 ? tab[size_t(t + i) * kRxBins + d] : 0u;
#pragma unroll
    for (uint32_t i = 0; i < kColBatch; ++i) {
      if (t + i < t1) tab[size_t(t + i) * kRxBins + d] = run;
      run += v[i];
    }
  }
}

// Stable scatter of one tile by one digit; offs = [tile][digit] offset table.
// vin == null stands for the identity (pass 0).  Payload: three planes of T
// (positions) that travel with the keys, pin (stride si) in the input order,
// pout (stride so) in the output order; null = none.  Keys, values and each
// payload plane are read in input order (coalesced), placed in LDS at their
// staged slot, and written out in staged order: runs of one digit each.
template <class T>
__global__ __launch_bounds__(kRxBlock) void k_rx_scatter(const uint32_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin, uint32_t n_host,
                                                         const uint32_t* __restrict__ n_dev, int shift,
                                                         int bits, uint32_t n_tiles,
                                                         const uint32_t* __restrict__ offs,
                                                         uint32_t* __restrict__ kout,
                                                         uint32_t* __restrict__ vout, const T* __restrict__ pin,
                                                         size_t si, T* __restrict__ pout, size_t so,
                                                         const unsigned long long* __restrict__ disorder) {
  __shared__ uint32_t wh[kRxWaves][kRxBins];  // running count per wave, then wave prefix
  __shared__ uint32_t tstart[kRxBins];        // first staged slot of each digit
  __shared__ uint32_t goff[kRxBins];          // first output slot of each digit
  __shared__ uint32_t wsum[kRxWaves + 1];
  __shared__ uint32_t sk[kRxTile];                       // staged keys, then their output slots
  __shared__ uint32_t sv[kRxTile * (sizeof(T) / 4)];     // staged values, then one payload plane
  T* const sp = reinterpret_cast<T*>(sv);
  const uint32_t n = count_of(n_host, n_dev);
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const uint64_t tile0 = uint64_t(t) * kRxTile;
  if (tile0 >= n) return;  // block-uniform
  const int nb = 1 << bits;
  const uint32_t mask = uint32_t(nb) - 1u;
  for (int d = threadIdx.x; d < kRxBins; d += kRxBlock) {
#pragma unroll
    for (int w = 0; w < kRxWaves; ++w) wh[w][d] = 0;
    goff[d] = d < nb ? offs[size_t(t) * kRxBins + d] : 0u;
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint64_t base = tile0 + uint64_t(w) * (kRxIters * kWave);
  uint32_t key[kRxIters], lr[kRxIters], val[kRxIters];
  // all loads of the wave's rounds first
#pragma unroll
  for (int it = 0; it < kRxIters; ++it) {
    const uint64_t e = base + uint64_t(it) * kWave + uint64_t(lane);
    key[it] = 0u;
    val[it] = uint32_t(e);
    if (e < n) {
      key[it] = kin[e];
      if (vin) val[it] = vin[e];
    }
  }
#pragma unroll
  for (int it = 0; it < kRxIters; ++it) {
    const uint64_t e = base + uint64_t(it) * kWave + uint64_t(lane);
    const bool v = e < n;
    const unsigned long long act = __ballot(v);
    const uint32_t k = key[it];
    const uint32_t d = (k >> shift) & mask;
    uint32_t r = 0;
    if (act != 0ull) {
      const unsigned long long peers = rx_peers(d, bits, act);
      const uint32_t below = lanes_below(peers);
      const uint32_t prior = v ? wh[w][d] : 0u;
      r = prior + below;
      if (v && below == 0u) wh[w][d] = prior + uint32_t(__popcll(peers));
    }
    lr[it] = r;
  }
  __syncthreads();
  // wave prefixes per digit and the tile's digit totals
  uint32_t tot = 0;
  if (threadIdx.x < kRxBins) {
    const int d = threadIdx.x;
#pragma unroll
    for (int ww = 0; ww < kRxWaves; ++ww) {
      const uint32_t c = wh[ww][d];
      wh[ww][d] = tot;
      tot += c;
    }
  }
  // exclusive scan of the digit totals (one digit per thread)
  const uint32_t inc = wave_incl_scan(tot);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int ww = 0; ww < kRxWaves; ++ww) {
      const uint32_t c = wsum[ww];
      wsum[ww] = s;
      s += c;
    }
  }
  __syncthreads();
  if (threadIdx.x < kRxBins) tstart[threadIdx.x] = wsum[w] + inc - tot;
  __syncthreads();
  // lr[it] becomes the staged slot of element it
#pragma unroll
  for (int it = 0; it < kRxIters; ++it) {
    const uint64_t e = base + uint64_t(it) * kWave + uint64_t(lane);
    if (e < n) {
      const uint32_t d = (key[it] >> shift) & mask;
      const uint32_t slot = tstart[d] + wh[w][d] + lr[it];
      lr[it] = slot;
      sk[slot] = key[it];
      sv[slot] = val[it];
    }
  }
  __syncthreads();
  const uint32_t len = uint32_t(n - tile0 < kRxTile ? n - tile0 : kRxTile);
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  for (uint32_t s = threadIdx.x; s < len; s += kRxBlock) {
    const uint32_t k = sk[s];
    const uint32_t d = (k >> shift) & mask;
    const uint32_t o = goff[d] + (s - tstart[d]);
    // o < n always holds for a consistent table; the test keeps a bad one in bounds
    if (o < n) {
      kout[o] = k;
      vout[o] = sv[s];
    }
    sk[s] = o < n ? o : kNone;
  }
  if (!pin || !carry_positions(disorder, n)) return;  // grid-uniform
  for (int q = 0; q < 3; ++q) {
    __syncthreads();  // sv / the previous plane has been read
#pragma unroll
    for (int it = 0; it < kRxIters; ++it) {
      const uint64_t e = base + uint64_t(it) * kWave + uint64_t(lane);
      if (e < n) sp[lr[it]] = pin[size_t(q) * si + e];
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < len; s += kRxBlock) {
      const uint32_t o = sk[s];
      if (o != kNone) pout[size_t(q) * so + o] = sp[s];
    }
  }
}

// Per occupied cell, from the sorted keys (packed[] cleared before): (domain
// particles << 32 | particles).  Each wave covers 64 consecutive sorted slots;
// ballots of "a run of equal keys starts here" and "domain particle" split the
// wave into run segments, whose lengths and domain counts are bit counts.  A
// segment that is a whole run is stored; the pieces of a run that crosses a
// wave boundary are added atomically.  One coalesced read of keys and
// sort_to_nonsort; empty cells keep 0.  A wave takes kRunRounds rounds of 64
// slots with all loads issued first (the work per slot is a few bit operations,
// so the loads' latency is what there is to hide).
constexpr int kRunRounds = 16;

// The runs of one wave's kRunRounds rounds of 64 consecutive sorted slots from
// w0: kk = keys, kp = the previous slot's key, dmf = domain particle.  The run
// that reaches the end of a round is carried (wave-uniform) into the next
// round: only a run that crosses the wave's first or last slot is added
// atomically, every other run is stored once.
__device__ __forceinline__ void runs_of_wave(const uint32_t (&kk)[kRunRounds], const uint32_t (&kp)[kRunRounds],
                                             const bool (&dmf)[kRunRounds], uint64_t w0, uint32_t n,
                                             uint64_t n_cells, unsigned long long* __restrict__ packed) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t ck = 0;              // carried run: key,
  unsigned long long cval = 0;  // (domain << 32 | count) so far,
  bool copen = false;           // open,
  bool cown = false;            // started inside this wave
  auto seg_val = [](unsigned long long D, int a, int e) {
    const unsigned long long upto = e == kWave - 1 ? ~0ull : (1ull << (e + 1)) - 1ull;
    const unsigned long long seg = upto & ~((1ull << a) - 1ull);
    return (static_cast<unsigned long long>(__popcll(D & seg)) << 32) | static_cast<unsigned long long>(e - a + 1);
  };
  auto close = [&]() {
    if (copen && lane == 0 && uint64_t(ck) < n_cells) {
      if (cown)
        packed[ck] = cval;
      else
        atomicAdd(&packed[ck], cval);
    }
    copen = false;
  };
#pragma unroll
  for (int r = 0; r < kRunRounds; ++r) {
    const uint64_t f = w0 + uint64_t(r) * kWave + uint64_t(lane);
    const bool v = f < n;
    const uint32_t k = kk[r];
    const bool st = v && (f == 0 || kp[r] != k);
    const bool dm = v && dmf[r];
    const unsigned long long S = __ballot(st), D = __ballot(dm), V = __ballot(v);
    if (V == 0ull) break;  // wave-uniform: this and later rounds are past n
    const int last = 63 - __builtin_clzll(V);
    const uint64_t f0 = w0 + uint64_t(r) * kWave;  // slot of lane 0
    const bool ends_at_n = f0 + uint64_t(last) + 1 == n;
    if (S & 1ull) {
      close();  // the carried run ended at the previous round's last slot
    } else {
      // lanes [0, e0] continue the run of the previous slot
      const unsigned long long up = S & ~1ull;
      const int e0 = up ? __builtin_ctzll(up) - 1 : last;
      const unsigned long long v0 = seg_val(D, 0, e0);
      if (copen) {
        cval += v0;
      } else {
        ck = uint32_t(__builtin_amdgcn_readfirstlane(int(k)));
        cval = v0;
        cown = false;
        copen = true;
      }
      if (up || ends_at_n) close();
    }
    // runs starting in this round: stored here when they end in the round, the
    // last one carried otherwise
    const int hi = S ? 63 - __builtin_clzll(S) : -1;
    const bool hi_open = hi >= 0 && !ends_at_n;
    if (st && uint64_t(k) < n_cells && !(hi_open && lane == hi)) {
      const unsigned long long above = lane == kWave - 1 ? 0ull : (S >> (lane + 1)) << (lane + 1);
      const int e = above ? __builtin_ctzll(above) - 1 : last;
      packed[k] = seg_val(D, lane, e);
    }
    if (hi_open) {
      ck = uint32_t(__builtin_amdgcn_readlane(int(k), hi));
      cval = seg_val(D, hi, last);
      cown = true;
      copen = true;
    }
  }
  // a run open at the wave's end may go on in the next wave
  if (copen) {
    cown = false;
    close();
  }
}

// Per occupied cell, from the sorted keys (packed[] cleared before): (domain
// particles << 32 | particles).  Each wave covers 64 consecutive sorted slots
// per round; ballots of "a run of equal keys starts here" and "domain particle"
// split the wave into run segments, whose lengths and domain counts are bit
// counts (runs_of_wave).  One coalesced read of keys and sort_to_nonsort; empty
// cells keep 0.  All loads of the wave's kRunRounds rounds are issued first
// (the work per slot is a few bit operations, so the loads' latency is what
// there is to hide).
__global__ __launch_bounds__(256) void k_cl_runs(const uint32_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ s2n, uint32_t n_host,
                                                 const uint32_t* __restrict__ n_dev, uint32_t gm,
                                                 uint64_t n_cells, unsigned long long* __restrict__ packed) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w0 = (uint64_t(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave) * (kRunRounds * kWave);
  const uint32_t n = count_of(n_host, n_dev);
  uint32_t kk[kRunRounds], kp[kRunRounds];
  bool dm[kRunRounds];
#pragma unroll
  for (int r = 0; r < kRunRounds; ++r) {
    const uint64_t f = w0 + uint64_t(r) * kWave + uint64_t(lane);
    kk[r] = 0u;
    kp[r] = ~0u;
    dm[r] = false;
    if (f < n) {
      kk[r] = keys[f];
      dm[r] = s2n[f] < gm;
      if (f > 0) kp[r] = keys[f - 1];
    }
  }
  runs_of_wave(kk, kp, dm, w0, n, n_cells, packed);
}

template <class T>
struct EmitArgs {
  const T* pos;
  size_t stride;
  T* pos_sorted;
  size_t sorted_stride;
  int4* quant;
  double qlo[3], qsp[3], qinv[3];  // cell origin, spacing, 2^28 / spacing
  uint32_t divc[3];
  uint32_t gm;
  uint32_t* n2s;
  uint32_t* dom;
  PlaneArgs pa;
  // ensembles: grid of each particle's system
  const SysGrid<T>* sys;
  const uint32_t* sid;
  uint32_t* sid_sorted;
  // the passes carried the positions into pos_sorted when *disorder says so
  const unsigned long long* disorder;
};

template <class T>
__device__ __forceinline__ void cl_emit(const EmitArgs<T>& A, uint32_t o, uint32_t f, uint32_t c,
                                        uint32_t dstart, uint32_t rank, bool pos_done) {
  A.n2s[o] = f;
  if (o < A.gm) A.dom[dstart + rank] = f;
  uint32_t s = 0;
  if (A.sys) {
    s = A.sid[o];
    A.sid_sorted[f] = s;
  }
  if (A.pos_sorted || A.quant) {
    T x, y, z;
    if (pos_done) {
      x = A.pos_sorted[f];
      y = A.pos_sorted[A.sorted_stride + f];
      z = A.pos_sorted[2 * A.sorted_stride + f];
    } else {
      x = A.pos[o];
      y = A.pos[A.stride + o];
      z = A.pos[2 * A.stride + o];
    }
    if (A.pos_sorted && !pos_done) {
      A.pos_sorted[f] = x;
      A.pos_sorted[A.sorted_stride + f] = y;
      A.pos_sorted[2 * A.sorted_stride + f] = z;
    }
    if (A.quant) {
      const double xs[3] = {double(x), double(y), double(z)};
      if (A.sys) {
        const SysGrid<T>& G = A.sys[s];
        const uint32_t cl = c - G.cbase;
        const uint32_t cc[3] = {cl % G.divc[0], (cl / G.divc[0]) % G.divc[1], cl / (G.divc[0] * G.divc[1])};
        A.quant[f] = quantise(xs, cc, G.qlo, G.qsp, G.qinv);
      } else {
        const uint32_t cc[3] = {c % A.divc[0], (c / A.divc[0]) % A.divc[1], c / (A.divc[0] * A.divc[1])};
        A.quant[f] = quantise(xs, cc, A.qlo, A.qsp, A.qinv);
      }
    }
  }
  copy_planes(A.pa, o, f);
}

template <class T>
__global__ __launch_bounds__(256) void k_cl_emit(EmitArgs<T> A, const uint32_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ s2n, uint32_t n_host,
                                                 const uint32_t* __restrict__ n_dev,
                                                 const unsigned long long* __restrict__ cs) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = count_of(n_host, n_dev);
  if (f >= n) return;
  const bool pos_done = A.pos_sorted && carry_positions(A.disorder, n);
  const uint32_t c = keys[f];
  const unsigned long long s0 = cs[c];
  cl_emit(A, s2n[f], f, c, uint32_t(s0 >> 32), f - uint32_t(s0), pos_done);
}

// The build of input already in cell order, speculatively, with the check of
// that order.  Per tile of kRxTile particles: the cell keys and the descents
// k[f-1] > k[f] (the key before a wave's first slot recomputed from its
// position), added to *disorder.  A tile without descents writes what the
// build gives for its particles when the whole input is in cell order: the
// identity permutations, the positions and planes copied, the domain sort
// ids, the quantised stream, and its cells' (domain << 32 | count) into
// packed (cleared before).  If *disorder ends at zero these are the build's
// results; otherwise the radix passes (from the positions: neither keys nor
// histograms are kept here) and k_cl_emit overwrite every one of them.
template <class T>
__global__ __launch_bounds__(kRxBlock) void k_cl_first(CellGeom<T> g, EmitArgs<T> A, uint32_t n,
                                                       uint32_t* __restrict__ s2n,
                                                       unsigned long long* __restrict__ disorder,
                                                       unsigned long long* __restrict__ packed, uint64_t n_cells) {
  static_assert(kRunRounds == kRxIters, "one wave's rounds are its run rounds");
  __shared__ uint32_t wdesc[kRxWaves];
  const uint32_t t = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const uint64_t base = uint64_t(t) * kRxTile + uint64_t(w) * (kRxIters * kWave);
  const T* pos = A.pos;
  const size_t st = A.stride;
  uint32_t kk[kRxIters], kp[kRxIters];
  // keys, four rounds' loads at a time; the positions stay in registers for
  // the identity build below (read once from memory)
  T x[kRxIters], y[kRxIters], z[kRxIters];
#pragma unroll
  for (int b = 0; b < kRxIters; b += 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint64_t e = base + uint64_t(b + i) * kWave + uint64_t(lane);
      x[b + i] = y[b + i] = z[b + i] = T(0);
      if (e < n) {
        x[b + i] = pos[e];
        y[b + i] = pos[st + e];
        z[b + i] = pos[2 * st + e];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint64_t e = base + uint64_t(b + i) * kWave + uint64_t(lane);
      uint32_t cx, cy, cz;
      kk[b + i] = e < n ? cell_of(g, x[b + i], y[b + i], z[b + i], &cx, &cy, &cz) : 0u;
    }
  }
  // the key of the slot before the wave's first one
  uint32_t before = ~0u;
  if (base > 0 && base - 1 < n) {
    uint32_t cx, cy, cz;
    const uint64_t e = base - 1;
    before = cell_of(g, pos[e], pos[st + e], pos[2 * st + e], &cx, &cy, &cz);
  }
  uint32_t desc = 0;
#pragma unroll
  for (int r = 0; r < kRxIters; ++r) {
    const uint64_t e = base + uint64_t(r) * kWave + uint64_t(lane);
    const uint32_t up = __shfl_up(kk[r], 1, kWave);
    const uint32_t prev_last = r == 0 ? before : uint32_t(__builtin_amdgcn_readlane(int(kk[r - 1]), kWave - 1));
    kp[r] = lane == 0 ? prev_last : up;
    desc += (e < n && e > 0 && kp[r] > kk[r]) ? 1u : 0u;
  }
  desc = wave_sum(desc);
  if (lane == 0) wdesc[w] = desc;
  __syncthreads();
  uint32_t tile_desc = 0;
  for (int ww = 0; ww < kRxWaves; ++ww) tile_desc += wdesc[ww];
  if (threadIdx.x == 0 && tile_desc) atomicAdd(disorder, static_cast<unsigned long long>(tile_desc));
  if (tile_desc != 0) return;  // block-uniform
  // the identity build of this tile
  bool dm[kRxIters];
#pragma unroll
  for (int r = 0; r < kRxIters; ++r) {
    const uint64_t e = base + uint64_t(r) * kWave + uint64_t(lane);
    dm[r] = e < A.gm;
    if (e >= n) continue;
    const uint32_t f = uint32_t(e);
    s2n[f] = f;
    A.n2s[f] = f;
    if (f < A.gm) A.dom[f] = f;
    if (A.pos_sorted) {
      A.pos_sorted[f] = x[r];
      A.pos_sorted[A.sorted_stride + f] = y[r];
      A.pos_sorted[2 * A.sorted_stride + f] = z[r];
    }
    if (A.quant) {
      const uint32_t c = kk[r];
      const uint32_t cc[3] = {c % A.divc[0], (c / A.divc[0]) % A.divc[1], c / (A.divc[0] * A.divc[1])};
      const double xs[3] = {double(x[r]), double(y[r]), double(z[r])};
      A.quant[f] = quantise(xs, cc, A.qlo, A.qsp, A.qinv);
    }
    copy_planes(A.pa, f, f);
  }
  runs_of_wave(kk, kp, dm, base, n, n_cells, packed);
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

// Digit widths of the radix passes for keys < n_cells: as few passes of at most
// 8 bits as cover the key, widths as even as possible.
static int radix_plan(uint64_t n_cells, int* widths) {
  int bits = 1;
  while (bits < 32 && (uint64_t(1) << bits) < n_cells) ++bits;
  const int passes = (bits + kRxMaxBits - 1) / kRxMaxBits;
  for (int p = 0; p < passes; ++p) widths[p] = bits / passes + (p < bits % passes ? 1 : 0);
  return passes;
}

// n_cap: particles the buffers hold (the count itself is n_host or *n_dev).
// g1 (one grid, count on the host): k_cl_first first checks the input's order
// while writing the build of ordered input; the host reads its descent count
// (one scalar) and, for input already in cell order, the build ends with the
// scan of the cell counts it wrote; otherwise the radix build follows.
// Ensembles (g1 null) keep every step on the device (their rebuilds are
// captured in graphs).
template <class T, class Key>
static int cl_build(ofp_cl* h, const Key& key, const EmitArgs<T>& A0, size_t n_cap, uint32_t n_host,
                    const uint32_t* n_dev, size_t gm, hipStream_t s, const CellGeom<T>* g1 = nullptr) {
  const uint64_t C = h->n_cells;
  const uint32_t n_tiles = uint32_t((n_cap + kRxTile - 1) / kRxTile);
  const size_t hist_n = size_t(kRxBins) * (n_tiles ? n_tiles : 1);
  OFP_TRY(h->packed.ensure(C * 8));
  OFP_TRY(h->scan64.ensure((C + 1) * 8));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(C)));
  OFP_TRY(h->colpart.ensure(size_t(kColChunks + 1) * kRxBins * 4));
  // the [tile][digit] table in chunks of ch tiles
  const uint32_t ch = n_tiles ? (n_tiles + kColChunks - 1) / kColChunks : 1;
  const uint32_t n_chunks = (n_tiles + ch - 1) / ch;
  uint32_t* part = h->colpart.as<uint32_t>();
  OFP_TRY(h->keys.ensure((n_cap + 1) * 4));
  OFP_TRY(h->keys2.ensure((n_cap + 1) * 4));
  OFP_TRY(h->s2n.ensure((n_cap + 1) * 4));
  OFP_TRY(h->vals2.ensure((n_cap + 1) * 4));
  OFP_TRY(h->n2s.ensure((n_cap + 1) * 4));
  OFP_TRY(h->dom.ensure((gm + 1) * 4));
  OFP_TRY(h->hist.ensure((hist_n + 1) * 4));
  const bool want_quant = (h->opt & OFP_CL_QUANT) != 0;
  if (want_quant) OFP_TRY(h->quant.ensure((n_cap + 1) * 16));
  if (h->n_sys) OFP_TRY(h->sid_sorted.ensure((n_cap + 1) * 4));

  int widths[4];
  const int passes = radix_plan(C, widths);
  // With sorted positions requested they may travel with the keys as a payload
  // (the passes before the last alternate between two payload buffers, the last
  // one writes the sorted positions); the kernels carry them only when pass 0
  // finds the input far from cell order (carry_positions), else k_cl_emit
  // gathers them.
  const bool payload = A0.pos_sorted != nullptr;
  if (payload && passes > 1) OFP_TRY(h->payload.ensure(2 * 3 * n_cap * sizeof(T)));
  T* pbuf[2] = {h->payload.as<T>(), h->payload.as<T>() + 3 * n_cap};
  OFP_TRY(h->ctl.ensure(sizeof(unsigned long long)));
  auto* disorder = h->ctl.as<unsigned long long>();
  // ping-pong buffers chosen so that the last pass lands in keys / s2n
  uint32_t* kbuf[2] = {h->keys.as<uint32_t>(), h->keys2.as<uint32_t>()};
  uint32_t* vbuf[2] = {h->s2n.as<uint32_t>(), h->vals2.as<uint32_t>()};
  uint32_t* hist = h->hist.as<uint32_t>();
  OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(disorder), 0, 1, s));
  auto* packed = h->packed.as<unsigned long long>();
  auto* cs = h->scan64.as<unsigned long long>();
  const bool want_quant_ = want_quant;
  auto emit_args = [&]() {
    EmitArgs<T> A = A0;
    A.quant = want_quant_ ? h->quant.as<int4>() : nullptr;
    A.gm = uint32_t(gm);
    A.n2s = h->n2s.as<uint32_t>();
    A.dom = h->dom.as<uint32_t>();
    A.sid_sorted = h->n_sys ? h->sid_sorted.as<uint32_t>() : nullptr;
    A.disorder = payload ? disorder : nullptr;
    return A;
  };
  auto finish = [&]() {
    h->n_part = n_dev ? n_cap : n_host;
    h->gm = gm;
    h->n_dev = n_dev;
    h->built = true;
    h->sparse_valid = false;
    h->items_valid = false;
    return OFP_OK;
  };
  const bool first_fused = g1 != nullptr && n_dev == nullptr && n_tiles > 0;
  if (first_fused) {
    OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(packed), 0, C, s));
    k_cl_first<T><<<n_tiles, kRxBlock, 0, s>>>(*g1, emit_args(), n_host, vbuf[0], disorder, packed, C);
    OFP_LAUNCHED();
    unsigned long long desc = 0;
    OFP_HIP(hipMemcpyAsync(&desc, disorder, sizeof(desc), hipMemcpyDeviceToHost, s));
    OFP_HIP(hipStreamSynchronize(s));
    if (desc == 0) {  // input in cell order: every tile wrote its identity build
      OFP_TRY(scan_excl_u64(reinterpret_cast<const uint64_t*>(packed), reinterpret_cast<uint64_t*>(cs), C,
                            h->scantmp.p, s));
      return finish();
    }
    // not in cell order: the radix build from the positions (its pass 0
    // counts the descents again)
    OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(disorder), 0, 1, s));
  }
  if (n_tiles > 0) {
    uint32_t* kin = kbuf[passes % 2];  // pass-0 keys
    const uint32_t* vin = nullptr;
    const T* pin = payload ? A0.pos : nullptr;
    size_t si = A0.stride;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
      const int bits = widths[p];
      uint32_t* kout = kbuf[(passes - 1 - p) % 2];
      uint32_t* vout = vbuf[(passes - 1 - p) % 2];
      T* pout = nullptr;
      size_t so = 0;
      if (payload) {
        const bool last = p == passes - 1;
        pout = last ? A0.pos_sorted : pbuf[p % 2];
        so = last ? A0.sorted_stride : n_cap;
      }
      if (p == 0) {
        k_rx_hist<Key><<<n_tiles, kRxBlock, 0, s>>>(key, nullptr, kin, n_host, n_dev, shift, bits, n_tiles, hist,
                                                    disorder);
        OFP_LAUNCHED();
      } else {
        k_rx_hist<Key><<<n_tiles, kRxBlock, 0, s>>>(key, kin, nullptr, n_host, n_dev, shift, bits, n_tiles, hist,
                                                    nullptr);
        OFP_LAUNCHED();
      }
      k_rx_colsum<<<n_chunks, kRxBins, 0, s>>>(hist, n_tiles, ch, 1 << bits, part);
      OFP_LAUNCHED();
      k_rx_colscan<<<1, kRxBins, 0, s>>>(part, n_chunks);
      OFP_LAUNCHED();
      k_rx_colapply<<<n_chunks, kRxBins, 0, s>>>(hist, n_tiles, ch, 1 << bits, part);
      OFP_LAUNCHED();
      k_rx_scatter<T><<<n_tiles, kRxBlock, 0, s>>>(kin, vin, n_host, n_dev, shift, bits, n_tiles, hist, kout, vout,
                                                   pin, si, pout, so, disorder);
      OFP_LAUNCHED();
      kin = kout;
      vin = vout;
      pin = pout;
      si = so;
      shift += bits;
    }
  }
  OFP_TRY(fill_u64(reinterpret_cast<uint64_t*>(packed), 0, C, s));
  if (n_tiles > 0) {
    k_cl_runs<<<grid_for(n_cap, 256 * kRunRounds), 256, 0, s>>>(h->keys.as<uint32_t>(), h->s2n.as<uint32_t>(),
                                                                n_host, n_dev, uint32_t(gm), C, packed);
    OFP_LAUNCHED();
  }
  OFP_TRY(scan_excl_u64(reinterpret_cast<const uint64_t*>(packed), reinterpret_cast<uint64_t*>(cs), C,
                        h->scantmp.p, s));
  if (n_tiles > 0) {
    k_cl_emit<T><<<grid_for(n_cap, 256), 256, 0, s>>>(emit_args(), h->keys.as<uint32_t>(), h->s2n.as<uint32_t>(),
                                                       n_host, n_dev, cs);
    OFP_LAUNCHED();
  }
  return finish();
}

template <class T>
static int cl_construct(ofp_cl* h, const CellGeom<T>& g, const void* pos_v, size_t stride,
                        size_t n_tot, size_t gm, void* pos_sorted_v, size_t sorted_stride,
                        const PlaneArgs& pa, hipStream_t s) {
  const T* pos = static_cast<const T*>(pos_v);
  EmitArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  A.pos = pos;
  A.stride = stride;
  A.pos_sorted = static_cast<T*>(pos_sorted_v);
  A.sorted_stride = sorted_stride;
  for (int k = 0; k < 3; ++k) {
    A.qlo[k] = double(g.lo[k]);
    A.qsp[k] = double(g.sp[k]);
    A.qinv[k] = double(kQuantOne) / double(g.sp[k]);
    A.divc[k] = g.divc[k];
  }
  A.pa = pa;
  KeyOne<T> key{g, pos, stride};
  return cl_build<T>(h, key, A, n_tot, uint32_t(n_tot), nullptr, gm, s, &g);
}

template <class T>
static int cl_construct_many_t(ofp_cl* h, const void* pos_v, size_t stride, size_t n_cap, const uint32_t* n_dev,
                             size_t gm, void* pos_sorted_v, size_t sorted_stride, hipStream_t s) {
  const T* pos = static_cast<const T*>(pos_v);
  EmitArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  A.pos = pos;
  A.stride = stride;
  A.pos_sorted = static_cast<T*>(pos_sorted_v);
  A.sorted_stride = sorted_stride;
  A.sys = static_cast<const SysGrid<T>*>(h->sys);
  A.sid = h->sid;
  KeyMany<T> key{A.sys, h->sid, pos, stride};
  return cl_build<T>(h, key, A, n_cap, uint32_t(n_cap), n_dev, gm, s);
}

// Target items for the Verlet build: one thread per x-row of cells groups
// consecutive cells greedily while they hold <= 64 domain particles in total
// (spanning at most kItemMaxCells cells); a cell with more than 64 is split
// into chunks of 64.  Every domain particle is in exactly one item.  In an
// ensemble the rows of all grids are numbered consecutively and the item
// carries its system in ncx >> 8.
constexpr uint32_t kItemTargets = 64;

__global__ __launch_bounds__(256) void k_cl_items(const unsigned long long* __restrict__ cs,
                                                  uint32_t divc0_one, uint32_t nrows, uint32_t n_sys,
                                                  const uint32_t* __restrict__ rbase,
                                                  const uint32_t* __restrict__ sysdesc, uint32_t sys_words,
                                                  TargetItem* __restrict__ items,
                                                  uint32_t* __restrict__ nitems) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  uint32_t base = r * divc0_one, divc0 = divc0_one, tag = 0;
  if (n_sys) {
    // system of row r: last s with rbase[s] <= r
    uint32_t lo = 0, hi = n_sys;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (rbase[mid] <= r)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t* G = sysdesc + size_t(lo) * sys_words;  // divc[3], cbase at word offset 0..3
    divc0 = G[0];
    base = G[3] + (r - rbase[lo]) * divc0;
    tag = lo << 8;
  }
  auto emit = [&](uint32_t row0, uint32_t n, uint32_t cell0, uint32_t ncx) {
    uint32_t k = atomicAdd(nitems, 1u);
    items[k] = TargetItem{row0, n, cell0, ncx | tag};
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

// Sparse view: flag[c] = 1 for a non-empty cell (cell_start is the packed
// scan: low 32 bits = first sorted index).
__global__ void k_nz_flag(const uint64_t* __restrict__ cs, uint64_t n_cells, uint32_t* __restrict__ flag) {
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < n_cells) flag[c] = uint32_t(cs[c + 1]) > uint32_t(cs[c]) ? 1u : 0u;
}

// pos = exclusive scan of the flags: the non-empty cell c is entry pos[c];
// inv[c] = that entry or -1.  The entry after the last holds the particle total.
__global__ void k_nz_scatter(const uint64_t* __restrict__ cs, uint64_t n_cells, const uint32_t* __restrict__ pos,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ start, int32_t* __restrict__ inv) {
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c > n_cells) return;
  if (c == n_cells) {
    start[pos[n_cells]] = uint32_t(cs[n_cells]);
    return;
  }
  const bool nz = uint32_t(cs[c + 1]) > uint32_t(cs[c]);
  inv[c] = nz ? int32_t(pos[c]) : -1;
  if (nz) {
    key[pos[c]] = uint32_t(c);
    start[pos[c]] = uint32_t(cs[c]);
  }
}

// The (2B + 1)^3 neighbour cells of each non-empty cell, B = the list's box_nn
// (offsets dz, dy, dx in -B..B, x fastest), as entries of the sparse view, -1
// where empty or off the grid.  One thread per (cell, offset).
__global__ void k_nz_nn(const uint32_t* __restrict__ key, uint64_t n_nz, uint32_t d0, uint32_t d1, uint32_t d2,
                        int B, const int32_t* __restrict__ inv, int32_t* __restrict__ nn) {
  const int W = 2 * B + 1, K = W * W * W;
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n_nz * uint64_t(K)) return;
  const uint64_t m = t / uint64_t(K);
  const int k = int(t - m * uint64_t(K));
  const int dx = k % W - B, dy = (k / W) % W - B, dz = k / (W * W) - B;
  const uint32_t c = key[m];
  const int x = int(c % d0) + dx, y = int((c / d0) % d1) + dy, z = int(c / (uint64_t(d0) * d1)) + dz;
  const bool in = x >= 0 && y >= 0 && z >= 0 && x < int(d0) && y < int(d1) && z < int(d2);
  nn[t] = in ? inv[uint64_t(x) + uint64_t(d0) * (uint64_t(y) + uint64_t(d1) * uint64_t(z))] : -1;
}

}  // namespace ofp

using namespace ofp;

// ------------------------------------------------------- ensemble internals
int ofp::cl_create_many(ofp_cl** out, int dtype, uint32_t n_sys, const void* sys_dev, const uint32_t* rbase_dev,
                        uint32_t n_rows, uint64_t n_cells, uint32_t opt) {
  OFP_CHECK(out && sys_dev && rbase_dev && n_sys > 0, OFP_E_ARG, "cl_create_many: null argument");
  OFP_CHECK(n_cells < (1ull << 31), OFP_E_LIMIT, "too many cells (%llu)", (unsigned long long)n_cells);
  ofp_cl* h = new ofp_cl();
  h->dtype = dtype;
  h->opt = opt;
  h->n_cells = n_cells;
  h->n_sys = n_sys;
  h->sys = sys_dev;
  h->rbase = rbase_dev;
  h->n_rows = n_rows;
  *out = h;
  return OFP_OK;
}

int ofp::cl_construct_many(ofp_cl* h, const void* pos, size_t stride, const uint32_t* sid, size_t n_cap,
                           const uint32_t* n_dev, size_t gm, void* pos_sorted, size_t sorted_stride,
                           hipStream_t s) {
  OFP_CHECK(h && h->n_sys, OFP_E_ARG, "cl_construct_many: not an ensemble cell list");
  OFP_CHECK(n_cap < 0xFFFFFFF0ull, OFP_E_LIMIT, "particle capacity must fit 32 bits");
  h->sid = sid;
  if (h->dtype == OFP_F32)
    return ofp::cl_construct_many_t<float>(h, pos, stride, n_cap, n_dev, gm, pos_sorted, sorted_stride, s);
  return ofp::cl_construct_many_t<double>(h, pos, stride, n_cap, n_dev, gm, pos_sorted, sorted_stride, s);
}

int ofp::cl_sid_sorted(const ofp_cl* h, const uint32_t** sid_sorted) {
  OFP_CHECK(h && h->built && h->n_sys, OFP_E_STATE, "ensemble cell list not constructed");
  *sid_sorted = h->sid_sorted.as<uint32_t>();
  return OFP_OK;
}

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
  h->n_rows = h->divc[1] * h->divc[2];
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
  DevBuf* bufs[] = {&h->packed, &h->scan64, &h->keys,    &h->keys2, &h->s2n,        &h->vals2, &h->n2s,
                    &h->dom,    &h->quant,  &h->scantmp, &h->hist,  &h->sid_sorted, &h->items, &h->nitems,
                    &h->payload, &h->ctl, &h->colpart, &h->nz_flag, &h->nz_key, &h->nz_start, &h->nz_nn,
                    &h->nz_inv};
  for (DevBuf* b : bufs) b->release();
  delete h;
  return OFP_OK;
}

int ofp_cl_construct(ofp_cl* h, const void* pos, size_t stride, size_t n_tot, size_t gm,
                     void* pos_sorted, size_t sorted_stride, const ofp_plane* planes_in,
                     const ofp_plane* planes_out, int n_planes, void* stream) {
  OFP_CHECK(h, OFP_E_ARG, "ofp_cl_construct: null handle");
  OFP_CHECK(!h->n_sys, OFP_E_STATE, "ofp_cl_construct: ensemble cell lists are built by their ensemble");
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
    const uint32_t words = h->dtype == OFP_F32 ? uint32_t(sizeof(SysGrid<float>) / 4)
                                               : uint32_t(sizeof(SysGrid<double>) / 4);
    const uint32_t off = h->dtype == OFP_F32 ? uint32_t(offsetof(SysGrid<float>, divc) / 4)
                                             : uint32_t(offsetof(SysGrid<double>, divc) / 4);
    const uint32_t* desc = h->n_sys ? static_cast<const uint32_t*>(h->sys) + off : nullptr;
    k_cl_items<<<grid_for(h->n_rows, 256), 256, 0, s>>>(h->scan64.as<unsigned long long>(), h->divc[0],
                                                        h->n_rows, h->n_sys, h->rbase, desc, words,
                                                        h->items.as<TargetItem>(), h->nitems.as<uint32_t>());
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

int ofp_cl_aux_(const ofp_cl* h, size_t bytes, void** p) {
  OFP_CHECK(h && p, OFP_E_ARG, "ofp_cl_aux_: null argument");
  ofp_cl* m = const_cast<ofp_cl*>(h);
  OFP_TRY(m->aux.ensure(bytes));
  *p = m->aux.p;
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

int ofp_cl_build_sparse(ofp_cl* h, void* stream) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "ofp_cl_build_sparse: cell list not constructed");
  OFP_CHECK(h->n_sys == 0, OFP_E_ARG, "ofp_cl_build_sparse: ensemble cell lists are not supported");
  OFP_CHECK(h->n_cells < (uint64_t(1) << 31), OFP_E_LIMIT, "ofp_cl_build_sparse: more than 2^31 cells");
  hipStream_t s = as_stream(stream);
  const uint64_t C = h->n_cells;
  OFP_TRY(h->nz_flag.ensure((C + 1) * sizeof(uint32_t)));
  OFP_TRY(h->nz_inv.ensure((C + 1) * sizeof(int32_t)));
  OFP_TRY(h->nz_key.ensure((C + 1) * sizeof(uint32_t)));
  OFP_TRY(h->nz_start.ensure((C + 1) * sizeof(uint32_t)));
  OFP_TRY(h->scantmp.ensure(scan_tmp_bytes(C)));
  DevBuf& pos = h->nz_nn;  // holds the scanned flags first, the neighbour table after
  OFP_TRY(pos.ensure((C + 1) * sizeof(uint32_t)));
  const uint64_t* cs = h->scan64.as<uint64_t>();
  k_nz_flag<<<grid_for(C, 256), 256, 0, s>>>(cs, C, h->nz_flag.as<uint32_t>());
  OFP_LAUNCHED();
  OFP_TRY(scan_excl_u32(h->nz_flag.as<uint32_t>(), pos.as<uint32_t>(), C, h->scantmp.p, s));
  k_nz_scatter<<<grid_for(C + 1, 256), 256, 0, s>>>(cs, C, pos.as<uint32_t>(), h->nz_key.as<uint32_t>(),
                                                     h->nz_start.as<uint32_t>(), h->nz_inv.as<int32_t>());
  OFP_LAUNCHED();
  uint32_t m = 0;
  OFP_HIP(hipMemcpyAsync(&m, pos.as<uint32_t>() + C, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  OFP_HIP(hipStreamSynchronize(s));
  const uint32_t W = 2 * h->box_nn + 1, K = W * W * W;
  OFP_TRY(h->nz_nn.ensure((size_t(m) * K + 1) * sizeof(int32_t)));
  if (m > 0) {
    k_nz_nn<<<grid_for(size_t(m) * K, 256), 256, 0, s>>>(h->nz_key.as<uint32_t>(), m, h->divc[0], h->divc[1],
                                                         h->divc[2], int(h->box_nn), h->nz_inv.as<int32_t>(),
                                                         h->nz_nn.as<int32_t>());
    OFP_LAUNCHED();
  }
  h->n_nz = m;
  h->nz_box_nn = h->box_nn;
  h->sparse_valid = true;
  return OFP_OK;
}

int ofp_cl_get_sparse_view(const ofp_cl* h, ofp_cl_sparse_view* v) {
  OFP_CHECK(h && v, OFP_E_ARG, "ofp_cl_get_sparse_view: null argument");
  OFP_CHECK(h->built && h->sparse_valid, OFP_E_STATE, "ofp_cl_get_sparse_view: no sparse view of this build");
  v->cell_key = h->nz_key.as<uint32_t>();
  v->cell_start = h->nz_start.as<uint32_t>();
  v->cell_nn = h->nz_nn.as<int32_t>();
  v->n_cells = h->n_nz;
  v->box_nn = h->nz_box_nn;
  v->nn_per_cell = (2 * h->nz_box_nn + 1) * (2 * h->nz_box_nn + 1) * (2 * h->nz_box_nn + 1);
  return OFP_OK;
}

int ofp_cl_restore(const ofp_cl* h, const ofp_plane* sorted_in, const ofp_plane* orig_out,
                   int n_planes, void* stream) {
  OFP_CHECK(h && h->built, OFP_E_STATE, "ofp_cl_restore: cell list not constructed");
  OFP_CHECK(!h->n_dev, OFP_E_STATE, "ofp_cl_restore: particle count is on the device");
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
  OFP_CHECK(!h->n_dev, OFP_E_STATE, "ofp_cl_reorder: particle count is on the device");
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
