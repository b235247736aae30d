// ofp_common.hip -- error state, plane arguments, fills and exclusive scans.
//
// The scans are three-phase (block reduce -> scan of block sums in one block ->
// block scan + offset): no inter-block waiting, so no kernel can spin.
#include "ofp_internal.hpp"

#include <atomic>
#include <cstring>

namespace ofp {

static std::atomic<uint64_t> g_realloc_epoch{0};

uint64_t realloc_epoch() { return g_realloc_epoch.load(std::memory_order_acquire); }
void bump_realloc_epoch() { g_realloc_epoch.fetch_add(1, std::memory_order_acq_rel); }

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

int make_plane_args(const ofp_plane* in, const ofp_plane* out, int n, PlaneArgs* a) {
  OFP_CHECK(n >= 0 && n <= kMaxPlanes, OFP_E_LIMIT, "at most %d planes per call (got %d)",
            kMaxPlanes, n);
  OFP_CHECK(n == 0 || in != nullptr, OFP_E_ARG, "null plane list");
  std::memset(a, 0, sizeof(*a));
  a->n = n;
  for (int k = 0; k < n; ++k) {
    OFP_CHECK(in[k].elem_bytes == 4 || in[k].elem_bytes == 8, OFP_E_ARG,
              "plane %d: elem_bytes must be 4 or 8 (got %u)", k, in[k].elem_bytes);
    OFP_CHECK(in[k].ptr != nullptr, OFP_E_ARG, "plane %d: null pointer", k);
    a->src[k] = reinterpret_cast<uint64_t>(in[k].ptr);
    a->esz[k] = in[k].elem_bytes;
    if (out) {
      OFP_CHECK(out[k].ptr != nullptr && out[k].elem_bytes == in[k].elem_bytes, OFP_E_ARG,
                "plane %d: output pointer null or element size differs", k);
      a->dst[k] = reinterpret_cast<uint64_t>(out[k].ptr);
    } else {
      a->dst[k] = a->src[k];
    }
  }
  return OFP_OK;
}

// ------------------------------------------------------------------ fills
template <class U>
__global__ void k_fill(U* p, U v, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t step = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += step) p[i] = v;
}

template <class U>
static int fill_impl(U* p, U v, size_t n, hipStream_t s) {
  if (n == 0) return OFP_OK;
  unsigned g = grid_for(n, 256);
  if (g > 8192) g = 8192;
  k_fill<U><<<g, 256, 0, s>>>(p, v, n);
  OFP_LAUNCHED();
  return OFP_OK;
}

int fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) { return fill_impl(p, v, n, s); }
int fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t s) { return fill_impl(p, v, n, s); }

// ------------------------------------------------------------------ scans
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr size_t kScanTile = size_t(kScanBlock) * kScanItems;  // 4096 per block

size_t scan_tmp_bytes(size_t n) {
  size_t nb = (n + kScanTile - 1) / kScanTile;
  return (nb + 1) * sizeof(uint64_t) * 2 + 256;
}

// Block-wide exclusive scan of one value per thread; returns the block total in
// *total (all threads).
template <class V>
__device__ __forceinline__ V block_excl_scan(V v, V* lds, V* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  constexpr int nw = kScanBlock / kWave;
  V inc = wave_incl_scan(v);
  if (lane == kWave - 1) lds[w] = inc;
  __syncthreads();
  if (w == 0) {
    V t = lane < nw ? lds[lane] : V(0);
    V ti = wave_incl_scan(t);
    if (lane < nw) lds[lane] = ti - t;
    if (lane == nw - 1) lds[nw] = ti;
  }
  __syncthreads();
  V r = lds[w] + inc - v;
  *total = lds[nw];
  __syncthreads();
  return r;
}

template <class In, class V>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const In* in, size_t n,
                                                            V* partial) {
  __shared__ V lds[kScanBlock / kWave + 1];
  size_t base = size_t(blockIdx.x) * kScanTile;
  V acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    size_t i = base + size_t(k) * kScanBlock + threadIdx.x;
    if (i < n) acc += V(in[i]);
  }
  V tot;
  (void)block_excl_scan<V>(acc, lds, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One block scans the nb partial sums in place (exclusive): each thread sums a
// contiguous segment of them, the block scans the segment sums, and each
// thread writes its segment's prefixes.  The segment loads are independent, so
// a thread has several in flight.
template <class V>
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(V* partial, size_t nb) {
  __shared__ V lds[kScanBlock / kWave + 1];
  const size_t per = (nb + kScanBlock - 1) / kScanBlock;
  const size_t a = size_t(threadIdx.x) * per;
  const size_t b = a + per < nb ? a + per : nb;
  V sum = 0;
#pragma unroll 8
  for (size_t i = a; i < b; ++i) sum += partial[i];
  V tot;
  V run = block_excl_scan<V>(sum, lds, &tot);
#pragma unroll 8
  for (size_t i = a; i < b; ++i) {
    const V x = partial[i];
    partial[i] = run;
    run += x;
  }
  if (threadIdx.x == 0) partial[nb] = tot;
}

// Each thread owns kScanItems consecutive elements, read and written as whole
// 16-byte vectors in full tiles (the arrays are at least 4-byte aligned, which
// gfx950's multi-dword loads and stores need).
template <class In, class V>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const In* in, V* out, size_t n,
                                                           const V* partial, size_t nb) {
  __shared__ V lds[kScanBlock / kWave + 1];
  const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
  const bool full = base + kScanItems <= n;
  In x[kScanItems];
  if (full) {
    __builtin_memcpy(x, in + base, sizeof(x));
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) x[k] = base + k < n ? in[base + k] : In(0);
  }
  V acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) acc += V(x[k]);
  V tot;
  const V ex = block_excl_scan<V>(acc, lds, &tot);
  V run = ex + partial[blockIdx.x];
  V y[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    y[k] = run;
    run += V(x[k]);
  }
  if (full) {
    __builtin_memcpy(out + base, y, sizeof(y));
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < n) out[base + k] = y[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[nb];
}

template <class In, class V>
static int scan_impl(const In* in, V* out, size_t n, void* tmp, hipStream_t s) {
  size_t nb = (n + kScanTile - 1) / kScanTile;
  V* partial = reinterpret_cast<V*>(tmp);
  if (n == 0) {
    return fill_impl<V>(out, V(0), 1, s);
  }
  k_scan_reduce<In, V><<<unsigned(nb), kScanBlock, 0, s>>>(in, n, partial);
  OFP_LAUNCHED();
  k_scan_partials<V><<<1, kScanBlock, 0, s>>>(partial, nb);
  OFP_LAUNCHED();
  k_scan_apply<In, V><<<unsigned(nb), kScanBlock, 0, s>>>(in, out, n, partial, nb);
  OFP_LAUNCHED();
  return OFP_OK;
}

int scan_excl_u32(const uint32_t* in, uint32_t* out, size_t n, void* tmp, hipStream_t s) {
  return scan_impl<uint32_t, uint32_t>(in, out, n, tmp, s);
}
int scan_excl_u64(const uint64_t* in, uint64_t* out, size_t n, void* tmp, hipStream_t s) {
  return scan_impl<uint64_t, uint64_t>(in, out, n, tmp, s);
}
int scan_excl_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, void* tmp,
                         hipStream_t s) {
  return scan_impl<uint32_t, uint64_t>(in, out, n, tmp, s);
}

}  // namespace ofp

extern "C" {

const char* ofp_last_error(void) { return ofp::g_last_error.c_str(); }

int ofp_abi_version(void) { return OFP_ABI_VERSION; }

int ofp_device_info(int* device, int* n_cu, char* arch, size_t arch_len) {
  int dev = 0;
  OFP_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  OFP_HIP(hipGetDeviceProperties(&prop, dev));
  if (device) *device = dev;
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (arch && arch_len) {
    std::strncpy(arch, prop.gcnArchName, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return OFP_OK;
}

}  // extern "C"
