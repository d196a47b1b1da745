// Device-wide exclusive scan and stable LSD radix sort (32-bit keys, int32
// payload) for gfx950.
//
// Radix sort, per 8-bit digit pass: (1) per-tile digit histograms in LDS,
// (2) exclusive scan of the digit-major histogram table, (3) stable scatter:
// each wavefront ranks its 64 keys among equal digits with 8 ballots
// (match mask = AND over digit bits), LDS per-wave digit counts give the
// cross-wave prefix, and a running per-digit count carries the order across
// the tile's rounds.  Element order inside a tile is (round, wave, lane) =
// index order, so every pass is stable and the sort equals a stable argsort.
//
// Sizes may live on the device (n_dev != nullptr): kernels are launched for
// n_max elements and read the actual count from n_dev, so a data-dependent
// sort needs no host synchronisation.
#include <mutex>

#include "slx_common.h"

namespace slx {

constexpr int kScanBS = 256;
constexpr int kScanIPT = 16;                       // 4096 elements per tile
constexpr int kScanTile = kScanBS * kScanIPT;

constexpr int kRsBS = 256;
constexpr int kRsIPT = 8;                          // 2048 keys per tile
constexpr int kRsTile = kRsBS * kRsIPT;
constexpr int kRadix = 256;

__device__ __forceinline__ int read_n(int n_max, const int* n_dev) {
  if (!n_dev) return n_max;
  int n = *n_dev;
  return n < n_max ? (n < 0 ? 0 : n) : n_max;
}

// block-wide exclusive scan of one int per thread (256 threads); returns the
// block total through *total
__device__ int block_exclusive_scan(int v, int* s_tmp, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  int wsum = 0, tot = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
    int t = s_tmp[i];
    if (i < w) wsum += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return x - v + wsum;
}

// ---- exclusive scan: tile sums, scan of sums, tile-local scan plus offset ----
__global__ void __launch_bounds__(kScanBS)
k_scan_tiles(int n_max, const int* n_dev, const int* __restrict__ in, int* __restrict__ tile_sum) {
  __shared__ int s_tmp[4];
  const int n = read_n(n_max, n_dev);
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanIPT;
  int s = 0;
  for (int k = 0; k < kScanIPT; ++k) {
    long long i = base + k;
    if (i < n) s += in[i];
  }
  int tot;
  block_exclusive_scan(s, s_tmp, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// single block: exclusive scan of the tile sums in place, total to *total_dev
__global__ void __launch_bounds__(kScanBS)
k_scan_top(int ntiles, int* __restrict__ tile_sum, int* __restrict__ total_dev) {
  __shared__ int s_tmp[4];
  int carry = 0;
  for (int b = 0; b < ntiles; b += kScanBS) {
    int i = b + threadIdx.x;
    int v = i < ntiles ? tile_sum[i] : 0;
    int tot;
    int ex = block_exclusive_scan(v, s_tmp, &tot);
    if (i < ntiles) tile_sum[i] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_dev) *total_dev = carry;
}

__global__ void __launch_bounds__(kScanBS)
k_scan_apply(int n_max, const int* n_dev, const int* __restrict__ in, int* __restrict__ out,
             const int* __restrict__ tile_off) {
  __shared__ int s_tmp[4];
  const int n = read_n(n_max, n_dev);
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanIPT;
  int v[kScanIPT];
  int s = 0;
  for (int k = 0; k < kScanIPT; ++k) {
    long long i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int tot;
  int run = block_exclusive_scan(s, s_tmp, &tot) + tile_off[blockIdx.x];
  for (int k = 0; k < kScanIPT; ++k) {
    long long i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

size_t scan_ws_bytes(int n_max) {
  int tiles = blocks_for(n_max > 0 ? n_max : 1, kScanTile);
  return sizeof(int) * (size_t)(tiles + 64);
}

int exclusive_scan(int n_max, const int* n_dev, const int* in, int* out, int* total_dev, void* ws,
                   size_t ws_bytes, hipStream_t st) {
  if (n_max <= 0) {
    if (total_dev) {
      // total = 0 through a one-thread scan of zero tiles
      SLX_LAUNCH("scan_top", k_scan_top, dim3(1), dim3(kScanBS), 0, st, 0, (int*)ws, total_dev);
    }
    return check_launch("exclusive_scan(empty)");
  }
  if (ws_bytes < scan_ws_bytes(n_max)) return fail(SLX_EWS, "exclusive_scan: workspace too small");
  int tiles = blocks_for(n_max, kScanTile);
  int* tile_sum = (int*)ws;
  SLX_LAUNCH("scan_tiles", k_scan_tiles, dim3(tiles), dim3(kScanBS), 0, st, n_max, n_dev, in, tile_sum);
  SLX_LAUNCH("scan_top", k_scan_top, dim3(1), dim3(kScanBS), 0, st, tiles, tile_sum, total_dev);
  SLX_LAUNCH("scan_apply", k_scan_apply, dim3(tiles), dim3(kScanBS), 0, st, n_max, n_dev, in, out, tile_sum);
  return check_launch("exclusive_scan");
}

// ---- radix sort ----

__global__ void __launch_bounds__(kRsBS)
k_rs_hist(int n_max, const int* n_dev, const uint32_t* __restrict__ keys, int shift,
          int* __restrict__ hist, int ntiles) {
  __shared__ int s_h[kRadix];
  const int n = read_n(n_max, n_dev);
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const long long tb = (long long)blockIdx.x * kRsTile;
  for (int r = 0; r < kRsIPT; ++r) {
    long long i = tb + (long long)r * kRsBS + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & (kRadix - 1)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

__global__ void __launch_bounds__(kRsBS)
k_rs_scatter(int n_max, const int* n_dev, const uint32_t* __restrict__ kin,
             const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
             int32_t* __restrict__ vout, int shift, const int* __restrict__ hist_scan, int ntiles) {
  __shared__ int s_base[kRadix];
  __shared__ int s_run[kRadix];
  __shared__ int s_wave[kRsBS / 64][kRadix];
  const int n = read_n(n_max, n_dev);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  s_base[tid] = hist_scan[tid * ntiles + blockIdx.x];
  s_run[tid] = 0;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const long long tb = (long long)blockIdx.x * kRsTile;
  for (int r = 0; r < kRsIPT; ++r) {
#pragma unroll
    for (int q = 0; q < kRsBS / 64; ++q) s_wave[q][tid] = 0;
    __syncthreads();
    const long long i = tb + (long long)r * kRsBS + tid;
    const bool valid = i < n;
    uint32_t key = 0;
    int32_t val = 0;
    int d = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      d = (int)((key >> shift) & (kRadix - 1));
    }
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const int rank = __popcll(m & lt);
    if (valid && rank == 0) s_wave[w][d] = __popcll(m);
    __syncthreads();
    {
      int run = s_run[tid];
#pragma unroll
      for (int q = 0; q < kRsBS / 64; ++q) {
        int c = s_wave[q][tid];
        s_wave[q][tid] = run;
        run += c;
      }
      s_run[tid] = run;
    }
    __syncthreads();
    if (valid) {
      const int pos = s_base[d] + s_wave[w][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
  }
}

__global__ void k_copy_pairs(int n_max, const int* n_dev, const uint32_t* __restrict__ ks,
                             const int32_t* __restrict__ vs, uint32_t* __restrict__ kd,
                             int32_t* __restrict__ vd) {
  const int n = read_n(n_max, n_dev);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

size_t radix_sort_ws_bytes(int n_max) {
  size_t n = (size_t)(n_max > 0 ? n_max : 1);
  size_t tiles = (size_t)blocks_for((long long)n, kRsTile);
  size_t hist = (size_t)kRadix * tiles;
  return 2 * n * 4 + 2 * hist * 4 + scan_ws_bytes((int)hist) + 256;
}

// Sorts (keys, vals) in place by the low key_bits bits of keys, stably.
int radix_sort_pairs(int n_max, const int* n_dev, uint32_t* keys, int32_t* vals, int key_bits,
                     void* ws, size_t ws_bytes, hipStream_t st) {
  if (n_max <= 0) return SLX_OK;
  if (key_bits <= 0) return SLX_OK;
  if (ws_bytes < radix_sort_ws_bytes(n_max)) return fail(SLX_EWS, "radix_sort: workspace too small");
  const int tiles = blocks_for(n_max, kRsTile);
  const int hist_n = kRadix * tiles;
  char* p = (char*)ws;
  uint32_t* k2 = (uint32_t*)p;
  p += sizeof(uint32_t) * (size_t)n_max;
  int32_t* v2 = (int32_t*)p;
  p += sizeof(int32_t) * (size_t)n_max;
  int* hist = (int*)p;
  p += sizeof(int) * (size_t)hist_n;
  int* hscan = (int*)p;
  p += sizeof(int) * (size_t)hist_n;
  void* sws = p;
  size_t sws_bytes = scan_ws_bytes(hist_n);
  const int passes = (key_bits + 7) / 8;
  uint32_t *ka = keys, *kb = k2;
  int32_t *va = vals, *vb = v2;
  for (int ps = 0; ps < passes; ++ps) {
    const int shift = 8 * ps;
    SLX_LAUNCH("rs_hist", k_rs_hist, dim3(tiles), dim3(kRsBS), 0, st, n_max, n_dev, ka, shift, hist, tiles);
    int rc = exclusive_scan(hist_n, nullptr, hist, hscan, nullptr, sws, sws_bytes, st);
    if (rc) return rc;
    SLX_LAUNCH("rs_scatter", k_rs_scatter, dim3(tiles), dim3(kRsBS), 0, st, n_max, n_dev, ka, va, kb, vb,
               shift, hscan, tiles);
    uint32_t* tk = ka; ka = kb; kb = tk;
    int32_t* tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    SLX_LAUNCH("rs_copy", k_copy_pairs, dim3(blocks_for(n_max, 256)), dim3(256), 0, st, n_max, n_dev, ka, va,
               keys, vals);
  }
  return check_launch("radix_sort_pairs");
}

// ---- radix sort without allocation ----
// The double buffer, digit histograms and scan tile sums of a sort of up to
// kNoAllocPairs pairs, for callers that cannot allocate (code under graph
// capture, serving loops).  One copy per GPU, held until the process ends
// (18.9 MB).  Host code never passes the array's name to a kernel: it is a
// host-side stand-in, and each device's copy is reached through the address
// hipGetSymbolAddress returns on that device.
constexpr int kNoAllocPairs = 1 << 21;
constexpr size_t noalloc_scratch_bytes() {
  size_t n = (size_t)kNoAllocPairs;
  size_t tiles = (n + kRsTile - 1) / kRsTile;
  size_t hist = (size_t)kRadix * tiles;
  size_t scan_tiles = (hist + kScanTile - 1) / kScanTile;
  return 2 * n * 4 + 2 * hist * 4 + sizeof(int) * (scan_tiles + 64) + 256;
}
constexpr size_t kNoAllocScratchBytes = noalloc_scratch_bytes();

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up once per device
int scratch_on_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SLX_EHIP, "slx_radix_sort_pairs_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(SLX_EHIP, "slx_radix_sort_pairs_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SLX_EHIP, "slx_radix_sort_pairs_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SLX_OK;
}

}  // namespace slx

using namespace slx;

extern "C" size_t slx_scan_ws_bytes(int n_max) { return scan_ws_bytes(n_max); }

extern "C" int slx_exclusive_scan(int n_max, const int32_t* n_dev, const int32_t* in, int32_t* out,
                                  int32_t* total_dev, void* ws, size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(n_max >= 0, "slx_exclusive_scan: n_max < 0");
  SLX_REQUIRE(n_max == 0 || (in && out && ws), "slx_exclusive_scan: null pointer");
  return exclusive_scan(n_max, n_dev, in, out, total_dev, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" size_t slx_radix_sort_ws_bytes(int n_max) { return radix_sort_ws_bytes(n_max); }

extern "C" int slx_radix_sort_pairs(int n_max, const int32_t* n_dev, uint32_t* keys, int32_t* vals,
                                    int key_bits, void* ws, size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(n_max >= 0 && key_bits >= 0 && key_bits <= 32, "slx_radix_sort_pairs: bad sizes");
  SLX_REQUIRE(n_max == 0 || (keys && vals && ws), "slx_radix_sort_pairs: null pointer");
  return radix_sort_pairs(n_max, n_dev, keys, vals, key_bits, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int slx_radix_sort_pairs_noalloc(int n_max, const int32_t* n_dev, uint32_t* keys, int32_t* vals,
                                            int key_bits, slx_stream_t stream) {
  SLX_REQUIRE(n_max >= 0 && key_bits >= 0 && key_bits <= 32, "slx_radix_sort_pairs_noalloc: bad sizes");
  SLX_REQUIRE(n_max <= kNoAllocPairs, "slx_radix_sort_pairs_noalloc: n_max %d exceeds the scratch limit of %d pairs",
              n_max, kNoAllocPairs);
  SLX_REQUIRE(n_max == 0 || (keys && vals), "slx_radix_sort_pairs_noalloc: null pointer");
  if (n_max == 0 || key_bits == 0) return SLX_OK;
  const size_t need = radix_sort_ws_bytes(n_max);
  if (need > kNoAllocScratchBytes) return fail(SLX_EWS, "slx_radix_sort_pairs_noalloc: scratch too small");
  void* ws = nullptr;
  int rc = scratch_on_device(&ws);
  if (rc) return rc;
  return radix_sort_pairs(n_max, n_dev, keys, vals, key_bits, ws, kNoAllocScratchBytes, (hipStream_t)stream);
}
