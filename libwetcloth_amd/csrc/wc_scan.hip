// Device-wide exclusive prefix sum of uint32 (three-phase: per-block scan,
// recursive scan of block totals, add-back).  Used for the bucket offsets of
// the particle sort, pressure-DOF numbering in bucket-major order (the
// std::partial_sum of Pressure.cpp:421) and seeding compaction.
#include "wc_ctx.h"

namespace wc {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_tile(uint32_t* data,
                                                            int64_t n,
                                                            uint32_t* bsum) {
  __shared__ uint32_t tile[kScanTile + kScanTile / 32];
  __shared__ uint32_t tsum[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  auto pad = [](int i) { return i + (i >> 5); };
  for (int q = 0; q < kScanItems; ++q) {
    const int li = q * kScanThreads + threadIdx.x;
    const int64_t gi = base + li;
    tile[pad(li)] = gi < n ? data[gi] : 0u;
  }
  __syncthreads();
  uint32_t run = 0;
  const int my0 = threadIdx.x * kScanItems;
  for (int q = 0; q < kScanItems; ++q) {
    const uint32_t v = tile[pad(my0 + q)];
    tile[pad(my0 + q)] = run;
    run += v;
  }
  tsum[threadIdx.x] = run;
  __syncthreads();
  // Hillis-Steele inclusive scan of the thread totals
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t add = (int)threadIdx.x >= off ? tsum[threadIdx.x - off] : 0u;
    __syncthreads();
    tsum[threadIdx.x] += add;
    __syncthreads();
  }
  const uint32_t excl = threadIdx.x ? tsum[threadIdx.x - 1] : 0u;
  for (int q = 0; q < kScanItems; ++q) tile[pad(my0 + q)] += excl;
  __syncthreads();
  for (int q = 0; q < kScanItems; ++q) {
    const int li = q * kScanThreads + threadIdx.x;
    const int64_t gi = base + li;
    if (gi < n) data[gi] = tile[pad(li)];
  }
  if (threadIdx.x == kScanThreads - 1) bsum[blockIdx.x] = tsum[kScanThreads - 1];
}

__global__ void k_scan_add(uint32_t* data, int64_t n, const uint32_t* boff) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  data[i] += boff[i / kScanTile];
}

void scan_rec(hipStream_t st, uint32_t* data, int64_t n, uint32_t* scratch,
              uint32_t* total_dev) {
  const int64_t nblk = (n + kScanTile - 1) / kScanTile;
  uint32_t* bsum = scratch;
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)nblk), dim3(kScanThreads), 0,
                     st, data, n, bsum);
  WC_LAUNCH_CHECK();
  if (nblk == 1) {
    WC_HIP(hipMemcpyAsync(total_dev, bsum, sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, st));
    return;
  }
  scan_rec(st, bsum, nblk, scratch + nblk + 1, total_dev);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, st, data, n, bsum);
  WC_LAUNCH_CHECK();
}

}  // namespace

void exclusive_scan_async(hipStream_t st, DevBuf<uint32_t>& scratch, uint32_t* data, int64_t n) {
  if (n <= 0) return;
  // scratch: block sums of every recursion level plus the total
  int64_t need = 2;
  for (int64_t m = n; m > 1; m = (m + kScanTile - 1) / kScanTile) need += m / kScanTile + 2;
  scratch.ensure((size_t)need + 8);
  scan_rec(st, data, n, scratch.p + 2, scratch.p);
}

uint32_t exclusive_scan(hipStream_t st, DevBuf<uint32_t>& scratch, uint32_t* data, int64_t n) {
  if (n <= 0) return 0;
  exclusive_scan_async(st, scratch, data, n);
  const uint32_t* total_dev = scratch.p;  // left there by the scan
  uint32_t total = 0;
  WC_HIP(hipMemcpyAsync(&total, total_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  WC_HIP(hipStreamSynchronize(st));
  return total;
}

uint32_t exclusive_scan(wc_ctx* ctx, uint32_t* data, int64_t n) {
  return exclusive_scan(ctx->stream, ctx->scan_tmp, data, n);
}

}  // namespace wc
