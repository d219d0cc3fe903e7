// Device-wide exclusive prefix sum (three phases: per-block scan, scan of
// block totals, add-back), used for ragged offsets (ghosts, boxes, pairs).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_buffer.hh"

namespace rgpu {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class In>
__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const In* __restrict__ in,
                                                                  int64_t* __restrict__ out,
                                                                  int64_t* __restrict__ tile_sums,
                                                                  int64_t n) {
  __shared__ int64_t wsum[kScanThreads / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int64_t v[kScanItems];
  int64_t tsum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t idx = base + k;
    v[k] = idx < n ? int64_t(in[idx]) : 0;
    tsum += v[k];
  }
  // inclusive scan of thread sums within the wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t incl = tsum;
  for (int off = 1; off < 64; off <<= 1) {
    int64_t o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int64_t wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += wsum[w];
  int64_t run = wbase + incl - tsum;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = wbase + incl;
}

__global__ static void scan_add_kernel(int64_t* __restrict__ out, const int64_t* __restrict__ tile_off,
                                       int64_t n) {
  int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += tile_off[idx / kScanTile];
}

// out[i] = sum_{k<i} in[k] for i < n, and out[n] = total (out has n+1
// elements).  Returns the total (synchronises the stream).
template <class In>
int64_t exclusive_scan(const In* in, int64_t* out, int64_t n, hipStream_t stream,
                       DeviceBuffer<int64_t>& scratch) {
  if (n == 0) {
    RGPU_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), stream));
    RGPU_HIP(hipStreamSynchronize(stream));
    return 0;
  }
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  // level storage: tile sums and their scan
  std::vector<int64_t> level_sizes;
  int64_t t = tiles;
  size_t need = 0;
  while (true) {
    level_sizes.push_back(t);
    need += size_t(2 * t + 1);
    if (t <= 1) break;
    t = (t + kScanTile - 1) / kScanTile;
  }
  int64_t* buf = scratch.reserve(need);
  int64_t* sums = buf;
  int64_t* sums_scan = buf + tiles;
  hipLaunchKernelGGL(scan_tiles_kernel<In>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, stream, in,
                     out, sums, n);
  RGPU_LAUNCH_CHECK();
  int64_t total;
  if (tiles > 1) {
    DeviceBuffer<int64_t> inner;
    total = exclusive_scan<int64_t>(sums, sums_scan, tiles, stream, inner);
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(scan_add_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, out, sums_scan,
                       n);
    RGPU_LAUNCH_CHECK();
  } else {
    RGPU_HIP(hipMemcpyAsync(&total, sums, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    RGPU_HIP(hipStreamSynchronize(stream));
  }
  RGPU_HIP(hipMemcpyAsync(out + n, &total, sizeof(int64_t), hipMemcpyHostToDevice, stream));
  RGPU_HIP(hipStreamSynchronize(stream));
  return total;
}

}  // namespace rgpu
