// Device-wide exclusive prefix sum (three phases: per-block scan, scan of
// block totals, add-back), used for ragged offsets (ghosts, boxes, pairs).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

// Exclusive scan of per-tile sums by one workgroup, chunk by chunk with a
// running carry; out_excl[i] = sum_{k<i} in[k], *total = sum of all.
__global__ static __launch_bounds__(kScanThreads) void scan_small_kernel(
    const int64_t* __restrict__ in, int64_t* __restrict__ out_excl, int64_t n,
    int64_t* __restrict__ total) {
  __shared__ int64_t wsum[kScanThreads / 64];
  __shared__ int64_t carry_sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_sh = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < n; c0 += kScanThreads) {
    const int64_t idx = c0 + threadIdx.x;
    const int64_t v = idx < n ? in[idx] : 0;
    int64_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int64_t wbase = carry_sh;
    for (int w = 0; w < wave; ++w) wbase += wsum[w];
    if (idx < n) out_excl[idx] = wbase + incl - v;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry_sh = wbase + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry_sh;
}

// Whole scan in one workgroup of 1024 threads (8 items each per chunk,
// running carry between chunks), out[n] = total; for short arrays, where
// one launch beats three.
constexpr int kScan1Threads = 1024;
template <class In>
__global__ __launch_bounds__(kScan1Threads) void scan_single_kernel(const In* __restrict__ in,
                                                                    int64_t* __restrict__ out,
                                                                    int64_t n) {
  __shared__ int64_t wsum[kScan1Threads / 64];
  __shared__ int64_t carry_sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_sh = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < n; c0 += int64_t(kScan1Threads) * kScanItems) {
    const int64_t base = c0 + int64_t(threadIdx.x) * kScanItems;
    int64_t v[kScanItems];
    int64_t tsum = 0;
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? int64_t(in[base + k]) : 0;
      tsum += v[k];
    }
    int64_t incl = tsum;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int64_t run = carry_sh;
    for (int w = 0; w < wave; ++w) run += wsum[w];
    const int64_t last = run + incl;  // inclusive end of this thread
    run += incl - tsum;
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) out[base + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == kScan1Threads - 1) carry_sh = last;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry_sh;
}

// scratch that exclusive_scan_device(n) uses (reserved ahead of a graph
// capture, so that nothing is allocated inside it)
inline void exclusive_scan_reserve(int64_t n, DeviceBuffer<int64_t>& scratch) {
  if (n > 8 * int64_t(kScan1Threads) * kScanItems) scratch.reserve(size_t(2 * ((n + kScanTile - 1) / kScanTile)));
}

// out[i] = sum_{k<i} in[k] for i < n and out[n] = total, all on the device
// (no synchronisation); out has n+1 elements.
template <class In>
void exclusive_scan_device(const In* in, int64_t* out, int64_t n, hipStream_t stream,
                           DeviceBuffer<int64_t>& scratch) {
  if (n == 0) {
    RGPU_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), stream));
    return;
  }
  if (n <= 8 * int64_t(kScan1Threads) * kScanItems) {
    hipLaunchKernelGGL(scan_single_kernel<In>, dim3(1), dim3(kScan1Threads), 0, stream, in, out, n);
    RGPU_LAUNCH_CHECK();
    return;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  int64_t* sums = scratch.reserve(size_t(2 * tiles));
  int64_t* sums_scan = sums + tiles;
  hipLaunchKernelGGL(scan_tiles_kernel<In>, dim3(unsigned(tiles)), dim3(kScanThreads), 0, stream, in,
                     out, sums, n);
  RGPU_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(kScanThreads), 0, stream, sums, sums_scan, tiles,
                     out + n);
  RGPU_LAUNCH_CHECK();
  if (tiles > 1) {
    hipLaunchKernelGGL(scan_add_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, out,
                       sums_scan, n);
    RGPU_LAUNCH_CHECK();
  }
}

}  // namespace rgpu
