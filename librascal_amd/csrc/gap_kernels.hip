// Sparse GAP kernel: LDS-tiled GEMMs with the zeta-power / energy / weight
// epilogue fused into the K_NM tile (generic precision path).
#include <hip/hip_runtime.h>

#include "common.hh"
#include "gap_kernels.hh"

namespace rgpu {

namespace {

constexpr int kTM = 64, kTN = 64, kTK = 16, kGT = 256;

template <class T>
__device__ __forceinline__ T pow_zeta_dev(T v, int zeta) {
  // models/kernels.hh:48-63 (integer power by repeated squaring)
  if (zeta == 1) return v;
  if (zeta == 2) return v * v;
  if (zeta == 3) return v * v * v;
  T r = T(1), x = v;
  unsigned n = unsigned(zeta);
  do {
    if (n & 1) r *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return r;
}

template <class T>
__global__ __launch_bounds__(kGT) void gap_kernel_tile(GapBlock<T> blk, int64_t M, double* energy,
                                                       T* wt, double* knm_atom) {
  __shared__ T As[kTK][kTM + 1];
  __shared__ T Bs[kTK][kTN + 1];
  __shared__ double esum[kTM][kTN / 4 + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int64_t r0 = int64_t(blockIdx.y) * kTM, c0 = int64_t(blockIdx.x) * kTN;
  T acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = T(0);
  const int lr = threadIdx.x / 4, lk = (threadIdx.x % 4) * 4;
  const int64_t arow = r0 + lr, bcol = c0 + lr;
  const T* Ap = arow < blk.n_rows ? blk.X + int64_t(blk.row_idx[arow]) * blk.D : nullptr;
  const T* Bp = bcol < blk.n_cols ? blk.Tm + (blk.col0 + bcol) * blk.D : nullptr;
  for (int64_t k0 = 0; k0 < blk.D; k0 += kTK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t k = k0 + lk + q;
      As[lk + q][lr] = (Ap && k < blk.D) ? Ap[k] : T(0);
      Bs[lk + q][lr] = (Bp && k < blk.D) ? Bp[k] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[kk][ty * 4 + u];
        b[u] = Bs[kk][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
  // epilogue
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t r = r0 + ty * 4 + u;
    double erow = 0;
    if (r < blk.n_rows) {
      const int c = blk.row_idx[r];
      double inv = 1.0;
      if (blk.normalize) {
        const double n2 = blk.norm2[c];
        inv = n2 > 0 ? 1.0 / sqrt(n2) : 1.0;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t m = c0 + tx * 4 + v;
        if (m >= blk.n_cols) continue;
        const T k = T(double(acc[u][v]) * inv);
        const T kz = pow_zeta_dev(k, blk.zeta);
        const int64_t mg = blk.col0 + m;
        if (blk.alpha) {
          const double a = blk.alpha[mg];
          erow += a * double(kz);
          if (wt) {
            const T g = blk.zeta > 1 ? T(blk.zeta) * pow_zeta_dev(k, blk.zeta - 1) : T(1);
            wt[r * blk.n_cols + m] = T(a) * g;
          }
        }
        if (knm_atom) knm_atom[int64_t(c) * M + mg] = double(kz);
      }
    }
    esum[ty * 4 + u][tx] = erow;
  }
  if (energy) {
    __syncthreads();
    if (threadIdx.x < kTM) {
      const int64_t r = r0 + threadIdx.x;
      if (r < blk.n_rows) {
        double t = 0;
        for (int x = 0; x < 16; ++x) t += esum[threadIdx.x][x];
        atomicAdd(energy + blk.row_idx[r], t);
      }
    }
  }
}

template <class T>
__global__ __launch_bounds__(kGT) void gap_force_tile(GapBlock<T> blk, const T* __restrict__ wt,
                                                      T* __restrict__ G) {
  __shared__ T As[kTK][kTM + 1];
  __shared__ T Bs[kTK][kTN + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int64_t r0 = int64_t(blockIdx.y) * kTM, d0 = int64_t(blockIdx.x) * kTN;
  T acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = T(0);
  const int lr = threadIdx.x / 4, lk = (threadIdx.x % 4) * 4;  // A: row, k
  const int bk = threadIdx.x / 16, bd = (threadIdx.x % 16) * 4;  // B: k, d
  const int64_t K = blk.n_cols;
  for (int64_t k0 = 0; k0 < K; k0 += kTK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = r0 + lr, k = k0 + lk + q;
      As[lk + q][lr] = (r < blk.n_rows && k < K) ? wt[r * K + k] : T(0);
      const int64_t kb = k0 + bk, d = d0 + bd + q;
      Bs[bk][bd + q] = (kb < K && d < blk.D) ? blk.Tm[(blk.col0 + kb) * blk.D + d] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[kk][ty * 4 + u];
        b[u] = Bs[kk][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t r = r0 + ty * 4 + u;
    if (r >= blk.n_rows) continue;
    T* g = G + int64_t(blk.row_idx[r]) * blk.D;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t d = d0 + tx * 4 + v;
      if (d < blk.D) g[d] = acc[u][v];
    }
  }
}

}  // namespace

template <class T>
void launch_gap_kernel(const GapBlock<T>& blk, int64_t M, double* energy, T* wt, double* knm_atom,
                       hipStream_t stream) {
  if (blk.n_rows == 0 || blk.n_cols == 0) return;
  dim3 grid(unsigned((blk.n_cols + kTN - 1) / kTN), unsigned((blk.n_rows + kTM - 1) / kTM));
  hipLaunchKernelGGL(gap_kernel_tile<T>, grid, dim3(kGT), 0, stream, blk, M, energy, wt, knm_atom);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_gap_force_contraction(const GapBlock<T>& blk, const T* wt, T* G, hipStream_t stream) {
  if (blk.n_rows == 0 || blk.n_cols == 0) return;
  dim3 grid(unsigned((blk.D + kTN - 1) / kTN), unsigned((blk.n_rows + kTM - 1) / kTM));
  hipLaunchKernelGGL(gap_force_tile<T>, grid, dim3(kGT), 0, stream, blk, wt, G);
  RGPU_LAUNCH_CHECK();
}

namespace {

__global__ void energy_tiles_kernel(EnergyTiles et, const double* __restrict__ part, int64_t ld,
                                    double* __restrict__ energy) {
  int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int pi = 0;
  while (pi < et.n && r >= et.m[pi]) r -= et.m[pi++];
  if (pi >= et.n) return;
  const int c = et.rows[pi][r];
  double s = 0.0;
  for (int t = 0; t < et.tiles[pi]; ++t) s += part[int64_t(t) * ld + c];
  energy[c] += s;
}

// out[s][m] = sum of in[a][m] over the rows a of segment s, in row order
__global__ void rows_segment_sum_kernel(const double* __restrict__ in, int64_t w,
                                        const int64_t* __restrict__ seg_off, double* __restrict__ out) {
  const int64_t m = int64_t(blockIdx.y) * blockDim.x + threadIdx.x, s = blockIdx.x;
  if (m >= w) return;
  double t = 0.0;
  for (int64_t a = seg_off[s]; a < seg_off[s + 1]; ++a) t += in[a * w + m];
  out[s * w + m] = t;
}

}  // namespace

void launch_rows_segment_sum(const double* in, int64_t w, const int64_t* seg_off, int n_seg, double* out,
                             hipStream_t stream) {
  if (w == 0 || n_seg == 0) return;
  // segments along x (no 65535 limit), column chunks along y
  if ((w + 255) / 256 > 65535) throw InvalidArgument("rows of more than 16776960 columns");
  hipLaunchKernelGGL(rows_segment_sum_kernel, dim3(unsigned(n_seg), unsigned((w + 255) / 256)), dim3(256), 0,
                     stream, in, w, seg_off, out);
  RGPU_LAUNCH_CHECK();
}

void launch_energy_tiles(const EnergyTiles& et, const double* part, int64_t ld, double* energy,
                         hipStream_t stream) {
  int64_t rows = 0;
  for (int i = 0; i < et.n; ++i) rows += et.m[i];
  if (rows == 0) return;
  hipLaunchKernelGGL(energy_tiles_kernel, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, stream, et, part,
                     ld, energy);
  RGPU_LAUNCH_CHECK();
}

template void launch_gap_kernel<double>(const GapBlock<double>&, int64_t, double*, double*, double*,
                                        hipStream_t);
template void launch_gap_kernel<float>(const GapBlock<float>&, int64_t, double*, float*, double*, hipStream_t);
template void launch_gap_force_contraction<double>(const GapBlock<double>&, const double*, double*,
                                                   hipStream_t);
template void launch_gap_force_contraction<float>(const GapBlock<float>&, const float*, float*,
                                                  hipStream_t);

}  // namespace rgpu
