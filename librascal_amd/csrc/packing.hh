// Packed layout of the power spectrum used by the mixed-precision path.
//
// Diagonal keys (a == a) hold symmetric blocks, p[n1][n2][l] = p[n2][n1][l].
// The packed row keeps an off-diagonal key (a < b) as its full N*N*L1 block
// and a diagonal key as the N(N+1)/2 pairs n1 <= n2 (row-major) times L1,
// the n1 < n2 entries multiplied by sqrt 2.  The dot product of two packed
// rows then equals the dot product of the full rows of a symmetric feature
// vector with any other vector t whose diagonal-key entries are packed as
// (t[n1][n2] + t[n2][n1]) / sqrt 2 (n1 < n2) and t[n][n]: the GAP kernel
// X . T and its gradients are computed on 14400 -> 11760 columns for four
// species, n_max 12, l_max 9.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rgpu {

// species pair (a <= b) of key k (keys ordered (0,0), (0,1), ..., (1,1), ...)
__host__ __device__ inline void key_pair(int key, int S, int* a, int* b) {
  int aa = 0, kk = key;
  while (kk >= S - aa) {
    kk -= S - aa;
    ++aa;
  }
  *a = aa;
  *b = aa + kk;
}

__host__ __device__ inline int packed_diag_block(int N, int L1) { return N * (N + 1) / 2 * L1; }

// first packed column of key (a, b)
__host__ __device__ inline int64_t packed_key_offset(int a, int b, int S, int N, int L1) {
  const int key = a * S - a * (a - 1) / 2 + (b - a);
  const int nd = a + (b > a ? 1 : 0);  // diagonal keys before this one
  return int64_t(nd) * packed_diag_block(N, L1) + int64_t(key - nd) * N * N * L1;
}

__host__ __device__ inline int64_t packed_feature_size(int S, int N, int L1) {
  const int K = S * (S + 1) / 2;
  return int64_t(S) * packed_diag_block(N, L1) + int64_t(K - S) * N * N * L1;
}

// Row stride of packed rows in memory: the packed width rounded up to a
// multiple of 32 (the K step of the matrix-core GEMMs), pad columns zero.
__host__ __device__ inline int64_t packed_row_stride(int S, int N, int L1) {
  return (packed_feature_size(S, N, L1) + 31) / 32 * 32;
}

// (n1, n2), n1 <= n2, of packed pair index t of a diagonal key (row x
// starts at x N - x (x - 1) / 2; the smaller root of that quadratic, then
// exact integer corrections)
__host__ __device__ inline void packed_pair(int t, int N, int* n1, int* n2) {
  const float disc = float((2 * N + 1) * (2 * N + 1) - 8 * t);
  int x = int((float(2 * N + 1) - sqrtf(disc > 0.f ? disc : 0.f)) * 0.5f);
  x = x < 0 ? 0 : (x > N - 1 ? N - 1 : x);
  while (x > 0 && x * N - x * (x - 1) / 2 > t) --x;
  while (x + 1 <= N - 1 && (x + 1) * N - (x + 1) * x / 2 <= t) ++x;
  *n1 = x;
  *n2 = x + (t - (x * N - x * (x - 1) / 2));
}

}  // namespace rgpu
