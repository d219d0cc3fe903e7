// FP64 GAP GEMMs on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64), the
// parity path of the sparse GAP prediction.
//
//   kernel:  K[r][m] = sum_d X[row(r)][d] T[col0+m][d]   (B rows K-contiguous)
//   forces:  G[row(r)][d] = sum_m W~[r][m] T[col0+m][d]  (B given as [K][N],
//            transposed while staging through LDS)
//
// Tiling: 256 threads (4 wavefronts, 2 x 2), block tile 64 x 64, K step 16;
// each wavefront owns 32 x 32 = 2 x 2 MFMA tiles of 16 x 16.  Grouped launch
// over species blocks, tiles dealt to the XCDs as the mixed-precision
// kernel's (groups of kGroupM row panels walked column by column, a
// contiguous range of them per XCD, so the tiles resident on one XCD share
// operand panels through its L2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hh"
#include "gap_f64.hh"

namespace rgpu {

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4;

constexpr int BM = 64, BN = 64, BK = 16, LDK = BK + 1, NT = 256;
constexpr int kMaxProblems = 16;
constexpr int kGroupM = 8;  // row panels per tile group

struct Problem {
  int64_t M, N, K;
  const double* A;
  int64_t lda;
  const int* a_rows;
  const double* B;   // kernel: rows n (K contiguous); force: rows k (N contiguous)
  int64_t ldb;
  const int* c_rows;
  int64_t tile0;
  int tiles_m, tiles_n;
  const double* norm2;
  const double* alpha;
  double* epart;  // energy partials [tiles_n][ldp], by centre
  int64_t ldp;
  double* wt;
  double* G;
  int64_t ldg;
};

struct ProblemSet {
  Problem p[kMaxProblems];
  int64_t total_tiles;
  int n;
  int normalize;
  int zeta;
};

__device__ __forceinline__ double pow_int(double v, int n) {
  if (n == 0) return 1.0;
  if (n == 1) return v;
  if (n == 2) return v * v;
  if (n == 3) return v * v * v;
  double r = 1.0, x = v;
  unsigned u = unsigned(n);
  do {
    if (u & 1) r *= x;
    u >>= 1;
    x *= x;
  } while (u);
  return r;
}

template <int EPI>  // 0: kernel (B K-contiguous) + energy epilogue; 1: force (B N-contiguous)
__global__ __launch_bounds__(NT, 2) void f64_gemm(ProblemSet ps) {
  __shared__ double sA[BM * LDK];
  __shared__ double sB[BN * LDK];
  __shared__ double esum[2][BM];

  // the workgroups go to the 8 XCDs round robin: XCD x takes linear tiles
  // [x per_xcd, (x + 1) per_xcd)
  const int64_t per_xcd = (int64_t(gridDim.x) + 7) / 8;
  const int64_t bid = int64_t(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (bid >= ps.total_tiles) return;  // whole workgroup, before any barrier
  int pi = 0;
  while (pi + 1 < ps.n && bid >= ps.p[pi + 1].tile0) ++pi;
  const Problem& P = ps.p[pi];
  const int64_t local = bid - P.tile0;
  const int64_t per_group = int64_t(kGroupM) * P.tiles_n;
  const int64_t grp = local / per_group, in = local - grp * per_group;
  const int64_t gm = std::min<int64_t>(kGroupM, P.tiles_m - grp * kGroupM);
  const int64_t tn = in / gm, tm = grp * kGroupM + (in - tn * gm);
  const int64_t r0 = tm * BM, c0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // A: 64 rows x 16 k, four consecutive k per thread
  const int lr = tid >> 2, lk = (tid & 3) * 4;
  const int64_t arow = r0 + lr;
  const bool a_ok = arow < P.M;
  const double* Ap = a_ok ? P.A + (P.a_rows ? int64_t(P.a_rows[arow]) : arow) * P.lda : nullptr;
  // B kernel form: 64 rows (n) x 16 k like A; force form: 16 k rows x 64 n
  const int64_t brow = c0 + lr;
  const bool b_ok = EPI == 0 ? brow < P.N : true;
  const double* Bp = (EPI == 0 && b_ok) ? P.B + brow * P.ldb : nullptr;
  const int fk = tid >> 4, fn = (tid & 15) * 4;  // force form: k row, 4 consecutive n

  double ra[4], rb[4];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t k = k0 + lk + q;
      ra[q] = (a_ok && k < P.K) ? Ap[k] : 0.0;
    }
    if (EPI == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t k = k0 + lk + q;
        rb[q] = (b_ok && k < P.K) ? Bp[k] : 0.0;
      }
    } else {
      const int64_t k = k0 + fk;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t n = c0 + fn + q;
        rb[q] = (k < P.K && n < P.N) ? P.B[k * P.ldb + n] : 0.0;
      }
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) sA[lr * LDK + lk + q] = ra[q];
    if (EPI == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) sB[lr * LDK + lk + q] = rb[q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) sB[(fn + q) * LDK + fk] = rb[q];
    }
  };

  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[i][j][e] = 0.0;

  const int fr = lane & 15, fq = lane >> 4;
  gload(0);
  for (int64_t k0 = 0; k0 < P.K; k0 += BK) {
    __syncthreads();
    lstore();
    __syncthreads();
    if (k0 + BK < P.K) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = sA[(wm * 32 + i * 16 + fr) * LDK + ks + fq];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = sB[(wn * 32 + j * 16 + fr) * LDK + ks + fq];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }

  // element e of a 16x16 tile: row lane/16 + 4*e, column lane%16
  if (EPI == 0) {
    double rowsum[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        rowsum[i][e] = 0.0;
        const int64_t r = r0 + wm * 32 + i * 16 + fq + 4 * e;
        if (r >= P.M) continue;
        const int centre = P.c_rows[r];
        double inv = 1.0;
        if (ps.normalize) {
          const double n2 = P.norm2[centre];
          inv = n2 > 0 ? 1.0 / sqrt(n2) : 1.0;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int64_t c = c0 + wn * 32 + j * 16 + fr;
          if (c >= P.N) continue;
          const double k = acc[i][j][e] * inv;
          const double a = P.alpha[c];
          rowsum[i][e] += a * pow_int(k, ps.zeta);
          if (P.wt) P.wt[r * P.N + c] = a * (ps.zeta > 1 ? ps.zeta * pow_int(k, ps.zeta - 1) : 1.0);
        }
      }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double v = rowsum[i][e];
        for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        rowsum[i][e] = v;
      }
    if (fr == 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) esum[wn][wm * 32 + i * 16 + fq + 4 * e] = rowsum[i][e];
    }
    __syncthreads();
    if (P.epart && tid < BM) {
      const int64_t r = r0 + tid;
      if (r < P.M) P.epart[tn * P.ldp + P.c_rows[r]] = esum[0][tid] + esum[1][tid];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = r0 + wm * 32 + i * 16 + fq + 4 * e;
        if (r >= P.M) continue;
        double* g = P.G + int64_t(P.c_rows[r]) * P.ldg;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int64_t c = c0 + wn * 32 + j * 16 + fr;
          if (c < P.N) g[c] = acc[i][j][e];
        }
      }
  }
}

void launch(ProblemSet& ps, int64_t tiles, int epi, hipStream_t stream) {
  if (tiles == 0) return;
  ps.total_tiles = tiles;
  const unsigned grid = unsigned((tiles + 7) / 8 * 8);  // whole rounds over the 8 XCDs
  if (epi == 0)
    hipLaunchKernelGGL(f64_gemm<0>, dim3(grid), dim3(NT), 0, stream, ps);
  else
    hipLaunchKernelGGL(f64_gemm<1>, dim3(grid), dim3(NT), 0, stream, ps);
  RGPU_LAUNCH_CHECK();
}

// out[c][r] = in[r][c] for in [rows][cols], 32 x 32 tiles through LDS
__global__ void transpose_f64_kernel(int64_t rows, int64_t cols, const double* __restrict__ in,
                                     double* __restrict__ out) {
  __shared__ double t[32][33];
  const int64_t r0 = int64_t(blockIdx.y) * 32, c0 = int64_t(blockIdx.x) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) t[k][tx] = in[r * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) out[c * rows + r] = t[tx][k];
  }
}

}  // namespace

int gap_f64_energy_tiles(int64_t n_cols) { return int((n_cols + BN - 1) / BN); }

void run_gap_f64_kernel(const GapBlock<double>* blks, int n_blocks, double* energy, double* part, int64_t ld,
                        double* const* wts, hipStream_t stream) {
  static_assert(kMaxProblems <= EnergyTiles::kMax, "one energy reduction per grouped launch");
  for (int b0 = 0; b0 < n_blocks; b0 += kMaxProblems) {
    ProblemSet ps{};
    EnergyTiles et;
    int64_t tiles = 0;
    for (int b = b0; b < std::min(n_blocks, b0 + kMaxProblems); ++b) {
      const GapBlock<double>& blk = blks[b];
      if (blk.n_rows == 0 || blk.n_cols == 0) continue;
      if (!blk.alpha) throw InvalidArgument("the GAP energy kernel needs weights");
      Problem& P = ps.p[ps.n++];
      P.M = blk.n_rows;
      P.N = blk.n_cols;
      P.K = blk.D;
      P.A = blk.X;
      P.lda = blk.D;
      P.a_rows = blk.row_idx;
      P.B = blk.Tm + blk.col0 * blk.D;
      P.ldb = blk.D;
      P.c_rows = blk.row_idx;
      P.tile0 = tiles;
      P.tiles_m = int((P.M + BM - 1) / BM);
      P.tiles_n = int((P.N + BN - 1) / BN);
      P.norm2 = blk.norm2;
      P.alpha = blk.alpha + blk.col0;
      P.epart = energy ? part : nullptr;
      P.ldp = ld;
      P.wt = wts ? wts[b] : nullptr;
      tiles += ((P.M + BM - 1) / BM) * P.tiles_n;
      ps.normalize = blk.normalize;
      ps.zeta = blk.zeta;
      et.rows[et.n] = blk.row_idx;
      et.m[et.n] = blk.n_rows;
      et.tiles[et.n++] = P.tiles_n;
    }
    launch(ps, tiles, 0, stream);
    if (energy) launch_energy_tiles(et, part, ld, energy, stream);
  }
}

void run_gram_f64(int64_t rows, int64_t n, const double* A, double* At, const int* iota, double* C,
                  hipStream_t stream) {
  if (rows == 0 || n == 0) return;
  const dim3 grid(unsigned((n + 31) / 32), unsigned((rows + 31) / 32));
  hipLaunchKernelGGL(transpose_f64_kernel, grid, dim3(256), 0, stream, rows, n, A, At);
  RGPU_LAUNCH_CHECK();
  // C[i][:] = sum_r At[i][r] A[r][:], the force product with W~ = A^T
  GapBlock<double> blk{};
  blk.n_rows = n;
  blk.row_idx = iota;
  blk.n_cols = rows;
  blk.col0 = 0;
  blk.Tm = A;
  blk.D = n;
  double* const wts[1] = {At};
  run_gap_f64_force(&blk, 1, wts, C, stream);
}

void run_gap_f64_force(const GapBlock<double>* blks, int n_blocks, double* const* wts, double* G,
                       hipStream_t stream) {
  for (int b0 = 0; b0 < n_blocks; b0 += kMaxProblems) {
    ProblemSet ps{};
    int64_t tiles = 0;
    for (int b = b0; b < std::min(n_blocks, b0 + kMaxProblems); ++b) {
      const GapBlock<double>& blk = blks[b];
      if (blk.n_rows == 0 || blk.n_cols == 0) continue;
      Problem& P = ps.p[ps.n++];
      P.M = blk.n_rows;
      P.N = blk.D;
      P.K = blk.n_cols;
      P.A = wts[b];
      P.lda = blk.n_cols;
      P.a_rows = nullptr;
      P.B = blk.Tm + blk.col0 * blk.D;  // [K = sparse points][N = D]
      P.ldb = blk.D;
      P.c_rows = blk.row_idx;
      P.tile0 = tiles;
      P.tiles_m = int((P.M + BM - 1) / BM);
      P.tiles_n = int((P.N + BN - 1) / BN);
      P.G = G;
      P.ldg = blk.D;
      tiles += ((P.M + BM - 1) / BM) * P.tiles_n;
    }
    launch(ps, tiles, 1, stream);
  }
}

}  // namespace rgpu
