// Mixed-precision GAP GEMMs on the gfx950 matrix cores.
//
// Both products of the GAP prediction are computed as C = A B^T with A and B
// stored row-major with the contraction index contiguous:
//   kernel:  K[r][m] = sum_d X[row(r)][d] T[col0+m][d]       (B = T, [M][D])
//   forces:  G[row(r)][d] = sum_m W~[r][m] T[col0+m][d]      (B = T^T, [D][M])
// Every operand arrives split into BF16 planes, x = hi + lo with
// hi = bf16(x), lo = bf16(x - hi) (the power spectrum and this kernel's own
// energy epilogue write them that way; the sparse points are split once at
// load time), so the main loop does no conversions: each algorithmic
// multiply-add is three v_mfma_f32_32x32x16_bf16 products (hi.hi + hi.lo +
// lo.hi) accumulated in FP32.
//
// Tiling: 256 threads (4 wavefronts, 2 x 2), block tile 128 x 128, K step 32;
// each wavefront owns 64 x 64 = 2 x 2 MFMA tiles.  Two LDS stages (one
// barrier per K step), the next K step's global loads issued before the
// MFMAs of the current one.  Where every row is 16-byte aligned and K is a
// multiple of 32 (the packed layout guarantees both), the loads go straight
// to LDS (global_load_lds_dwordx4) in a swizzled layout; otherwise they pass
// through registers with a guarded K tail.  All species blocks of one
// product go in one grouped launch, dealt to the XCDs so that a row panel
// stays in one L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "common.hh"
#include "gap_mfma.hh"
#include "packing.hh"

namespace rgpu {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int BM = 128, BN = 128, BK = 32, NT = 256;
constexpr int kMaxProblems = 16;
constexpr int kGroupM = 8;  // row panels per tile group

__device__ __forceinline__ unsigned short bf16_bits(__bf16 h) {
  return __builtin_bit_cast(unsigned short, h);
}

struct Problem {
  int64_t M, N, K;             // C (M x N) = A (M x K) B (N x K)^T
  const unsigned short* Ahi;   // A rows, K contiguous
  const unsigned short* Alo;
  int64_t lda;
  const int* a_rows;           // gathered A rows (null: identity)
  const unsigned short* Bhi;   // B rows, K contiguous, already offset
  const unsigned short* Blo;
  int64_t ldb;
  const int* c_rows;           // output row ids (centre indices)
  int64_t tile0;               // first linear tile of this problem
  int tiles_m, tiles_n;        // tiles along M and N
  int fast;                    // K % 32 == 0 and every row 16-byte aligned
  // kernel epilogue
  const double* norm2;
  const double* alpha;         // offset to the block's first sparse point
  double* epart;               // energy partials [tiles_n][ldp], by centre
  int64_t ldp;
  unsigned short* wth;         // [M][N] W~ split (may be null)
  unsigned short* wtl;
  // force epilogue
  float* G;
  int64_t ldg;
};

struct ProblemSet {
  Problem p[kMaxProblems];
  int64_t total_tiles;
  int n;
  int normalize;
  int zeta;
};

template <class T>
__device__ __forceinline__ T pow_int(T v, int n) {
  if (n == 0) return T(1);
  if (n == 1) return v;
  if (n == 2) return v * v;
  if (n == 3) return v * v * v;
  T r = T(1), x = v;
  unsigned u = unsigned(n);
  do {
    if (u & 1) r *= x;
    u >>= 1;
    x *= x;
  } while (u);
  return r;
}

// 32-row block rb, 16-wide K chunk c, lane l (row l%32, K 8*(l/32) .. +8):
// fragment slot in an LDS stage, XOR-swizzled by 4 c + 2 (l / 32) inside
// aligned groups of 8 so the K-coalesced stores and the fragment reads hit
// distinct banks
__device__ __forceinline__ int frag_slot(int rb, int c, int l) {
  return (rb * (BK / 16) + c) * 64 + (l ^ (4 * c + 2 * (l >> 5)));
}

constexpr int FR = (BM / 32) * (BK / 16) * 64;  // fragments per operand per LDS stage

// LDS of one workgroup: two stages of the four operand planes
struct GemmLds {
  bf16x8 Ah[2][FR];
  bf16x8 Al[2][FR];
  bf16x8 Bh[2][FR];
  bf16x8 Bl[2][FR];
  double esum[2][BM];
};

// Epilogue of a TBM x 128 block tile whose wavefronts form (TBM / 64) x 2
// and own 64 x 64 = 2 x 2 MFMA tiles each; esum: LDS [2][TBM]
template <int EPI, int TBM>
__device__ __forceinline__ void gemm_epilogue(const ProblemSet& ps, const Problem& P, int64_t r0,
                                              int64_t c0, f32x16 (&acc)[2][2], double (*esum)[TBM]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // epilogue; element e = 4*g + q of a 32x32 tile sits at
  // row 8*g + 4*(lane/32) + q, column lane%32
  const int half = lane >> 5, fr = lane & 31;
  if (EPI == 0) {
    double rowsum[2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e) rowsum[i][e] = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int rl = wm * 64 + i * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
        const int64_t r = r0 + rl;
        if (r >= P.M) continue;
        const int centre = P.c_rows[r];
        double inv = 1.0;
        if (ps.normalize) {
          const double n2 = P.norm2[centre];
          inv = n2 > 0 ? 1.0 / sqrt(n2) : 1.0;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int64_t c = c0 + wn * 64 + j * 32 + fr;
          if (c >= P.N) continue;
          const float k = float(double(acc[i][j][e]) * inv);
          const double a = P.alpha[c];
          rowsum[i][e] += a * double(pow_int(k, ps.zeta));
          if (P.wth) {
            const float g = ps.zeta > 1 ? float(ps.zeta) * pow_int(k, ps.zeta - 1) : 1.f;
            const float v = float(a) * g;
            const __bf16 h = (__bf16)v;
            P.wth[r * P.N + c] = bf16_bits(h);
            P.wtl[r * P.N + c] = bf16_bits((__bf16)(v - (float)h));
          }
        }
      }
    }
    // reduce over the 32 columns held by the lanes of each half
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        double v = rowsum[i][e];
        for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        rowsum[i][e] = v;
      }
    if (fr == 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e)
          esum[wn][wm * 64 + i * 32 + 8 * (e >> 2) + 4 * half + (e & 3)] = rowsum[i][e];
    }
    __syncthreads();
    if (P.epart && tid < TBM) {
      const int64_t r = r0 + tid;
      if (r < P.M) P.epart[(c0 / BN) * P.ldp + P.c_rows[r]] = esum[0][tid] + esum[1][tid];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t r = r0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
        if (r >= P.M) continue;
        float* g = P.G + int64_t(P.c_rows[r]) * P.ldg;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int64_t c = c0 + wn * 64 + j * 32 + fr;
          if (c < P.N) g[c] = acc[i][j][e];
        }
      }
  }
}

template <int EPI, bool FAST>
__device__ __forceinline__ void gemm_tile(const ProblemSet& ps, const Problem& P, int64_t r0, int64_t c0,
                                          GemmLds& sm) {
  auto& sAh = sm.Ah;
  auto& sAl = sm.Al;
  auto& sBh = sm.Bh;
  auto& sBl = sm.Bl;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // Global loads are coalesced along K: four lanes read the 32 K values
  // (64 B per plane) of one row, rows rs + 64 j (j < 2).  Rows past the
  // matrix edge repeat its last row (their results are never stored).
  const int rs = tid >> 2, kq = tid & 3;
  const unsigned short *Ah[2], *Al[2], *Bh[2], *Bl[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    int64_t ra = r0 + rs + 64 * j;
    ra = ra < P.M ? ra : P.M - 1;
    if (P.a_rows) ra = P.a_rows[ra];
    Ah[j] = P.Ahi + ra * P.lda;
    Al[j] = P.Alo + ra * P.lda;
    int64_t rb = c0 + rs + 64 * j;
    rb = rb < P.N ? rb : P.N - 1;
    Bh[j] = P.Bhi + rb * P.ldb;
    Bl[j] = P.Blo + rb * P.ldb;
  }
  const int sc = kq >> 1, shalf = kq & 1;
  int fs[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int r = rs + 64 * j;
    fs[j] = frag_slot(r >> 5, sc, (r & 31) + 32 * shalf);
  }

  // two register stages: the global loads of K step kt + 2 are issued while
  // step kt computes, and stored to LDS at the end of step kt + 1
  uint4 rah[2][2], ral[2][2], rbh[2][2], rbl[2][2];
  auto load8 = [&](const unsigned short* p, int64_t k) -> uint4 {
    if (FAST) return *reinterpret_cast<const uint4*>(p + k);
    unsigned short t[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) t[q] = (k + q < P.K) ? p[k + q] : (unsigned short)0;
    uint4 v;
    memcpy(&v, t, 16);
    return v;
  };
  auto gload = [&](int rs_, int64_t k0) {
    const int64_t k = k0 + 8 * kq;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      rah[rs_][j] = load8(Ah[j], k);
      ral[rs_][j] = load8(Al[j], k);
      rbh[rs_][j] = load8(Bh[j], k);
      rbl[rs_][j] = load8(Bl[j], k);
    }
  };
  auto lstore = [&](int rs_, int st) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      sAh[st][fs[j]] = __builtin_bit_cast(bf16x8, rah[rs_][j]);
      sAl[st][fs[j]] = __builtin_bit_cast(bf16x8, ral[rs_][j]);
      sBh[st][fs[j]] = __builtin_bit_cast(bf16x8, rbh[rs_][j]);
      sBl[st][fs[j]] = __builtin_bit_cast(bf16x8, rbl[rs_][j]);
    }
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

  auto compute = [&](int st) {
#pragma unroll
    for (int c = 0; c < BK / 16; ++c) {
      bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int f = frag_slot(wm * 2 + i, c, lane);
        ah[i] = sAh[st][f];
        al[i] = sAl[st][f];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int f = frag_slot(wn * 2 + j, c, lane);
        bh[j] = sBh[st][f];
        bl[j] = sBl[st][f];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
  };

  const int64_t nk = (P.K + BK - 1) / BK;
  // K step of a prefetch, clamped to the last one: the loads and LDS stores
  // are issued unconditionally (the surplus ones at the end are never read),
  // which keeps the wait before each LDS store at the loads it consumes
  // rather than at every load in flight
  auto step = [&](int64_t kt) { return (kt < nk ? kt : nk - 1) * BK; };
  gload(0, 0);
  gload(1, step(1));
  lstore(0, 0);
  __syncthreads();
  // K steps in pairs so that the register stage indices are constants;
  // step kt reads LDS stage kt & 1 and register stage kt & 1 is refilled
  for (int64_t kt = 0; kt < nk; kt += 2) {
    gload(0, step(kt + 2));
    compute(0);
    // the other LDS stage was last read before the previous barrier
    lstore(1, 1);
    __syncthreads();
    if (kt + 1 >= nk) break;
    gload(1, step(kt + 3));
    compute(1);
    lstore(0, 0);
    __syncthreads();
  }

  gemm_epilogue<EPI, BM>(ps, P, r0, c0, acc, sm.esum);
}

// s_waitcnt vmcnt(N), the other counters left alone (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Direct-to-LDS variant (global_load_lds_dwordx4: the loads write LDS
// without passing through VGPRs).  A stage of one operand plane holds the
// tile's rows of one K step (BK = 32 values, 64 B) as groups of 16 rows, one
// wavefront load instruction (64 lanes x 16 B) per group; inside a group
// row r keeps its K quarter q (8 values) at slot 4 r + (q ^ dma_swz(r)), so
// that each instruction reads whole 64-byte row segments and the fragment
// reads (16 rows of one quarter per 16 lanes) hit distinct banks.
__device__ __forceinline__ int dma_swz(int r) { return (r ^ (r >> 2)) & 3; }
// slot of the fragment of lane l for 32-row block rb, 16-wide K chunk c
__device__ __forceinline__ int dma_frag_slot(int rb, int c, int l) {
  const int r = l & 15;
  return (2 * rb + ((l >> 4) & 1)) * 64 + 4 * r + ((2 * c + (l >> 5)) ^ dma_swz(r));
}

template <int TBM, int NS>
struct DmaLds {
  bf16x8 A[NS][2][TBM / 16 * 64];  // [stage][hi, lo][slot]
  bf16x8 B[NS][2][BN / 16 * 64];
  double esum[2][TBM];
};

// TBM x 128 block tile, TBM / 32 wavefronts, NS LDS stages with NS - 1 K
// steps in flight; requires the fast layout (K % 32 == 0, aligned rows)
template <int EPI, int TBM, int NS>
__device__ __forceinline__ void gemm_tile_dma(const ProblemSet& ps, const Problem& P, int64_t r0,
                                              int64_t c0, DmaLds<TBM, NS>& sm) {
  constexpr int W = TBM / 32;
  constexpr int GA = TBM / 16 / W, GB = BN / 16 / W;  // groups per plane and wavefront
  static_assert(GA * W * 16 == TBM && GB * W * 16 == BN, "group partition");
  constexpr int PER = 2 * (GA + GB);  // load instructions per wavefront and K step
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // lane of a group load: row lane / 4, quarter q; rows past the matrix
  // edge repeat its last row (their results are never stored)
  const int rr = lane >> 2, q = (lane & 3) ^ dma_swz(rr);
  int64_t offa[GA], offb[GB];
#pragma unroll
  for (int j = 0; j < GA; ++j) {
    int64_t ra = r0 + (wave + W * j) * 16 + rr;
    ra = ra < P.M ? ra : P.M - 1;
    if (P.a_rows) ra = P.a_rows[ra];
    offa[j] = ra * P.lda + 8 * q;
  }
#pragma unroll
  for (int j = 0; j < GB; ++j) {
    int64_t rb = c0 + (wave + W * j) * 16 + rr;
    rb = rb < P.N ? rb : P.N - 1;
    offb[j] = rb * P.ldb + 8 * q;
  }
  const int64_t nk = P.K / BK;
  // loads of K step kt (clamped to the last one: the surplus loads at the
  // end land in stages that are never read) into stage s
  auto issue = [&](int s, int64_t kt) {
    const int64_t k = (kt < nk ? kt : nk - 1) * BK;
#pragma unroll
    for (int j = 0; j < GA; ++j) {
      const int g = (wave + W * j) * 64;
      __builtin_amdgcn_global_load_lds(P.Ahi + offa[j] + k, &sm.A[s][0][g], 16, 0, 0);
      __builtin_amdgcn_global_load_lds(P.Alo + offa[j] + k, &sm.A[s][1][g], 16, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < GB; ++j) {
      const int g = (wave + W * j) * 64;
      __builtin_amdgcn_global_load_lds(P.Bhi + offb[j] + k, &sm.B[s][0][g], 16, 0, 0);
      __builtin_amdgcn_global_load_lds(P.Blo + offb[j] + k, &sm.B[s][1][g], 16, 0, 0);
    }
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

  auto compute = [&](int s) {
#pragma unroll
    for (int c = 0; c < BK / 16; ++c) {
      bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int f = dma_frag_slot(wm * 2 + i, c, lane);
        ah[i] = sm.A[s][0][f];
        al[i] = sm.A[s][1][f];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int f = dma_frag_slot(wn * 2 + j, c, lane);
        bh[j] = sm.B[s][0][f];
        bl[j] = sm.B[s][1][f];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
  };

#pragma unroll
  for (int s = 0; s < NS - 1; ++s) issue(s, s);
  int s_read = 0, s_fill = NS - 1;
  for (int64_t kt = 0; kt < nk; ++kt) {
    // this wavefront's loads of step kt are done once at most the NS - 2
    // later steps' remain; the barrier extends that to every wavefront and
    // orders the previous step's fragment reads before the refill of its
    // stage
    wait_vmcnt<(NS - 2) * PER>();
    __syncthreads();
    issue(s_fill, kt + NS - 1);
    compute(s_read);
    s_read = s_read + 1 == NS ? 0 : s_read + 1;
    s_fill = s_fill + 1 == NS ? 0 : s_fill + 1;
  }
  // no load may still write this workgroup's LDS when it exits
  wait_vmcnt<0>();
  gemm_epilogue<EPI, TBM>(ps, P, r0, c0, acc, sm.esum);
}

// problem index and tile (row, column) of linear block bid.  Workgroups
// are dealt to the 8 XCDs round robin; each XCD takes a contiguous range of
// linear tiles, and the linear order walks groups of kGroupM row panels
// column by column, so the tiles resident on one XCD at a time form a
// square of row and column panels that share their operand panels through
// the XCD's L2.
__device__ __forceinline__ bool tile_of(const ProblemSet& ps, int* pi_out, int64_t* tm_out,
                                        int64_t* tn_out) {
  const int64_t per_xcd = (int64_t(gridDim.x) + 7) / 8;
  const int64_t bid = int64_t(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (bid >= ps.total_tiles) return false;
  int pi = 0;
  while (pi + 1 < ps.n && bid >= ps.p[pi + 1].tile0) ++pi;
  const Problem& P = ps.p[pi];
  const int64_t local = bid - P.tile0;
  const int64_t per_group = int64_t(kGroupM) * P.tiles_n;
  const int64_t grp = local / per_group, in = local - grp * per_group;
  const int64_t gm = std::min<int64_t>(kGroupM, P.tiles_m - grp * kGroupM);
  const int64_t tn = in / gm;
  *pi_out = pi;
  *tm_out = grp * kGroupM + (in - tn * gm);
  *tn_out = tn;
  return true;
}

template <int EPI>  // 0: GAP kernel epilogue, 1: store rows of G
__global__ __launch_bounds__(NT, 2) void bf16x3_gemm(ProblemSet ps) {
  int pi;
  int64_t tm, tn;
  if (!tile_of(ps, &pi, &tm, &tn)) return;
  const Problem& P = ps.p[pi];
  __shared__ __align__(16) GemmLds sm;
  if (P.fast)
    gemm_tile<EPI, true>(ps, P, tm * BM, tn * BN, sm);
  else
    gemm_tile<EPI, false>(ps, P, tm * BM, tn * BN, sm);
}

// every problem of the set in the fast layout
template <int EPI, int TBM, int NS>
__global__ __launch_bounds__(TBM * 2, 512 / (TBM * 2)) void bf16x3_gemm_dma(ProblemSet ps) {
  int pi;
  int64_t tm, tn;
  if (!tile_of(ps, &pi, &tm, &tn)) return;
  const Problem& P = ps.p[pi];
  __shared__ __align__(16) DmaLds<TBM, NS> sm;
  gemm_tile_dma<EPI, TBM, NS>(ps, P, tm * TBM, tn * BN, sm);
}

unsigned short bf16_rne_host(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
float bf16_to_float_host(unsigned short h) {
  unsigned u = unsigned(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

void finish_problem(Problem& P) {
  P.fast = (P.K % BK == 0) && (P.lda % 8 == 0) && (P.ldb % 8 == 0) && aligned16(P.Ahi) &&
           aligned16(P.Alo) && aligned16(P.Bhi) && aligned16(P.Blo);
}

// tile counts and first tiles of the set's problems for block rows bm
int64_t layout_tiles(ProblemSet& ps, int bm) {
  int64_t tiles = 0;
  for (int i = 0; i < ps.n; ++i) {
    Problem& P = ps.p[i];
    P.tile0 = tiles;
    P.tiles_m = int((P.M + bm - 1) / bm);
    P.tiles_n = int((P.N + BN - 1) / BN);
    tiles += int64_t(P.tiles_m) * P.tiles_n;
  }
  return tiles;
}

template <int EPI>
void launch_set(ProblemSet& ps, hipStream_t stream) {
  if (ps.n == 0) return;
  bool fast = true;
  for (int i = 0; i < ps.n; ++i) fast = fast && ps.p[i].fast;
  auto go = [&](auto kernel, int bm, int threads) {
    ps.total_tiles = layout_tiles(ps, bm);
    if (ps.total_tiles == 0) return;
    const unsigned grid = unsigned((ps.total_tiles + 7) / 8 * 8);  // whole rounds over the 8 XCDs
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, stream, ps);
    RGPU_LAUNCH_CHECK();
  };
  // 128 x 128 tiles either way; two stages of direct-to-LDS loads measured
  // faster than three (with 256-row tiles) and than the register-staged loop
  if (fast)
    go(bf16x3_gemm_dma<EPI, BM, 2>, BM, NT);
  else
    go(bf16x3_gemm<EPI>, BM, NT);
}

}  // namespace

bool mfma_available() { return true; }

void prepare_mixed_sparse(MixedSparse& ms, const std::vector<double>& T, int64_t M, int S, int N,
                          int L1, hipStream_t stream) {
  // rows of T in the packed layout (packing.hh); a sparse point's diagonal
  // blocks need not be symmetric: (t[n1][n2] + t[n2][n1]) / sqrt 2 keeps
  // every product with a symmetric feature row exact
  const int K = S * (S + 1) / 2, KB = N * N * L1, KD = packed_diag_block(N, L1);
  const int64_t D = int64_t(K) * KB;
  const int64_t Dq = packed_row_stride(S, N, L1);  // zero pad columns
  ms.M = M;
  ms.D = Dq;
  const size_t n = size_t(M * Dq);
  std::vector<unsigned short> hi(std::max<size_t>(n, 1)), lo(std::max<size_t>(n, 1)),
      thi(std::max<size_t>(n, 1)), tlo(std::max<size_t>(n, 1));
  std::vector<double> row(size_t(std::max<int64_t>(Dq, 1)), 0.0);
  for (int64_t m = 0; m < M; ++m) {
    const double* src = T.data() + m * D;
    for (int k = 0; k < K; ++k) {
      int a, b;
      key_pair(k, S, &a, &b);
      const double* sk = src + int64_t(k) * KB;
      double* dk = row.data() + packed_key_offset(a, b, S, N, L1);
      if (a < b) {
        std::copy(sk, sk + KB, dk);
      } else {
        for (int o = 0; o < KD; ++o) {
          const int t = o / L1, l = o - t * L1;
          int n1, n2;
          packed_pair(t, N, &n1, &n2);
          dk[o] = n1 < n2 ? (sk[(n1 * N + n2) * L1 + l] + sk[(n2 * N + n1) * L1 + l]) * 0.70710678118654752440
                          : sk[(n1 * N + n1) * L1 + l];
        }
      }
    }
    for (int64_t d = 0; d < Dq; ++d) {
      const float x = float(row[size_t(d)]);
      const unsigned short h = bf16_rne_host(x);
      const unsigned short l = bf16_rne_host(x - bf16_to_float_host(h));
      hi[size_t(m * Dq + d)] = h;
      lo[size_t(m * Dq + d)] = l;
      thi[size_t(d * M + m)] = h;
      tlo[size_t(d * M + m)] = l;
    }
  }
  auto up = [&](DeviceBuffer<unsigned short>& b, const std::vector<unsigned short>& v) {
    unsigned short* p = b.reserve(v.size());
    RGPU_HIP(hipMemcpyAsync(p, v.data(), v.size() * 2, hipMemcpyHostToDevice, stream));
  };
  up(ms.t_hi, hi);
  up(ms.t_lo, lo);
  up(ms.tt_hi, thi);
  up(ms.tt_lo, tlo);
  RGPU_HIP(hipStreamSynchronize(stream));
}

int gap_mixed_energy_tiles(int64_t n_cols) { return int((n_cols + BN - 1) / BN); }

void run_gap_mixed_kernel(const GapBlock<float>* blks, int n_blocks, const MixedSparse& ms,
                          SplitPlanes xs, double* energy, double* part, int64_t ld, const SplitPlanes* wts,
                          hipStream_t stream) {
  static_assert(kMaxProblems <= EnergyTiles::kMax, "one energy reduction per grouped launch");
  for (int b0 = 0; b0 < n_blocks; b0 += kMaxProblems) {
    ProblemSet ps{};
    EnergyTiles et;
    for (int b = b0; b < std::min(n_blocks, b0 + kMaxProblems); ++b) {
      const GapBlock<float>& blk = blks[b];
      if (blk.n_rows == 0 || blk.n_cols == 0) continue;
      Problem& P = ps.p[ps.n++];
      P.M = blk.n_rows;
      P.N = blk.n_cols;
      P.K = blk.D;
      P.Ahi = xs.hi;
      P.Alo = xs.lo;
      P.lda = blk.D;
      P.a_rows = blk.row_idx;
      P.Bhi = ms.t_hi.get() + blk.col0 * blk.D;
      P.Blo = ms.t_lo.get() + blk.col0 * blk.D;
      P.ldb = blk.D;
      P.c_rows = blk.row_idx;
      P.norm2 = blk.norm2;
      if (!blk.alpha) throw InvalidArgument("the GAP energy kernel needs weights");
      P.alpha = blk.alpha + blk.col0;
      P.epart = energy ? part : nullptr;
      P.ldp = ld;
      P.wth = wts ? wts[b].hi : nullptr;
      P.wtl = wts ? wts[b].lo : nullptr;
      finish_problem(P);
      ps.normalize = blk.normalize;
      ps.zeta = blk.zeta;
      et.rows[et.n] = blk.row_idx;
      et.m[et.n] = blk.n_rows;
      et.tiles[et.n++] = gap_mixed_energy_tiles(blk.n_cols);
    }
    launch_set<0>(ps, stream);
    if (energy) launch_energy_tiles(et, part, ld, energy, stream);
  }
}

void run_gap_mixed_force(const GapBlock<float>* blks, int n_blocks, const MixedSparse& ms,
                         const SplitPlanes* wts, float* G, hipStream_t stream) {
  for (int b0 = 0; b0 < n_blocks; b0 += kMaxProblems) {
    ProblemSet ps{};
    for (int b = b0; b < std::min(n_blocks, b0 + kMaxProblems); ++b) {
      const GapBlock<float>& blk = blks[b];
      if (blk.n_rows == 0 || blk.n_cols == 0) continue;
      Problem& P = ps.p[ps.n++];
      P.M = blk.n_rows;
      P.N = blk.D;
      P.K = blk.n_cols;
      P.Ahi = wts[b].hi;
      P.Alo = wts[b].lo;
      P.lda = blk.n_cols;
      P.a_rows = nullptr;
      P.Bhi = ms.tt_hi.get() + blk.col0;
      P.Blo = ms.tt_lo.get() + blk.col0;
      P.ldb = ms.M;
      P.c_rows = blk.row_idx;
      P.G = G;
      P.ldg = blk.D;
      finish_problem(P);
    }
    launch_set<1>(ps, stream);
  }
}

}  // namespace rgpu
