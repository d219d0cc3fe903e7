// Mixed-precision sparse GAP kernel on the gfx950 matrix cores: BF16 hi+lo
// split operands (x = hi + lo, three products hi.hi + hi.lo + lo.hi) with
// FP32 accumulation, fused zeta-power / energy / weight epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device_buffer.hh"
#include "gap_kernels.hh"

namespace rgpu {

// Sparse points in the packed layout of the invariants (packing.hh),
// pre-split into BF16 hi/lo planes, in the two layouts the two GEMMs read:
// [M][D] (K-contiguous rows for K = X T^T) and [D][M] (K-contiguous
// columns for T~ = W~ T); D here is the packed width.
struct MixedSparse {
  int64_t M = 0, D = 0;
  DeviceBuffer<unsigned short> t_hi, t_lo;    // [M][D]
  DeviceBuffer<unsigned short> tt_hi, tt_lo;  // [D][M]
};

// A BF16 hi/lo pair of planes (same shape); bit patterns of __bf16.
struct SplitPlanes {
  unsigned short* hi = nullptr;
  unsigned short* lo = nullptr;
};

bool mfma_available();
// T: [M][D] FP64 rows in the full layout of S species, N radial channels,
// L1 = max_angular + 1
void prepare_mixed_sparse(MixedSparse& ms, const std::vector<double>& T, int64_t M, int S, int N,
                          int L1, hipStream_t stream);
// K tile GEMM + epilogue for all species blocks in grouped launches.  The
// features are given split: X = xs.hi + xs.lo, [n_centres][D] (rows gathered
// through the blocks' row_idx, un-normalised, scaled by 1/|p| in the
// epilogue).  Outputs: energy (per centre, +=; summed over the column
// tiles through part, [gap_mixed_energy_tiles(n_cols) of the widest block]
// x [ld >= n_centres] doubles, see EnergyTiles) and the weighted kernel
// gradient W~ of each block, split, [n_rows][n_cols] (wts may be null).
int gap_mixed_energy_tiles(int64_t n_cols);
void run_gap_mixed_kernel(const GapBlock<float>* blks, int n_blocks, const MixedSparse& ms,
                          SplitPlanes xs, double* energy, double* part, int64_t ld, const SplitPlanes* wts,
                          hipStream_t stream);
// T~ rows: G[row_idx[r]][:] = sum_m W~_b[r][m] T[col0+m][:]
void run_gap_mixed_force(const GapBlock<float>* blks, int n_blocks, const MixedSparse& ms,
                         const SplitPlanes* wts, float* G, hipStream_t stream);

}  // namespace rgpu
