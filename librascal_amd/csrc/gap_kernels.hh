// Sparse GAP kernel on the GPU: K = (X T^T)^zeta per species block, energy,
// the weighted kernel gradient W~ = alpha zeta (X T^T)^(zeta-1) and the force
// contraction T~ = W~ T (models/sparse_kernels.hh:104-190,
// models/sparse_kernel_predict.hh:70-181).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rgpu {

// One species block: rows = centres of the species (gathered through
// row_idx), columns = sparse points of the species.
template <class T>
struct GapBlock {
  int64_t n_rows;         // centres of this species
  const int* row_idx;     // [n_rows] centre indices
  int64_t n_cols;         // sparse points of this species
  int64_t col0;           // first sparse point (global, species-sorted order)
  const T* X;             // [n_centres][D] un-normalised power spectrum
  const double* norm2;    // [n_centres]
  int normalize;
  const T* Tm;            // [M][D] sparse points (species-sorted)
  const double* alpha;    // [M] weights (may be null when no energy)
  int zeta;
  int64_t D;
};

// GEMM1 + epilogue.  Any of energy (per centre, +=), wt ([n_rows][n_cols]),
// knm_atom ([n_centres][M], written) may be null.
template <class T>
void launch_gap_kernel(const GapBlock<T>& blk, int64_t M, double* energy, T* wt, double* knm_atom,
                       hipStream_t stream);

// out[s][:] = sum over rows a in [seg_off[s], seg_off[s+1]) of in[a][:]
// (rows of w doubles, summed in row order; out written)
void launch_rows_segment_sum(const double* in, int64_t w, const int64_t* seg_off, int n_seg, double* out,
                             hipStream_t stream);

// Per-centre energies without atomics: a GEMM's energy epilogue stores the
// partial row sum of its column tile t at part[t * ld + centre], and
// launch_energy_tiles adds the sum over the tiles, taken in tile order, to
// energy[centre] (the same value on every run).  One entry per species
// block: its centres and its number of column tiles.
struct EnergyTiles {
  static constexpr int kMax = 16;
  int n = 0;
  const int* rows[kMax] = {};
  int64_t m[kMax] = {};
  int tiles[kMax] = {};
};
void launch_energy_tiles(const EnergyTiles& et, const double* part, int64_t ld, double* energy,
                         hipStream_t stream);

// GEMM2: G[row_idx[r]][:] = sum_m wt[r][m] Tm[col0+m][:]
template <class T>
void launch_gap_force_contraction(const GapBlock<T>& blk, const T* wt, T* G, hipStream_t stream);

}  // namespace rgpu
