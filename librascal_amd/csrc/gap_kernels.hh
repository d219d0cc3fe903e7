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
// knm_atom ([n_centres][M], written), knm_struct ([n_struct][M], +=, with
// centre_struct) may be null.
template <class T>
void launch_gap_kernel(const GapBlock<T>& blk, int64_t M, double* energy, T* wt, double* knm_atom,
                       double* knm_struct, const int* centre_struct, hipStream_t stream);

// GEMM2: G[row_idx[r]][:] = sum_m wt[r][m] Tm[col0+m][:]
template <class T>
void launch_gap_force_contraction(const GapBlock<T>& blk, const T* wt, T* G, hipStream_t stream);

}  // namespace rgpu
