// FP64 GAP GEMMs on the gfx950 matrix cores (parity path).
#pragma once

#include <hip/hip_runtime.h>

#include "gap_kernels.hh"

namespace rgpu {

// K GEMM + epilogue for all species blocks (grouped): energy per centre (+=;
// summed over the column tiles through part, [gap_f64_energy_tiles(n_cols)
// of the widest block] x [ld >= n_centres] doubles, see EnergyTiles) and
// wts[b] [n_rows][n_cols] (wts may be null)
int gap_f64_energy_tiles(int64_t n_cols);
void run_gap_f64_kernel(const GapBlock<double>* blks, int n_blocks, double* energy, double* part, int64_t ld,
                        double* const* wts, hipStream_t stream);
// G[row_idx[r]][:] = sum_m wts[b][r][m] T[col0+m][:]
void run_gap_f64_force(const GapBlock<double>* blks, int n_blocks, double* const* wts, double* G,
                       hipStream_t stream);

// C (n x n) = A^T A for A [rows][n] (device pointers; At: room for n x rows
// doubles, iota: 0 .. n-1 as ints)
void run_gram_f64(int64_t rows, int64_t n, const double* A, double* At, const int* iota, double* C,
                  hipStream_t stream);

}  // namespace rgpu
