// Launch interface of the SOAP kernels (expansion, power spectrum, adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_buffer.hh"
#include "radial.hh"

namespace rgpu {

// Pair records produced by the neighbour list, grouped per centre (self pair
// first).
struct PairView {
  int64_t n_centres;
  const int64_t* centre_off;  // [n_centres+1]
  const int* pair_j;          // original atom (global index)
  const double* pair_vec;     // [n_pairs][3]
  const double* pair_dist;    // [n_pairs]
  const int* species_of;      // [n_atoms] species index in the dense layout
  const double* pair_fc;      // [n_pairs] cutoff function at the pair distance
  const double* pair_dfc;     // [n_pairs] its derivative
  // adjoint outputs of the deterministic force sum: the force of pair p on
  // its neighbour [n_pairs][3], and minus the sum over a centre's pairs
  // (periodic images of the centre itself left out) [n_centres][3]
  double* pair_force = nullptr;
  double* centre_force = nullptr;
};

// Reverse pair lists for the deterministic force sum: for atom j, the pairs
// p (self pairs left out) whose neighbour is j, in no particular order:
// rlist[roff[j] .. roff[j+1]).  count: [n] ints of scratch; roff: n + 1;
// rlist has room for every pair of the list's capacity.
struct ReverseLists {
  int* count = nullptr;
  int64_t* roff = nullptr;
  int64_t* rlist = nullptr;
};
// (clear_count: count is not known to be zero; the fill returns every
// count to zero)
void launch_reverse_pairs(const PairView& pv, const ReverseLists& rl, DeviceBuffer<int64_t>& scan_scratch,
                          hipStream_t stream, bool clear_count);
// grad[j] = centre_force[j] + sum over rlist of j of pair_force[p], in
// ascending pair order (the forces of the adjoint, summed in a fixed order)
void launch_force_gather(const PairView& pv, const ReverseLists& rl, double* grad, hipStream_t stream);

// pair_fc / pair_dfc of the first *n_pairs (device value) of cap pairs
// (cutoff_functions.hh:69-283)
void launch_pair_cutoff(const RadialDev& rad, const int64_t* n_pairs, int64_t cap, const double* dist,
                        double* fc, double* dfc, hipStream_t stream);

// Work partition of the SOAP kernels; 0 selects the built-in choice.  Set
// per batch by the autotuner (rgpu_autotune).
struct SoapTiles {
  int pb_fwd = 0;  // pairs staged per batch in the expansion (<= 64)
  int kc = 0;      // key blocks staged per pass in the power spectrum
  int pb_adj = 0;  // pairs staged per batch in the adjoint (<= 64)
  int keys_adj = 0;  // key blocks staged per pass in the FP32 adjoint
  int adj_kind = 3;  // FP32 adjoint where it applies: 0 staged vector ALUs, 1 matrix cores,
                     // 2 register tiles (keys_adj: key blocks per staged chunk of rounds),
                     // 3 register tiles with the pair phase as a kernel of its own
                     //   (spline radial functions; else as 2).  Kinds 2 and 3 fall back
                     //   to 0 for sizes they are not built for.
  int half_nl = 0;   // expansion over the half neighbour list (1) or the full list (0)
  int fused = -1;    // power spectrum in the expansion's launch (1), separate (0), built-in (-1)
  int fwd_lds = 0;   // expansion accumulating c in LDS and writing it once (1)
  int fwd_group = 0; // FP32 expansion on pair groups where it applies (1) or per entry (0)
};

// c[i][s][n][l*l+l+m] (finalised) and the per-centre key mask.
template <class T>
void launch_expansion(const RadialDev& rad, int S, const PairView& pv, T* c,
                      unsigned long long* keymask, hipStream_t stream,
                      const SoapTiles& tiles = SoapTiles());

// Half neighbour list of a full list (device buffers sized for the full
// list; rows holds n_pairs x N (L+1)^2 values of the expansion precision)
// and the expansion computed over it: the same c and key masks as
// launch_expansion (adaptor_half_neighbour_list.hh:319-380;
// calculator_spherical_expansion.hh:2282-2299).  pos: the atom positions
// the list was built from.
struct HalfList {
  int* count;       // [n_centres] kept pairs
  int64_t* off;     // [n_centres + 1]
  int64_t* src;     // [n_pairs] row read by each full pair (< 0: mirror, -1 - row)
  int* j;           // kept pairs: neighbour atom
  double* vec;      //   [3] pair vector
  double* dist;     //   distance
  double* fc;       //   cutoff function
  double* dfc;      //   its derivative
  void* rows;       //   [N][(L+1)^2] contribution fc R Y
  DeviceBuffer<int64_t>* scan_scratch;
};
// scratch of the half list's scan for n centres (reserved before a capture)
void reserve_half_list_scan(int64_t n, DeviceBuffer<int64_t>& scratch);
template <class T>
void launch_expansion_half(const RadialDev& rad, int S, const PairView& pv, const double* pos,
                           const HalfList& hl, T* c, unsigned long long* keymask, hipStream_t stream,
                           const SoapTiles& tiles = SoapTiles());

// Persistent fused expansion and power spectrum (one launch): c and the
// key masks as launch_expansion, the invariants as launch_power_spectrum
// (T double: p) or launch_power_spectrum_split (T float: p_hi / p_lo), and
// norm2; c stays in LDS between the stages and is written to HBM once.
// With p and p_hi null only the expansion runs, c written once.
template <class T>
void launch_expansion_ps(const RadialDev& rad, int S, const PairView& pv, T* c,
                         unsigned long long* keymask, T* p, unsigned short* p_hi, unsigned short* p_lo,
                         double* norm2, hipStream_t stream, const SoapTiles& tiles = SoapTiles());

// p[i][key(a<=b)][n1*N+n2][l] UN-normalised, and norm2[i] = |p_i|^2 (FP64).
template <class T>
void launch_power_spectrum(const RadialDev& rad, int S, int64_t n_centres, const T* c, T* p,
                           double* norm2, hipStream_t stream, const SoapTiles& tiles = SoapTiles());

// The same power spectrum in the packed layout (packing.hh) written as the
// matrix-core operand planes p = p_hi + p_lo (BF16 bit patterns, each
// rounded to nearest even), rows of packed_row_stride() columns (zero pad);
// norm2 from the FP32 values.
void launch_power_spectrum_split(const RadialDev& rad, int S, int64_t n_centres, const float* c,
                                 unsigned short* p_hi, unsigned short* p_lo, double* norm2,
                                 hipStream_t stream, const SoapTiles& tiles = SoapTiles());

// Normalised copy X = p / |p| (or X = p), FP64 output for export: n rows,
// row r from row rows[r] of p (rows null: row r).
template <class T>
void launch_normalise(int64_t n, int64_t D, const T* p, const double* norm2, bool normalize,
                      double* out, hipStream_t stream, const int64_t* rows = nullptr);

// Stored expansion derivatives (FP64) for the feature-gradient export:
// dcp[p][3][N][(L+1)^2] per pair (the self slot unused) and the centre
// blocks dcs[i][3][S][N][(L+1)^2] (calculator_spherical_expansion.hh
// :2306-2401).
void launch_pair_gradients(const RadialDev& rad, int S, const PairView& pv, double* dcp, double* dcs,
                           hipStream_t stream);
// Feature-gradient rows out[3 n_pairs][Dout] (row 3 r + x = dX_i / dr_x of
// pair row r, self pair first per centre) from c (FP64, full layout), the
// un-normalised p of the output columns p_rows[n_centres][Dout], norm2 and
// the derivatives above; col (may be null) maps output columns to the full
// layout (-1: zero column).
void launch_feature_gradient_rows(int N, int L1, int S, const PairView& pv, int64_t n_pairs,
                                  const double* c, const double* p_rows, const double* norm2,
                                  bool normalize, const double* dcp, const double* dcs, int64_t Dout,
                                  const int64_t* col, double* out, hipStream_t stream);
// coefficient_subselection as device arrays: col[k] (Dsel; -1 for a
// coefficient outside the layout) and its inverse, the selections k naming
// full-layout column o in ascending order at inv_k[inv_off[o] ..
// inv_off[o+1])
struct SelectionMap {
  const int64_t* col = nullptr;
  const int64_t* inv_off = nullptr;
  const int64_t* inv_k = nullptr;
};
// xs[i][k] = p[i][col[k]], norm2[i] = |xs_i|^2, then p[i] = S^T xs_i (the
// selected coefficients, zero elsewhere)
void launch_sparsify(int64_t n, int64_t D, int64_t Dsel, const SelectionMap& sel, double* p, double* xs,
                     double* norm2, hipStream_t stream);
// dst[i][:D] = S^T src[i][:Dsel] (repeated selections summed in ascending
// selection order)
void launch_scatter_columns(int64_t n, int64_t Dsel, const SelectionMap& sel, const double* src, int64_t D,
                            double* dst, hipStream_t stream);

// Adjoint (FP64): given G = dL/dX per centre (X the exported features), write
// the pair and centre forces of dL/dr (pv.pair_force, pv.centre_force;
// launch_force_gather sums them per atom) and per-centre virials
// centre_virial[n_centres][6] (sum over the centre's pairs of r_ji (x)
// dL/dr_j, Voigt xx,yy,zz,yz,xz,xy, not divided by the volume).  xg, when
// not null, holds X_i . G_i per centre (times xg_scale); otherwise it is
// reduced from p and G.
template <class T>
void launch_backward(const RadialDev& rad, int S, const PairView& pv, const T* c, const T* p,
                     const double* norm2, bool normalize, const double* xg, double xg_scale,
                     const T* G, double* centre_virial, hipStream_t stream,
                     const SoapTiles& tiles = SoapTiles());

// launch_backward<double> for a rank-one dL/dX: G_i = Gscale[i] * Grow
// (Grow [D], one sparse point's features; centres with a zero scale are
// skipped and get a zero virial).  Used for the K_NM gradient rows.
void launch_backward_rank1(const RadialDev& rad, int S, const PairView& pv, const double* c,
                           const double* p, const double* norm2, bool normalize, const double* Grow,
                           const double* Gscale, double* centre_virial,
                           hipStream_t stream, const SoapTiles& tiles = SoapTiles());

// Pack dL/dX [n][D] (full layout, FP64) into the packed layout of the mixed
// path, rows of packed_row_stride() columns (FP32): diagonal-key entries
// (G[n1][n2] + G[n2][n1]) / sqrt 2
// for n1 < n2, G[n][n] on the diagonal.
void launch_pack_gradient(int S, int N, int L1, int64_t n, const double* G, float* Gp,
                          hipStream_t stream);

// Adjoint of the mixed path: dL/dX (G) and p in the packed layout, p as its
// BF16 operand planes (launch_power_spectrum_split).  work: room for
// n_centres S (L+1)^2 N floats (dL/dc between the key contraction and the
// pair kernel of adj_kind 3), or null.
void launch_backward_split(const RadialDev& rad, int S, const PairView& pv, const float* c,
                           const unsigned short* p_hi, const unsigned short* p_lo,
                           const double* norm2, bool normalize, const double* xg, double xg_scale,
                           const float* G, double* centre_virial, hipStream_t stream,
                           const SoapTiles& tiles = SoapTiles(), float* work = nullptr);

// scale[i] = zeta d[i*ld]^(zeta-1) for the centres of species s, else 0
void launch_knm_gradient_scale(int64_t n, const double* d, int64_t ld, const int* species_of, int s,
                               int zeta, double* scale, hipStream_t stream);
// p[0, n) = v (a kernel: no memset node when the stream is captured)
void launch_fill(double* p, int64_t n, double v, hipStream_t stream);
// dst[r * ld] = src[r], r < n
void launch_strided_copy(const double* src, int64_t n, double* dst, int64_t ld, hipStream_t stream);

// out[s][:w] += (scale ? scale[s] : 1) * sum of in[a][:w] over a in
// [seg_off[s], seg_off[s+1])
void launch_segment_sum(const double* in, int w, const int64_t* seg_off, int n_seg, double* out,
                        hipStream_t stream, const double* scale = nullptr);
// two segment sums over the same segments in one launch (each as
// launch_segment_sum)
void launch_segment_sum2(const double* in0, int w0, double* out0, const double* scale0, const double* in1, int w1,
                         double* out1, const double* scale1, const int64_t* seg_off, int n_seg,
                         hipStream_t stream);
// raw radial integrals (FP64 Hyp1f1) raw[pt][n][l] (+ derivatives, may be null)
void launch_radial_direct(const RadialDev& rad, const double* r, int64_t n_pts, double* raw,
                          double* draw, hipStream_t stream);
// out[pt][n][l] = sum_m O[m][n] raw[pt][m][l]
void launch_radial_finalise(const RadialDev& rad, const double* raw, int64_t n_pts, double* out,
                            hipStream_t stream);
// spline lookup in precision T: R[pt][n][l], dR (may be null)
template <class T>
void launch_radial_spline_eval(const RadialDev& rad, const double* r, int64_t n_pts, double* R,
                               double* dR, hipStream_t stream);
// Y_lm (+ gradients) at unit vectors in precision T
template <class T>
void launch_ylm_eval(const RadialDev& rad, const double* dirs, int64_t n, double* Y, double* dY,
                     hipStream_t stream);

}  // namespace rgpu
