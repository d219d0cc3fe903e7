// Radial basis tables shared by the host set-up and the HIP kernels.
//
// Host side (radial_host.cc): Gaussian-type-orbital constants, the overlap
// orthonormalisation O = diag(norm) S^{-1/2}, the per-(n,l) Hyp1f1 series /
// asymptotic coefficients and switching points, and the uniform cubic-spline
// tables of O^T R(r); for the DVR basis the Gauss-Legendre points and
// weights and the Hyp1f1 of its top two Bessel orders (O = identity).
// Device side (soap_device.hh): evaluation of R(r) directly (FP64) or by
// the spline (FP64 or FP32), Y_lm and cutoff functions.
#pragma once

#include <vector>

#include "common.hh"
#include "hypers.hh"

namespace rgpu {

constexpr int kMaxN = 24;
constexpr int kMaxL1 = 17;

// One confluent hypergeometric function 1F1(a, b, z), a = (n+l+3)/2,
// b = l + 3/2 (math/hyp1f1.hh:49-229).
struct HypCoef {
  double a, b;
  double prefac_s;  // Gamma(a)/Gamma(b)
  double prefac_a;  // Gamma(b)/Gamma(a)
  double z_switch;  // series for z <= z_switch, asymptotic above
  int is_exp;       // a == b: 1F1 = exp(z)
  int is_n_and_l;   // 2(a-b) integer
};

// Gradient of Y_lm on the unit sphere (no 1/r factor) from the harmonics of
// degree l - 1: the gradient of r^l Y_lm is a harmonic polynomial of degree
// l - 1, so dY_c[lm] = c0 Y[i0] + c1 Y[i1] - l r_c Y[lm] with constant
// coefficients (at most two terms per component for the real harmonics).
template <class T>
struct YGradTerm {
  T c0, c1;
  int i0, i1;
};

// Everything a kernel needs to evaluate the radial part, passed by value.
struct RadialDev {
  int N, L, L1, LL;       // max_radial, max_angular, L+1, (L+1)^2
  int use_spline;         // 1: spline table, 0: direct evaluation
  int basis;              // 0 GTO, 1 DVR
  int recursion;          // downward l-recursion (L >= 3)
  int cutoff_type;        // 0 ShiftedCosine, 1 RadialScaling
  double rc, sw, rs_rate, rs_scale, rs_exponent;
  double fac_a;           // 1 / (2 sigma^2)
  // device arrays
  const double* fac_b;    // [N]
  const double* O;        // [N][N] = diag(norm) S^{-1/2}
  const double* center;   // [N] centre term (finalised when use_spline)
  const HypCoef* hyp;     // [N][L+1]
  const double* hc;       // [N][L+1][2][kHypTerms] series coeff, derivative coeff
  const double* ha;       // [N][L+1][2][kHypTerms] asymptotic coeff, derivative coeff
  // DVR (calculator_spherical_expansion.hh:738-930): Gauss-Legendre points
  // x_n on [0, rc + 3 sigma] and factors sqrt(w_n) x_n; hyp / hc / ha hold
  // 1F1(o + 1, 2 o + 2, z) of the Bessel orders o = L, L + 1 at entries 0, 1
  const double* dvr_x;    // [N]
  const double* dvr_f;    // [N]
  // RadialDimReduction: n_tabs spline tables, each on its own grid of
  // tab_n[t] points (spacing tab_h[t], tab_pu[t] points per unit length,
  // points at tab_grid + tab_grid0[t]), stored one after another from cell
  // tab_cell0[t]; sp_tab [S] maps a batch species to its table (null with
  // one table, which then uses the scalar grid above); center holds n_tabs
  // centre terms of N values
  int n_tabs;
  const int* sp_tab;
  const int* tab_n;
  const int* tab_cell0;
  const int* tab_grid0;
  const double* tab_h;
  const double* tab_pu;
  const double* tab_grid;
  // spline on [x1, x2], grid_n points, packed per cell k and function
  // f = n*L1 + l as {y_k, d_k, y_k+1, d_k+1} with d = y'' h^2/6, so that one
  // vector load serves one evaluation
  int grid_n, search_size;
  double x1, h, per_unit;
  const double* grid;     // [grid_n] grid points
  const double* spl64;    // [grid_n-1][F][4]
  const float* spl32;     // [grid_n-1][F][4]
  // real spherical harmonics recursion coefficients, ylm_table_size(L+1)
  // values: [L1][L1][4] = {a_lm, b_lm, f_lm, f'_lm}, then d_m, sqrt(2l+1),
  // sqrt(l(l+1)) ([L1] each); see ylm_coefficients()
  const double* ylm64;
  const float* ylm32;
  // [LL][3] gradient terms (YGradTerm), FP64 and FP32
  const YGradTerm<double>* ygrad64;
  const YGradTerm<float>* ygrad32;
};

__host__ __device__ inline int ylm_table_size(int L1) { return 4 * L1 * L1 + 3 * L1; }

// Coefficients of the associated-Legendre recursion and of the harmonics'
// gradient (math/spherical_harmonics.cc:36-96, 249-300), all FP64:
//   a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = -sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))
//   (m < l-1); f_lm = sqrt((l-m)(l+m+1)), f'_lm = sqrt((l-m-1)(l+m+2))
//   (m < l); d_1 = -sqrt(3/2), d_m = -sqrt(1 + 1/(2m)).
std::vector<double> ylm_coefficients(int L);

// Host-side constants (all FP64).
struct RadialHost {
  int N = 0, L = 0, basis = 0;
  double fac_a = 0;
  std::vector<double> sigma, fac_b, norm, nfac, O, center_raw, center_fin;
  std::vector<double> dvr_x, dvr_f;  // DVR points and sqrt(w) x factors
  // RadialDimReduction: spline tables, one per species (1 without it), the
  // species' atomic numbers, projections [species][L+1][N][N] and centre
  // terms [species][N]
  int n_tabs = 1;
  std::vector<int> tab_species;
  std::vector<double> proj, tab_centers;
  std::vector<HypCoef> hyp;
  std::vector<double> hc, ha;
  // spline
  int grid_n = 0;
  double x1 = 0, x2 = 0, h = 0, per_unit = 0;
  std::vector<double> spl64;

  explicit RadialHost(const SoapHypers& hp);
  void set_projections(const SoapHypers& hp);
};

// Natural cubic spline second derivatives (times h^2/6) of the columns of
// y[n][F] on a uniform grid with spacing h (math/interpolator.cc:148-175).
void spline_second_derivatives(int n, int F, double h, const std::vector<double>& y,
                               std::vector<double>& d2);

// Eigen 3.3 LinSpaced for floating point (used for the spline grids).
std::vector<double> linspaced(int num, double lo, double hi);

// Gauss-Legendre points (ascending) and weights on [r0, r1]
void gauss_legendre(double r0, double r1, int order_n, std::vector<double>& x,
                    std::vector<double>& w);

}  // namespace rgpu
