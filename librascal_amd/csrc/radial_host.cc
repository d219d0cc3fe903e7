// Host set-up of the GTO radial basis (FP64).
//
// Follows representations/calculator_spherical_expansion.hh:322-728 for the
// GTO constants and the orthonormalisation, and math/hyp1f1.cc:51-331 for the
// Hyp1f1 coefficient tables and the series/asymptotic switching point.
#include <algorithm>
#include <cmath>

#include "radial.hh"

namespace rgpu {

namespace {

double pow_u(double x, size_t n) {  // math/utils.hh:85-100
  double v = 1.0;
  do {
    if (n & 1) v *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return v;
}
double pow_i(double x, int n) {
  return n < 0 ? pow_u(1.0 / x, size_t(-n)) : pow_u(x, size_t(n));
}

// Evaluation of 1F1 itself on the host, used only to place the switching
// point between the series and the asymptotic expansion.
struct HostHyp {
  double a, b, tol;
  bool is_exp, is_n_and_l;
  double prefac_a;
  std::vector<double> c, ac;
  HostHyp(double a_, double b_, double tol_) : a(a_), b(b_), tol(tol_) {
    is_exp = std::fabs(1 - b / a) < kDblFtol;
    double ip;
    is_n_and_l = std::fabs(std::modf(2 * (a - b), &ip)) < 1e-14;
    prefac_a = std::tgamma(b) / std::tgamma(a);
    if (!is_exp) {
      c.resize(kHypTerms);
      ac.resize(kHypTerms);
      c[0] = a / b;
      for (int i = 1; i < kHypTerms; ++i) c[i] = c[i - 1] * ((a + i) / ((i + 1) * (b + i)));
      double bma = b - a, oma = 1 - a;
      ac[0] = bma * oma;
      for (int i = 1; i < kHypTerms; ++i) ac[i] = ac[i - 1] * ((bma + i) * (oma + i) / (i + 1));
    }
  }
  double series(double z) const {
    if (is_exp) return std::exp(z);
    double res = 1.0, zpow = z, z4 = z * z;
    z4 *= z4;
    for (int i = 0; i < kHypTerms - 3; i += 4) {
      double a1 = zpow * (c[i] + z * (c[i + 1] + z * (c[i + 2] + z * c[i + 3])));
      if (a1 < tol * res) {
        res += a1;
        break;
      }
      zpow *= z4;
      res += a1;
    }
    return res;
  }
  double asym(double z) const {
    if (is_exp) return std::exp(z);
    double iz = 1.0 / z, res = 1.0, izpow = 1.0;
    for (int i = 0; i < kHypTerms; ++i) {
      izpow *= iz;
      double s = ac[i] * izpow;
      if (res > 0 && std::fabs(s) < tol * res) break;
      res += s;
    }
    double zab = is_n_and_l ? std::sqrt(pow_i(z, int(2 * (a - b)))) : std::pow(z, a - b);
    double r = prefac_a * std::exp(z) * zab * res;
    if (std::isnan(r)) r = INFINITY;
    return r;
  }
  // hyp1f1.cc:289-331
  double switching_point() const {
    if (std::fabs(1 - b / a) <= kDblFtol) return 1.0;
    const int max_it = 100;
    double z = 1.0;
    double hs = series(z), ha = asym(z);
    double z_below = z, z_above = z;
    if (std::fabs(1 - hs / ha) > tol) {
      for (int it = 0; std::fabs(1 - hs / ha) > 100 * tol && it < max_it; ++it) {
        z *= 1.5;
        hs = series(z);
        ha = asym(z);
      }
      z_above = z;
    } else {
      for (int it = 0; std::fabs(1 - hs / ha) < 100 * tol && it < max_it; ++it) {
        z *= 0.5;
        hs = series(z);
        ha = asym(z);
      }
      z_below = z;
    }
    z = (z_above + z_below) * 0.5;
    hs = series(z);
    ha = asym(z);
    for (int it = 0; z_above - z_below > tol && it < max_it; ++it) {
      if (std::fabs(1 - hs / ha) > tol)
        z_below = z;
      else
        z_above = z;
      z = (z_above + z_below) * 0.5;
      hs = series(z);
      ha = asym(z);
    }
    return z;
  }
};

// Cyclic Jacobi eigen-decomposition of a small symmetric matrix.
void jacobi(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

}  // namespace

std::vector<double> linspaced(int num, double lo, double hi) {
  std::vector<double> g(num);
  if (num == 1) {
    g[0] = hi;
    return g;
  }
  double step = (hi - lo) / double(num - 1);
  bool flip = std::fabs(hi) < std::fabs(lo);
  int s1 = num - 1;
  for (int i = 0; i < num; ++i) {
    if (flip)
      g[i] = (i == 0) ? lo : (hi - double(s1 - i) * step);
    else
      g[i] = (i == s1) ? hi : (lo + double(i) * step);
  }
  return g;
}

void spline_second_derivatives(int n, int F, double h, const std::vector<double>& y,
                               std::vector<double>& d2) {
  std::vector<double> y2(size_t(n) * F, 0.0), u(size_t(n) * F, 0.0);
  const double sig = 0.5;
  for (int i = 1; i < n - 1; ++i)
    for (int k = 0; k < F; ++k) {
      double p = sig * y2[size_t(i - 1) * F + k] + 2.0;
      y2[size_t(i) * F + k] = (sig - 1.0) / p;
      double uu = 3. * (y[size_t(i + 1) * F + k] - 2 * y[size_t(i) * F + k] +
                        y[size_t(i - 1) * F + k]) /
                  (h * h);
      uu -= sig * u[size_t(i - 1) * F + k];
      uu /= p;
      u[size_t(i) * F + k] = uu;
    }
  for (int k = 0; k < F; ++k) y2[size_t(n - 1) * F + k] = 0.0;
  for (int i = n - 2; i > 0; --i)
    for (int k = 0; k < F; ++k)
      y2[size_t(i) * F + k] = y2[size_t(i) * F + k] * y2[size_t(i + 1) * F + k] + u[size_t(i) * F + k];
  for (int k = 0; k < F; ++k) y2[k] = 0.0;
  const double h_sq_6 = h * h / 6.0;
  d2.resize(y2.size());
  for (size_t k = 0; k < y2.size(); ++k) d2[k] = y2[k] * h_sq_6;
}

// Series / asymptotic coefficients of 1F1(a, b, z) (and of its derivative)
// into entry idx of the tables (hyp1f1.hh:49-229).
static void fill_hyp(RadialHost& h, int idx, double a, double b, double tol) {
  HypCoef& c = h.hyp[idx];
  c.a = a;
  c.b = b;
  c.prefac_s = std::tgamma(a) / std::tgamma(b);
  c.prefac_a = std::tgamma(b) / std::tgamma(a);
  c.is_exp = std::fabs(1 - b / a) < kDblFtol;
  double ip;
  c.is_n_and_l = std::fabs(std::modf(2 * (a - b), &ip)) < 1e-14;
  HostHyp hh(a, b, tol);
  c.z_switch = hh.switching_point();
  if (c.is_exp) return;
  double* s0 = &h.hc[(size_t(idx) * 2 + 0) * kHypTerms];
  double* s1 = &h.hc[(size_t(idx) * 2 + 1) * kHypTerms];
  s0[0] = a / b;
  s1[0] = (a + 1) / (b + 1);
  for (int i = 1; i < kHypTerms; ++i) {
    s0[i] = s0[i - 1] * ((a + i) / ((i + 1) * (b + i)));
    s1[i] = s1[i - 1] * ((a + i + 1) / ((i + 1) * (b + i + 1)));
  }
  double* a0 = &h.ha[(size_t(idx) * 2 + 0) * kHypTerms];
  double* a1 = &h.ha[(size_t(idx) * 2 + 1) * kHypTerms];
  double bma = b - a, oma = 1 - a, oma1 = 1 - a - 1;
  a0[0] = bma * oma;
  a1[0] = bma * oma1;
  for (int i = 1; i < kHypTerms; ++i) {
    a0[i] = a0[i - 1] * ((bma + i) * (oma + i) / (i + 1));
    a1[i] = a1[i - 1] * ((bma + i) * (oma1 + i) / (i + 1));
  }
}

// Gauss-Legendre points and weights on [r0, r1], order_n points, ascending
// (math/gauss_legendre.cc:30-73: Newton iteration on the Legendre
// three-term recurrence from the starting guesses cos(pi (i + 3/4) / (n +
// 1/2)), the symmetric half mirrored).  The derivative pp is carried from
// one root to the next, as in the reference: for an odd order the middle
// guess cos(pi/2) ~ 6e-17 is already within the tolerance of z1 = 0, the
// Newton loop does not run, and the middle weight is formed with the
// previous root's derivative.  That weight is not the true quadrature
// weight, but DVR models built by the reference use it, so the basis
// (and every feature of such a model) is reproduced exactly.
void gauss_legendre(double r0, double r1, int order_n, std::vector<double>& x,
                    std::vector<double>& w) {
  if (order_n < 2)
    throw RuntimeError("there should be at least 2 integration points but: order_n < 2 == true");
  x.assign(order_n, 0.0);
  w.assign(order_n, 0.0);
  const double mid = 0.5 * (r0 + r1), half = 0.5 * (r1 - r0);
  const int m = (order_n + 1) / 2;
  double pp = 0;
  for (int i = 0; i < m; ++i) {
    double z = std::cos(3.14159265358979323846 * (i + 0.75) / (order_n + 0.5)), z1 = 0;
    while (std::fabs(z - z1) > kDblFtol) {
      double p1 = 1, p2 = 0;
      for (int j = 0; j < order_n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j + 1.0) * z * p2 - j * p3) / (j + 1);
      }
      pp = order_n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    }
    x[i] = mid - half * z;
    x[order_n - 1 - i] = mid + half * z;
    w[i] = (2.0 * half) / ((1.0 - z * z) * pp * pp);
    w[order_n - 1 - i] = w[i];
  }
}

// RadialDimReduction: the centre term of species k is P_k[l=0] times the
// finalised centre term (calculator_spherical_expansion.hh:1417-1428)
void RadialHost::set_projections(const SoapHypers& hp) {
  tab_species = hp.rdr_species;
  proj = hp.rdr_proj;
  n_tabs = tab_species.empty() ? 1 : int(tab_species.size());
  if (tab_species.empty()) {
    tab_centers = center_fin;
    return;
  }
  const int L1 = L + 1;
  tab_centers.assign(size_t(n_tabs) * N, 0.0);
  for (int k = 0; k < n_tabs; ++k)
    for (int c = 0; c < N; ++c) {
      const double* P0 = &proj[((size_t(k) * L1 + 0) * N + c) * N];
      double s = 0;
      for (int n = 0; n < N; ++n) s += P0[n] * center_fin[n];
      tab_centers[size_t(k) * N + c] = s;
    }
}

RadialHost::RadialHost(const SoapHypers& hp) : N(hp.max_radial), L(hp.max_angular) {
  basis = hp.radial_basis;
  fac_a = 0.5 * pow_i(hp.gaussian_sigma, -2);
  const double tol = 1e-13;
  if (basis == 1) {
    // DVR (calculator_spherical_expansion.hh:809-848): points and weights on
    // [0, rc + 3 sigma], factor sqrt(w) x, centre term sqrt(w) x exp(-a x^2),
    // no orthonormalisation
    std::vector<double> wts;
    gauss_legendre(0.0, hp.cutoff + 3 * hp.gaussian_sigma, N, dvr_x, wts);
    dvr_f.resize(N);
    center_raw.resize(N);
    for (int n = 0; n < N; ++n) {
      dvr_f[n] = std::sqrt(wts[n]) * dvr_x[n];
      center_raw[n] = dvr_f[n] * std::exp(-fac_a * dvr_x[n] * dvr_x[n]);
    }
    center_fin = center_raw;
    sigma.assign(N, 0.0);
    fac_b.assign(N, 0.0);
    norm.assign(N, 1.0);
    nfac.assign(N, 0.0);
    O.assign(N * N, 0.0);
    for (int n = 0; n < N; ++n) O[n * N + n] = 1.0;
    // 1F1(o + 1, 2 o + 2, .) of the two top Bessel orders o = L, L + 1 (one
    // order above L for the gradient recursion; math/bessel.cc:24-58)
    hyp.resize(2);
    hc.assign(size_t(2) * 2 * kHypTerms, 0.0);
    ha.assign(size_t(2) * 2 * kHypTerms, 0.0);
    for (int k = 0; k < 2; ++k) fill_hyp(*this, k, double(L + k + 1), double(2 * (L + k) + 2), tol);
    set_projections(hp);
    return;
  }
  // calculator_spherical_expansion.hh:621-636
  sigma.resize(N);
  fac_b.resize(N);
  norm.resize(N);
  nfac.resize(N);
  for (int n = 0; n < N; ++n) {
    sigma[n] = std::max(std::sqrt(double(n)), 1.0) * hp.cutoff / double(N);
    fac_b[n] = 0.5 * pow_i(sigma[n], -2);
    norm[n] = 0.25 * std::sqrt(2.0 / (std::tgamma(1.5 + n) * pow_u(sigma[n], 3 + 2 * n)));
    nfac[n] = std::tgamma(0.5 * (3.0 + n)) / std::tgamma(1.5);
  }
  // :643-680 overlap matrix and its inverse square root
  std::vector<double> S(N * N);
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      double sa = sigma[a], sb = sigma[b];
      S[a * N + b] = std::pow(0.5 / pow_i(sa, 2) + 0.5 / pow_i(sb, 2), -0.5 * (3.0 + a + b)) /
                     (pow_u(sa, a) * pow_u(sb, b)) * std::tgamma(0.5 * (3.0 + a + b)) /
                     (std::pow(sa * sb, 1.5) * std::sqrt(std::tgamma(1.5 + a) * std::tgamma(1.5 + b)));
    }
  std::vector<double> w, V;
  jacobi(N, S, w, V);
  for (double x : w)
    if (!(x > 0)) throw RuntimeError("Warning: Could not diagonalize radial overlap matrix.");
  O.assign(N * N, 0.0);
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      double s = 0;
      for (int k = 0; k < N; ++k) s += V[a * N + k] * (1.0 / std::sqrt(w[k])) * V[b * N + k];
      O[a * N + b] = norm[a] * s;
    }
  // :478-501 centre contribution (Y_00 and 1F1(a,b,0) = 1)
  center_raw.resize(N);
  for (int n = 0; n < N; ++n) {
    double abn = (n % 2 == 0) ? std::sqrt(pow_i(fac_a + fac_b[n], -1 * (3 + n)))
                              : pow_i(fac_a + fac_b[n], -1 * (3 + n) / 2);
    center_raw[n] = nfac[n] * abn;
  }
  center_fin.assign(N, 0.0);
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) center_fin[a] += O[b * N + a] * center_raw[b];
  set_projections(hp);
  // Hyp1f1 tables, tolerance 1e-13 and 200 terms (expansion.hh:697)
  const int L1 = L + 1;
  hyp.resize(N * L1);
  hc.assign(size_t(N) * L1 * 2 * kHypTerms, 0.0);
  ha.assign(size_t(N) * L1 * 2 * kHypTerms, 0.0);
  for (int n = 0; n < N; ++n)
    for (int l = 0; l < L1; ++l) fill_hyp(*this, n * L1 + l, 0.5 * (n + l + 3), l + 1.5, tol);
}

std::vector<double> ylm_coefficients(int L) {
  const int L1 = L + 1;
  std::vector<double> t(size_t(ylm_table_size(L1)), 0.0);
  for (int l = 0; l <= L; ++l)
    for (int m = 0; m <= L; ++m) {
      double* c = &t[size_t(4 * (l * L1 + m))];
      const double ll = double(l) * l, mm = double(m) * m, l1 = double(l - 1) * (l - 1);
      if (m < l - 1) {
        c[0] = std::sqrt((4.0 * ll - 1.0) / (ll - mm));
        c[1] = -std::sqrt((l1 - mm) / (4.0 * l1 - 1.0));
      }
      if (m < l) {
        c[2] = std::sqrt(double(l - m) * double(l + m + 1));
        c[3] = std::sqrt(double(l - m - 1) * double(l + m + 2));
      }
    }
  double* d = &t[size_t(4 * L1 * L1)];
  d[0] = 1.0;
  for (int m = 1; m <= L; ++m) d[m] = m == 1 ? -1.22474487139158904910 : -std::sqrt(1.0 + 0.5 / m);
  for (int l = 0; l <= L; ++l) {
    d[L1 + l] = std::sqrt(double(2 * l + 1));
    d[2 * L1 + l] = std::sqrt(double(l) * double(l + 1));
  }
  return t;
}

}  // namespace rgpu
