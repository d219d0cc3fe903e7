/*
 * FP64 CPU oracle: a plain serial restatement of librascal's SOAP + sparse
 * GAP algorithms.  TEST INFRASTRUCTURE ONLY (see rascal_oracle.h): it is the
 * checker for the HIP library and the timed CPU baseline, never part of the
 * product path.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the librascal source tree, src/rascal/...).  Results follow the reference's
 * arithmetic order where it matters for bit-identity (neighbour-list
 * geometry); elsewhere the formulas are restated and agree to rounding.
 */
#include "rascal_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr double kPi = 3.14159265358979323846;
// math/utils.hh:66
constexpr double kDblFtol = 100.0 * 2.220446049250313080847e-16;
constexpr double kInvSqrt2 = 0.70710678118654752440;
constexpr double kSqrt2 = 1.41421356237309504880;
constexpr double kSqrt3 = 1.73205080756887729353;
// math/bessel.hh:42, lower end of the radial spline range
constexpr double kSplineX1 = 1e-5;

// math/utils.hh:85-117 (repeated squaring integer powers)
double pow_u(double x, size_t n) {
  double v = 1.0;
  do {
    if (n & 1) v *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return v;
}
double pow_i(double x, int n) {
  if (n < 0) return pow_u(1.0 / x, static_cast<size_t>(-n));
  return pow_u(x, static_cast<size_t>(n));
}

/* ------------------------------------------------------------------------
 * Confluent hypergeometric 1F1 (math/hyp1f1.hh:49-229, hyp1f1.cc:41-331)
 * ---------------------------------------------------------------------- */
struct Hyp1f1Series {
  double a, b, prefac, tol;
  size_t mmax;
  bool is_exp = false;
  std::vector<double> c, cd;
  Hyp1f1Series(double a_, double b_, size_t mmax_, double tol_)
      : a(a_), b(b_), prefac(std::tgamma(a_) / std::tgamma(b_)), tol(tol_),
        mmax(mmax_) {
    // hyp1f1.cc:51-70
    if (std::fabs(1 - b / a) < kDblFtol) {
      is_exp = true;
    } else {
      c.resize(mmax);
      cd.resize(mmax);
      c[0] = a / b;
      cd[0] = (a + 1) / (b + 1);
      for (size_t i = 1; i < mmax; ++i) {
        c[i] = c[i - 1] * ((a + i) / ((i + 1) * (b + i)));
        cd[i] = cd[i - 1] * ((a + i + 1) / ((i + 1) * (b + i + 1)));
      }
    }
  }
  // hyp1f1.cc:123-153 (adaptive: four terms per step)
  double sum(double z, const std::vector<double>& co) const {
    double res = 1.0, zpow = z, z4 = z * z;
    z4 *= z4;
    for (size_t i = 0; i < mmax - 3; i += 4) {
      double a1 = zpow * (co[i] + z * (co[i + 1] + z * (co[i + 2] + z * co[i + 3])));
      if (a1 < tol * res) {
        res += a1;
        break;
      }
      zpow *= z4;
      res += a1;
    }
    return res;
  }
  // hyp1f1.cc:84-121
  double f(double z, bool der) const {
    if (is_exp) return std::exp(z);
    return der ? sum(z, cd) * a / b : sum(z, c);
  }
  // hyp1f1.cc:72-82: Gamma(a)/Gamma(b) exp(z2) 1F1
  double g(double z, double z2, double ez2, bool der) const {
    if (!is_exp) return prefac * f(z, der) * ez2;
    return prefac * std::exp(z + z2);
  }
};

struct Hyp1f1Asymptotic {
  double a, b, prefac, tol;
  size_t mmax;
  bool is_n_and_l = false, is_exp = false;
  std::vector<double> c, cd;
  Hyp1f1Asymptotic(double a_, double b_, size_t mmax_, double tol_)
      : a(a_), b(b_), prefac(std::tgamma(b_) / std::tgamma(a_)), tol(tol_),
        mmax(mmax_) {
    // hyp1f1.cc:156-184
    double ip;
    double f2 = std::modf(2 * (a - b), &ip);
    if (std::fabs(f2) < 1e-14) is_n_and_l = true;
    if (std::fabs(1 - b / a) < kDblFtol) {
      is_exp = true;
    } else {
      c.resize(mmax);
      cd.resize(mmax);
      double bma = b - a, oma = 1 - a, oma1 = 1 - a - 1;
      c[0] = bma * oma;
      cd[0] = bma * oma1;
      for (size_t i = 1; i < mmax; ++i) {
        c[i] = c[i - 1] * ((bma + i) * (oma + i) / (i + 1));
        cd[i] = cd[i - 1] * ((bma + i) * (oma1 + i) / (i + 1));
      }
    }
  }
  // hyp1f1.cc:186-195
  double zab(double z) const {
    if (is_n_and_l) return std::sqrt(pow_i(z, static_cast<int>(2 * (a - b))));
    return std::pow(z, a - b);
  }
  // hyp1f1.cc:248-270
  double sum(double z, const std::vector<double>& co) const {
    double iz = 1.0 / z, res = 1.0, izpow = 1.0;
    for (size_t i = 0; i < mmax; ++i) {
      izpow *= iz;
      double s = co[i] * izpow;
      if (res > 0 && std::fabs(s) < tol * res) break;
      res += s;
    }
    return res;
  }
  // hyp1f1.cc:210-222 (the 1F1 itself)
  double f(double z, bool der) const {
    if (is_exp) return std::exp(z);
    double r = prefac * std::exp(z) * zab(z) * sum(z, der ? cd : c);
    if (std::isnan(r)) r = INFINITY;
    return r;
  }
  // hyp1f1.cc:197-208 (G form, prefactors cancel)
  double g(double z, double z2, bool der) const {
    if (!is_exp) return sum(z, der ? cd : c) * std::exp(z + z2) * zab(z);
    return std::exp(z + z2) / prefac;
  }
};

struct Hyp1f1 {
  Hyp1f1Series s;
  Hyp1f1Asymptotic as;
  double a, b, tol, z_asympt = 1.0, hs = 0, ha = 0, z_above = 0, z_below = 0;
  Hyp1f1(double a_, double b_, size_t mmax, double tol_)
      : s(a_, b_, mmax, tol_), as(a_, b_, mmax, tol_), a(a_), b(b_), tol(tol_) {
    // hyp1f1.cc:272-287
    if (std::fabs(1 - b / a) > kDblFtol) {
      find_switching_point();
      update();
    }
  }
  void update() {
    hs = s.f(z_asympt, false);
    ha = as.f(z_asympt, false);
  }
  // hyp1f1.cc:289-331: bracket then bisect the largest z where series and
  // asymptotic forms agree within the tolerance
  void find_switching_point() {
    const int max_it = 100;
    update();
    z_below = z_above = z_asympt;
    if (std::fabs(1 - hs / ha) > tol) {
      for (int it = 0; std::fabs(1 - hs / ha) > 100 * tol && it < max_it; ++it) {
        z_asympt *= 1.5;
        update();
      }
      z_above = z_asympt;
    } else {
      for (int it = 0; std::fabs(1 - hs / ha) < 100 * tol && it < max_it; ++it) {
        z_asympt *= 0.5;
        update();
      }
      z_below = z_asympt;
    }
    z_asympt = (z_above + z_below) * 0.5;
    update();
    for (int it = 0; z_above - z_below > tol && it < max_it; ++it) {
      if (std::fabs(1 - hs / ha) > tol)
        z_below = z_asympt;
      else
        z_above = z_asympt;
      z_asympt = (z_above + z_below) * 0.5;
      update();
    }
  }
  // hyp1f1.hh:157-166
  double f(double z, bool der) const {
    return z > z_asympt ? as.f(z, der) : s.f(z, der);
  }
  // hyp1f1.cc:345-354
  double g(double z, double z2, double ez2, bool der) const {
    return z > z_asympt ? as.g(z, z2, der) : s.g(z, z2, ez2, der);
  }
};

// hyp1f1.cc:41-49
inline double rec_val_down(double a, double b, double z, double M1p2p,
                           double M1p1p) {
  return (z * (a - b) * M1p2p + M1p1p * b) / a;
}
inline double rec_der_down(double, double b, double z, double M2p3p,
                           double M1p2p) {
  return z * M2p3p + M1p2p * (b + 1);
}

// Hyp1f1SphericalExpansion (hyp1f1.hh:249-297, hyp1f1.cc:333-491)
struct Hyp1f1Sph {
  int N = 0, L = 0;
  bool recursion;
  double tol;
  size_t mmax;
  std::vector<Hyp1f1> h;
  std::vector<double> vals, ders;  // [N][L+1]
  Hyp1f1Sph(bool rec, double tol_, size_t mmax_)
      : recursion(rec), tol(tol_), mmax(mmax_) {}
  void precompute(int n_max, int l_max) {
    N = n_max;
    L = l_max;
    vals.assign(N * (L + 1), 0.0);
    ders.assign(N * (L + 1), 0.0);
    h.clear();
    for (int n = 0; n < N; ++n)
      for (int l = 0; l <= L; ++l)
        h.emplace_back(0.5 * (n + l + 3), l + 1.5, mmax, tol);
  }
  static double ga(int n, int l) { return 0.5 * (n + l + 3); }
  static double gb(int l) { return l + 1.5; }
  void calc(double r, double alpha, const double* fac_b, bool derivative) {
    if (!recursion || L < 3)
      calc_direct(r, alpha, fac_b, derivative);
    else
      calc_recursion(r, alpha, fac_b);
  }
  void calc_recursion(double r, double alpha, const double* fac_b) {
    double arij = alpha * r, z2 = -r * arij, ez2 = std::exp(z2);
    for (int n = 0; n < N; ++n) {
      double zz = (arij * alpha) * (1.0 / (alpha + fac_b[n]));
      double dzdr = zz * 2;
      double z = zz * r;
      double* v = &vals[n * (L + 1)];
      double* d = &ders[n * (L + 1)];
      int l = L;
      const Hyp1f1& hl = h[n * (L + 1) + l];
      double M1p2p = hl.g(z, z2, ez2, false), M2p3p = hl.g(z, z2, ez2, true);
      v[l] = M1p2p;
      d[l] = M2p3p;
      const Hyp1f1& hm = h[n * (L + 1) + l - 1];
      double MP1p2p = hm.g(z, z2, ez2, false), MP2p3p = hm.g(z, z2, ez2, true);
      v[l - 1] = MP1p2p;
      d[l - 1] = MP2p3p;
      for (l -= 2; l > 0; l -= 2) {
        double a = ga(n, l), b = gb(l);
        double M1p1p = rec_der_down(a, b, z, M2p3p, M1p2p);
        double Moo = rec_val_down(a, b, z, M1p2p, M1p1p);
        v[l] = Moo;
        d[l] = M1p1p;
        M2p3p = M1p1p;
        M1p2p = Moo;
        a = ga(n, l - 1);
        b = gb(l - 1);
        double MP1p1p = rec_der_down(a, b, z, MP2p3p, MP1p2p);
        double MPoo = rec_val_down(a, b, z, MP1p2p, MP1p1p);
        v[l - 1] = MPoo;
        d[l - 1] = MP1p1p;
        MP2p3p = MP1p1p;
        MP1p2p = MPoo;
      }
      if (L % 2 == 0) {
        double a = ga(n, 0), b = gb(0);
        double M1p1p = rec_der_down(a, b, z, M2p3p, M1p2p);
        double Moo = rec_val_down(a, b, z, M1p2p, M1p1p);
        v[0] = Moo;
        d[0] = M1p1p;
      }
      for (int k = 0; k <= L; ++k) d[k] *= dzdr;
    }
    for (int k = 0; k < N * (L + 1); ++k) ders[k] -= (2 * arij) * vals[k];
  }
  void calc_direct(double r, double alpha, const double* fac_b, bool der) {
    double arij = alpha * r, z2 = -r * arij, ez2 = std::exp(z2);
    for (int n = 0; n < N; ++n) {
      double zz = (arij * alpha) * (1.0 / (alpha + fac_b[n]));
      double z = zz * r;
      for (int l = 0; l <= L; ++l)
        vals[n * (L + 1) + l] = h[n * (L + 1) + l].g(z, z2, ez2, false);
    }
    if (der) {
      for (int n = 0; n < N; ++n) {
        double zz = (arij * alpha) * (1.0 / (alpha + fac_b[n]));
        double z = zz * r, dzdr = zz * 2;
        for (int l = 0; l <= L; ++l)
          ders[n * (L + 1) + l] = h[n * (L + 1) + l].g(z, z2, ez2, true) * dzdr;
      }
      for (int k = 0; k < N * (L + 1); ++k)
        ders[k] -= (2 * alpha * r) * vals[k];
    }
  }
};

/* ------------------------------------------------------------------------
 * Real spherical harmonics (math/spherical_harmonics.cc:36-363)
 * ---------------------------------------------------------------------- */
struct Ylm {
  int L = 0;
  std::vector<double> ca, cb, acoef1, acoef2, plm;  // plm[(L+1)*(L+1)]
  std::vector<double> alp;                          // [(L+1)][(L+2)]
  std::vector<double> cosm, sinm;                   // [L+1]
  std::vector<double> Y, dY;                        // [(L+1)^2], [3][(L+1)^2]
  void precompute(int lmax) {
    L = lmax;
    int W = 2 * L + 1;
    ca.assign((L + 1) * W, 0.0);
    cb.assign((L + 1) * W, 0.0);
    // spherical_harmonics.cc:56-67
    for (int l = 0; l <= L; ++l) {
      double lsq = double(l) * l, lm1sq = double(l - 1) * (l - 1);
      for (int m = 0; m <= l; ++m) {
        double msq = double(m) * m;
        ca[l * W + m] = std::sqrt((4 * lsq - 1.0) / (lsq - msq));
        cb[l * W + m] = -1.0 * std::sqrt((lm1sq - msq) / (4 * lm1sq - 1.0));
      }
    }
    acoef1.assign(L + 1, 0.0);
    acoef2.assign(L + 1, 0.0);
    for (int l = 0; l <= L; ++l) {
      acoef1[l] = std::sqrt(2 * l + 1);
      acoef2[l] = -std::sqrt(1.0 + 0.5 / l);
    }
    plm.assign((L + 1) * (L + 1), 0.0);
    for (int l = 0; l <= L; ++l)
      for (int m = 0; m <= l; ++m)
        plm[l * (L + 1) + m] = std::sqrt(double(l - m) * double(l + m + 1));
    alp.assign((L + 1) * (L + 2), 0.0);
    cosm.assign(L + 1, 0.0);
    sinm.assign(L + 1, 0.0);
    Y.assign((L + 1) * (L + 1), 0.0);
    dY.assign(3 * (L + 1) * (L + 1), 0.0);
  }
  double& P(int l, int m) { return alp[l * (L + 2) + m]; }
  // spherical_harmonics.cc:97-133
  void legendre(double cos_t) {
    double sin_t = std::sqrt(1.0 - cos_t * cos_t);
    const double s2pi = std::sqrt(0.5 / kPi);
    double acc = s2pi;
    P(0, 0) = s2pi;
    if (L > 0) {
      P(1, 0) = cos_t * kSqrt3 * s2pi;
      acc = acc * -std::sqrt(3.0 / 2.0) * sin_t;
      P(1, 1) = acc;
    }
    int W = 2 * L + 1;
    for (int l = 2; l <= L; ++l) {
      for (int m = 0; m < l - 1; ++m) {
        double t = cos_t * P(l - 1, m);
        t += cb[l * W + m] * P(l - 2, m);
        t *= ca[l * W + m];
        P(l, m) = t;
      }
      P(l, l - 1) = acc * cos_t * acoef1[l];
      acc = acc * sin_t * acoef2[l];
      P(l, l) = acc;
    }
  }
  // spherical_harmonics.cc:135-225
  void calc(const double* dir_in, bool deriv) {
    double dir[3] = {dir_in[0], dir_in[1], dir_in[2]};
    double n2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    if (std::fabs(n2 - 1.0) > kDblFtol) {
      double nn = std::sqrt(n2);
      for (int k = 0; k < 3; ++k) dir[k] /= nn;
    }
    double cos_t = dir[2];
    double sxy = std::hypot(dir[0], dir[1]);
    double cos_p = 1.0, sin_p = 0.0;
    if (sxy >= kDblFtol) {
      cos_p = dir[0] / sxy;
      sin_p = dir[1] / sxy;
    }
    legendre(cos_t);
    for (int m = 0; m <= L; ++m) {
      if (m == 0) {
        cosm[0] = 1.0;
        sinm[0] = 0.0;
      } else if (m == 1) {
        cosm[1] = cos_p;
        sinm[1] = sin_p;
      } else {
        cosm[m] = 2.0 * cos_p * cosm[m - 1] - cosm[m - 2];
        sinm[m] = 2.0 * cos_p * sinm[m - 1] - sinm[m - 2];
      }
    }
    for (int l = 0; l <= L; ++l) {
      int base = l * l + l;
      for (int m = 1; m <= l; ++m) {
        Y[base + m] = P(l, m) * cosm[m];
        Y[base - m] = P(l, m) * sinm[m];
      }
      Y[base] = P(l, 0) * kInvSqrt2;
    }
    if (deriv) derivatives(sxy, cos_t, sin_p, cos_p);
  }
  // spherical_harmonics.cc:227-363 (gradient on the unit sphere, no 1/r)
  void derivatives(double sin_t, double cos_t, double sin_p, double cos_p) {
    const int LL = (L + 1) * (L + 1);
    double* dx = &dY[0];
    double* dy = &dY[LL];
    double* dz = &dY[2 * LL];
    dx[0] = dy[0] = dz[0] = 0.0;
    std::vector<double> lpd(L + 1), pdf(L + 1);
    for (int l = 1; l <= L; ++l) {
      const double* pl = &plm[l * (L + 1)];
      for (int k = 0; k < l; ++k) {
        double p2 = (k + 2 <= L + 1) ? P(l, k + 2) : 0.0;
        double f1 = (k + 1 <= l) ? pl[k + 1] : 0.0;
        lpd[k] = pl[k] * P(l, k) - f1 * p2;
        if (sin_t > 0.1)
          pdf[k] = double(k + 1) / sin_t * P(l, k + 1);
        else
          pdf[k] = -0.5 / cos_t * (pl[k] * P(l, k) + f1 * p2);
      }
      int base = l * l + l;
      double q = pl[0] * kInvSqrt2 * P(l, 1);
      dx[base] = cos_t * cos_p * q;
      dy[base] = cos_t * sin_p * q;
      dz[base] = -1.0 * sin_t * q;
      for (int k = 0; k < l; ++k) {
        int m = k + 1;
        dx[base + m] = sin_p * pdf[k] * sinm[m] +
                       -0.5 * cos_t * cos_p * cosm[m] * lpd[k];
        dy[base + m] = -1.0 * cos_p * pdf[k] * sinm[m] +
                       -0.5 * cos_t * sin_p * cosm[m] * lpd[k];
        dz[base + m] = 0.5 * sin_t * cosm[m] * lpd[k];
        dx[base - m] = -1.0 * sin_p * pdf[k] * cosm[m] +
                       -0.5 * cos_t * cos_p * sinm[m] * lpd[k];
        dy[base - m] = cos_p * pdf[k] * cosm[m] +
                       -0.5 * cos_t * sin_p * sinm[m] * lpd[k];
        dz[base - m] = 0.5 * sin_t * sinm[m] * lpd[k];
      }
    }
  }
};

/* ------------------------------------------------------------------------
 * Cutoff functions (representations/cutoff_functions.hh:69-283)
 * ---------------------------------------------------------------------- */
double sw_cos(double r, double rc, double sw) {
  if (r <= rc - sw) return 1.0;
  if (r > rc) return 0.0;
  double rs = kPi * (r - rc + sw) / sw;
  return 0.5 * (1. + std::cos(rs));
}
double dsw_cos(double r, double rc, double sw) {
  if (r <= rc - sw) return 0.0;
  if (r > rc) return 0.0;
  double rs = kPi * (r - rc + sw) / sw;
  return -0.5 * kPi / sw * std::sin(rs);
}
double cutoff_fc(const orc_hypers& h, double r) {
  double f = sw_cos(r, h.cutoff, h.smooth_width);
  if (h.cutoff_type == 0) return f;
  double u;
  if (std::fabs(h.rs_rate) <= kDblFtol)
    u = 1. / std::pow(r / h.rs_scale, h.rs_exponent);
  else if (h.rs_exponent == 0)
    u = 1.;
  else
    u = h.rs_rate / (h.rs_rate + std::pow(r / h.rs_scale, h.rs_exponent));
  return u * f;
}
double cutoff_dfc(const orc_hypers& h, double r) {
  double df = dsw_cos(r, h.cutoff, h.smooth_width);
  if (h.cutoff_type == 0) return df;
  double f = sw_cos(r, h.cutoff, h.smooth_width);
  double u, du;
  if (std::fabs(h.rs_rate) <= kDblFtol) {
    u = 1. / std::pow(r / h.rs_scale, h.rs_exponent);
    du = -h.rs_exponent / r / std::pow(r / h.rs_scale, h.rs_exponent);
  } else if (h.rs_exponent == 0) {
    u = 1.;
    du = 0.;
  } else {
    double ff = std::pow(r / h.rs_scale, h.rs_exponent);
    u = h.rs_rate / (h.rs_rate + ff);
    du = -h.rs_rate * h.rs_exponent * ff / r / pow_u(h.rs_rate + ff, 2);
  }
  return du * f + u * df;
}

/* ------------------------------------------------------------------------
 * Symmetric eigen-decomposition (cyclic Jacobi) for S^{-1/2}
 * ---------------------------------------------------------------------- */
void jacobi_eigen(int n, std::vector<double> A, std::vector<double>& w,
                  std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (std::fabs(theta) + std::sqrt(theta * theta + 1));
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
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

/* ------------------------------------------------------------------------
 * Uniform cubic spline of a vector function (math/interpolator.hh:166-186,
 * 501-538, 705-712, 976-1282; interpolator.cc:148-230)
 * ---------------------------------------------------------------------- */
struct VecSpline {
  double x1 = 0, x2 = 0, h = 0, h_sq_6 = 0, per_unit = 0;
  int n = 0, F = 0, search_size = 0;
  std::vector<double> grid, y, d2;  // y, d2: [n][F]
  static std::vector<double> linspaced(int num, double lo, double hi) {
    // Eigen 3.3 LinSpaced for floating point (flip when |hi| < |lo|)
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
  template <class Fn>
  void build(Fn fn, int F_, double x1_, double x2_, double accuracy) {
    F = F_;
    x1 = x1_;
    x2 = x2_;
    int dof = 5;
    const int max_grid = 100000;
    std::vector<double> tmp(F);
    while (true) {
      int npts = 2 << dof;
      grid = linspaced(npts, x1, x2);
      n = npts;
      y.assign(size_t(n) * F, 0.0);
      for (int i = 0; i < n; ++i) {
        fn(grid[i], tmp.data());
        std::copy(tmp.begin(), tmp.end(), y.begin() + size_t(i) * F);
      }
      init_from_grid();
      int tdof = std::max(dof, 5);
      int tn = (2 << tdof);
      double off = (x2 - x1) / (2 * (tn - 1));
      std::vector<double> tg = linspaced(tn - 1, x1 + off, x2 - off);
      double err = 0;
      std::vector<double> iv(F);
      for (double x : tg) {
        eval(x, iv.data());
        fn(x, tmp.data());
        for (int k = 0; k < F; ++k) err += std::fabs(iv[k] - tmp[k]);
      }
      err /= double(tg.size()) * F;
      bool ok = err < accuracy;
      if (ok || n >= max_grid) break;
      ++dof;
    }
  }
  // interpolator.cc:148-175 (natural boundary conditions)
  void init_from_grid() {
    h = grid[1] - grid[0];
    h_sq_6 = h * h / 6.0;
    std::vector<double> y2(size_t(n) * F, 0.0), u(size_t(n) * F, 0.0),
        p(F, 0.0);
    const double sig = 0.5;
    for (int i = 1; i < n - 1; ++i) {
      for (int k = 0; k < F; ++k) {
        p[k] = sig * y2[size_t(i - 1) * F + k] + 2.0;
        y2[size_t(i) * F + k] = (sig - 1.0) / p[k];
        double uu = 3. * (y[size_t(i + 1) * F + k] - 2 * y[size_t(i) * F + k] +
                          y[size_t(i - 1) * F + k]) /
                    (h * h);
        uu -= sig * u[size_t(i - 1) * F + k];
        uu /= p[k];
        u[size_t(i) * F + k] = uu;
      }
    }
    for (int k = 0; k < F; ++k) y2[size_t(n - 1) * F + k] = 0.0;
    for (int i = n - 2; i > 0; --i)
      for (int k = 0; k < F; ++k)
        y2[size_t(i) * F + k] =
            y2[size_t(i) * F + k] * y2[size_t(i + 1) * F + k] + u[size_t(i) * F + k];
    for (int k = 0; k < F; ++k) y2[k] = 0.0;
    d2.resize(y2.size());
    for (size_t k = 0; k < y2.size(); ++k) d2[k] = y2[k] * h_sq_6;
    // interpolator.cc:214-223 (UniformSearchMethod::initialize, 2 support pts)
    per_unit = double(n) / (grid[n - 1] - grid[0]);
    search_size = n - 2;
  }
  // interpolator.cc:225-230: cell left of the nearest grid point
  int search(double x) const {
    return std::max(0, std::min(search_size, int((x - x1) * per_unit) - 1));
  }
  // interpolator.cc:177-193
  void eval(double x, double* out) const {
    int klo = search(x), khi = klo + 1;
    double a = (grid[khi] - x) / h, b = 1 - a;
    const double* ylo = &y[size_t(klo) * F];
    const double* yhi = &y[size_t(khi) * F];
    const double* dlo = &d2[size_t(klo) * F];
    const double* dhi = &d2[size_t(khi) * F];
    for (int k = 0; k < F; ++k)
      out[k] = a * (ylo[k] + (a * a - 1) * dlo[k]) +
               b * (yhi[k] + (b * b - 1) * dhi[k]);
  }
  // interpolator.cc:195-212
  void eval_der(double x, double* out) const {
    int klo = search(x), khi = klo + 1;
    double a = (grid[khi] - x) / h, b = 1 - a;
    const double* ylo = &y[size_t(klo) * F];
    const double* yhi = &y[size_t(khi) * F];
    const double* dlo = &d2[size_t(klo) * F];
    const double* dhi = &d2[size_t(khi) * F];
    for (int k = 0; k < F; ++k)
      out[k] = (yhi[k] - ylo[k] - (3 * a * a - 1) * dlo[k] +
                (3 * b * b - 1) * dhi[k]) /
               h;
  }
};

/* ------------------------------------------------------------------------
 * GTO radial contribution (representations/calculator_spherical_expansion.hh
 * :322-728) with the None / Spline handlers (:945-1156)
 * ---------------------------------------------------------------------- */
struct Radial {
  orc_hypers hp{};
  int N = 0, L = 0;
  std::vector<double> sigma, fac_b, norm, nfac, ortho, O;  // O = diag(norm) S^-1/2
  std::vector<double> center;       // [N] (finalised when spline)
  double fac_a = 0;
  Hyp1f1Sph hyp{true, 1e-13, 200};  // expansion.hh:697
  VecSpline spline;
  bool use_spline = false;
  std::vector<double> R, dR, abln, dfal;  // R[n][l]

  explicit Radial(const orc_hypers& h) : hp(h), N(h.max_radial), L(h.max_angular) {
    // :621-636
    sigma.resize(N);
    fac_b.resize(N);
    norm.resize(N);
    nfac.resize(N);
    for (int n = 0; n < N; ++n) {
      sigma[n] = std::max(std::sqrt(double(n)), 1.0) * hp.cutoff / double(N);
      fac_b[n] = 0.5 * pow_i(sigma[n], -2);
      norm[n] = 0.25 * std::sqrt(2.0 / (std::tgamma(1.5 + n) *
                                        pow_u(sigma[n], 3 + 2 * n)));
      nfac[n] = std::tgamma(0.5 * (3.0 + n)) / std::tgamma(1.5);
    }
    // :643-680 overlap and S^{-1/2}
    std::vector<double> S(N * N);
    for (int a = 0; a < N; ++a)
      for (int b = 0; b < N; ++b) {
        double sa = sigma[a], sb = sigma[b];
        S[a * N + b] = std::pow(0.5 / pow_i(sa, 2) + 0.5 / pow_i(sb, 2),
                                -0.5 * (3.0 + a + b)) /
                       (pow_u(sa, a) * pow_u(sb, b)) *
                       std::tgamma(0.5 * (3.0 + a + b)) /
                       (std::pow(sa * sb, 1.5) *
                        std::sqrt(std::tgamma(1.5 + a) * std::tgamma(1.5 + b)));
      }
    std::vector<double> w, V;
    jacobi_eigen(N, S, w, V);
    ortho.assign(N * N, 0.0);
    for (int a = 0; a < N; ++a)
      for (int b = 0; b < N; ++b) {
        double s = 0;
        for (int k = 0; k < N; ++k) s += V[a * N + k] * (1.0 / std::sqrt(w[k])) * V[b * N + k];
        ortho[a * N + b] = s;
      }
    O.assign(N * N, 0.0);
    for (int a = 0; a < N; ++a)
      for (int b = 0; b < N; ++b) O[a * N + b] = norm[a] * ortho[a * N + b];
    hyp.precompute(N, L);
    fac_a = 0.5 * pow_i(hp.gaussian_sigma, -2);
    R.assign(N * (L + 1), 0.0);
    dR.assign(N * (L + 1), 0.0);
    abln.assign(N * (L + 1), 0.0);
    dfal.assign(L + 1, 0.0);
    // :478-501 centre contribution
    center.assign(N, 0.0);
    for (int n = 0; n < N; ++n) {
      double abn;
      if (n % 2 == 0)
        abn = std::sqrt(pow_i(fac_a + fac_b[n], -1 * (3 + n)));
      else
        abn = pow_i(fac_a + fac_b[n], -1 * (3 + n) / 2);
      center[n] = nfac[n] * abn;
    }
    use_spline = hp.optimization == 1;
    if (use_spline) {
      // :1132-1134 the centre term is finalised at set-up for the spline
      std::vector<double> cf(N, 0.0);
      for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) cf[a] += O[b * N + a] * center[b];
      center = cf;
      // :1108-1135 spline of O^T R(r) on [1e-5, rc]
      const int F = N * (L + 1);
      spline.build(
          [this, F](double x, double* out) {
            direct(x, false);
            std::vector<double> fin(N * (L + 1));
            finalize_neighbour(R.data(), fin.data());
            // column-major flattening (index l*N + n)
            for (int n = 0; n < N; ++n)
              for (int l = 0; l <= L; ++l) out[l * N + n] = fin[n * (L + 1) + l];
            (void)F;
          },
          F, kSplineX1, hp.cutoff, hp.spline_accuracy);
    }
  }
  // :518-557 (and :560-584 for the derivative)
  void direct(double r, bool der) {
    dfal[0] = 1.;
    double dfa = r * fac_a;
    for (int l = 1; l <= L; ++l) dfal[l] = dfal[l - 1] * dfa;
    for (int n = 0; n < N; ++n) {
      double ab = 1.0 / std::sqrt(fac_a + fac_b[n]);
      abln[n * (L + 1)] = pow_u(ab, 3 + n);
      for (int l = 1; l <= L; ++l) abln[n * (L + 1) + l] = abln[n * (L + 1) + l - 1] * ab;
    }
    hyp.calc(r, fac_a, fac_b.data(), der);
    for (int n = 0; n < N; ++n)
      for (int l = 0; l <= L; ++l) {
        int k = n * (L + 1) + l;
        R[k] = abln[k] * hyp.vals[k] * dfal[l];
      }
    if (der) {
      for (int n = 0; n < N; ++n)
        for (int l = 0; l <= L; ++l) {
          int k = n * (L + 1) + l;
          dR[k] = abln[k] * hyp.ders[k] * dfal[l] + R[k] * (double(l) / r);
        }
    }
  }
  // out[n][l] = sum_m O[m][n] in[m][l]  (O^T R, :598-604)
  void finalize_neighbour(const double* in, double* out) const {
    for (int n = 0; n < N; ++n)
      for (int l = 0; l <= L; ++l) {
        double s = 0;
        for (int m = 0; m < N; ++m) s += O[m * N + n] * in[m * (L + 1) + l];
        out[n * (L + 1) + l] = s;
      }
  }
  // neighbour contribution (None: raw R, finalised later; Spline: O^T R)
  void neighbour(double r, bool der) {
    if (!use_spline) {
      direct(r, der);
      return;
    }
    std::vector<double> tmp(N * (L + 1));
    spline.eval(r, tmp.data());
    for (int n = 0; n < N; ++n)
      for (int l = 0; l <= L; ++l) R[n * (L + 1) + l] = tmp[l * N + n];
    if (der) {
      spline.eval_der(r, tmp.data());
      for (int n = 0; n < N; ++n)
        for (int l = 0; l <= L; ++l) dR[n * (L + 1) + l] = tmp[l * N + n];
    }
  }
};

/* ------------------------------------------------------------------------
 * Neighbour list (structure_managers/adaptor_neighbour_list.hh:904-1118,
 * adaptor_center_contribution.hh:334-403, adaptor_strict.hh:373-455)
 * ---------------------------------------------------------------------- */
struct Pairs {
  std::vector<int64_t> off;   // [n_centres+1]
  std::vector<int> j;         // original atom index
  std::vector<double> vec;    // [n][3]
  std::vector<double> dist;
};

void inverse3(const double* C, double* Ci) {
  // cofactor inverse of a 3x3 matrix (row-major C[d*3+k])
  double a = C[0], b = C[1], c = C[2], d = C[3], e = C[4], f = C[5], g = C[6],
         h = C[7], i = C[8];
  double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
  double det = a * A + b * B + c * Cc;
  double id = 1.0 / det;
  Ci[0] = A * id;
  Ci[3] = B * id;
  Ci[6] = Cc * id;
  Ci[1] = -(b * i - c * h) * id;
  Ci[4] = (a * i - c * g) * id;
  Ci[7] = -(a * h - b * g) * id;
  Ci[2] = (b * f - c * e) * id;
  Ci[5] = -(a * f - c * d) * id;
  Ci[8] = (a * e - b * d) * id;
}

void check_structure(int n_atoms, const double* pos, const int* types,
                     const double* cell) {
  // atomic_structure.hh:249-255 zero cell; structure_manager_centers.cc:67-99
  bool all_zero = true;
  for (int k = 0; k < 9; ++k)
    if (std::fabs(cell[k]) >= 1e-10) all_zero = false;
  if (all_zero)
    throw std::runtime_error("The unit cell should not be filled with zeros");
  double Ci[9];
  inverse3(cell, Ci);
  for (int a = 0; a < n_atoms; ++a) {
    for (int d = 0; d < 3; ++d) {
      double s = Ci[d * 3 + 0] * pos[a * 3 + 0] + Ci[d * 3 + 1] * pos[a * 3 + 1] +
                 Ci[d * 3 + 2] * pos[a * 3 + 2];
      if (s < -1e-10 || s > 1. + 1e-10)
        throw std::runtime_error(
            "Some of the atoms in the structure are not inside the unit cell");
    }
    if (types && (types[a] < 0 || types[a] >= 130))
      throw std::runtime_error("atom types should be in the range [0, 130]");
  }
}

void build_pairs(int n_atoms, const double* pos, const double* cell,
                 const int* pbc, double rc, Pairs& out) {
  // mesh and ghost bounds (:935-968)
  double mesh_min[3], ghost_min[3], ghost_max[3];
  int nboxes[3];
  for (int i = 0; i < 3; ++i) {
    double mn = 0.0, mx = 0.0;
    for (int col = 0; col < 3; ++col) {
      double v = cell[i * 3 + col];
      mn += v * static_cast<double>(v < 0.);
      mx += v * static_cast<double>(v > 0.);
    }
    mesh_min[i] = mn - 2. * rc;
    double lmesh = mx - mn + 4. * rc;
    double nb = std::ceil(lmesh / rc);
    nboxes[i] = static_cast<int>(nb);
    ghost_min[i] = mn - rc;
    double lghost = mx - mn + 2 * rc;
    double ngb = std::ceil(lghost / rc);
    ghost_max[i] = ngb * rc + ghost_min[i];
  }
  double maxlen = std::max({ghost_max[0] - ghost_min[0], ghost_max[1] - ghost_min[1],
                            ghost_max[2] - ghost_min[2]});
  double tol = maxlen * 1e-8;
  // image multipliers from the mesh corners (:974-1011)
  double lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    lo[i] = ghost_min[i] - tol;
    hi[i] = ghost_max[i] + tol;
  }
  double Ci[9];
  inverse3(cell, Ci);
  double xmin[3] = {INFINITY, INFINITY, INFINITY},
         xmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = ((c >> i) & 1) ? hi[i] : lo[i];
    for (int d = 0; d < 3; ++d) {
      double m = Ci[d * 3 + 0] * x[0] + Ci[d * 3 + 1] * x[1] + Ci[d * 3 + 2] * x[2];
      xmin[d] = std::min(xmin[d], m);
      xmax[d] = std::max(xmax[d], m);
    }
  }
  int pmin[3], rep[3];
  for (int i = 0; i < 3; ++i) {
    int mmin = static_cast<int>(std::floor(xmin[i]));
    int mmax = static_cast<int>(std::ceil(xmax[i])) - 1;
    if (pbc[i]) {
      pmin[i] = mmin;
      rep[i] = mmax - mmin + 1;
    } else {
      pmin[i] = 0;
      rep[i] = 1;
    }
  }
  // extended atom list: originals then ghosts (:1021-1072)
  std::vector<double> xpos(pos, pos + 3 * n_atoms);
  std::vector<int> orig(n_atoms);
  for (int a = 0; a < n_atoms; ++a) orig[a] = a;
  int ntot = rep[0] * rep[1] * rep[2];
  for (int a = 0; a < n_atoms; ++a) {
    for (int idx = 0; idx < ntot; ++idx) {
      int p[3];
      int f = 1;
      for (int i = 2; i >= 0; --i) {
        p[i] = idx / f % rep[i] + pmin[i];
        f *= rep[i];
      }
      if (p[0] == 0 && p[1] == 0 && p[2] == 0) continue;
      double g[3];
      bool inside = true;
      for (int d = 0; d < 3; ++d) {
        double t = cell[d * 3 + 0] * double(p[0]);
        t = t + cell[d * 3 + 1] * double(p[1]);
        t = t + cell[d * 3 + 2] * double(p[2]);
        g[d] = pos[a * 3 + d] + t;
        // position_in_bounds (:357-373)
        if (!((g[d] - ghost_min[d]) > tol && (g[d] - ghost_max[d]) < -tol))
          inside = false;
      }
      if (inside) {
        xpos.insert(xpos.end(), g, g + 3);
        orig.push_back(a);
      }
    }
  }
  int next = static_cast<int>(orig.size());
  // linked cells (:1076-1083, :327-336, :395-422)
  auto box_of = [&](const double* x, int* b) {
    for (int d = 0; d < 3; ++d) {
      double v = x[d] - mesh_min[d];
      b[d] = std::max(1, static_cast<int>(std::floor(v / rc)));
    }
  };
  std::vector<std::vector<int>> boxes(size_t(nboxes[0]) * nboxes[1] * nboxes[2]);
  auto lin = [&](const int* b) {
    return (size_t(b[0]) * nboxes[1] + b[1]) * nboxes[2] + b[2];
  };
  for (int t = 0; t < next; ++t) {
    int b[3];
    box_of(&xpos[3 * t], b);
    for (int d = 0; d < 3; ++d)
      if (b[d] >= nboxes[d] - 1 || b[d] <= 0)
        throw std::runtime_error(
            "this atom does not fall in one of the linked cell bin");
    boxes[lin(b)].push_back(t);
  }
  // candidates per centre, self pair first, strict filter
  out.off.assign(n_atoms + 1, 0);
  out.j.clear();
  out.vec.clear();
  out.dist.clear();
  double rc2 = rc * rc;
  auto add_pair = [&](int i, int t) {
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = xpos[3 * t + d] - pos[3 * i + d];
    double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (d2 <= rc2) {
      out.j.push_back(orig[t]);
      out.vec.insert(out.vec.end(), v, v + 3);
      out.dist.push_back(std::sqrt(d2));
    }
  };
  for (int i = 0; i < n_atoms; ++i) {
    add_pair(i, i);
    int b[3];
    box_of(&pos[3 * i], b);
    for (int s = 0; s < 27; ++s) {
      int c[3];
      int f = 1;
      for (int d = 2; d >= 0; --d) {
        c[d] = s / f % 3 + b[d] - 1;
        f *= 3;
      }
      for (int t : boxes[lin(c)])
        if (t != i) add_pair(i, t);
    }
    out.off[i + 1] = static_cast<int64_t>(out.j.size());
  }
}

/* ------------------------------------------------------------------------
 * Calculator: expansion + power spectrum + GAP
 * ---------------------------------------------------------------------- */
struct Calc {
  orc_hypers hp;
  Radial rad;
  Ylm ylm;
  int N, L, L1, LL;
  explicit Calc(const orc_hypers& h)
      : hp(h), rad(h), N(h.max_radial), L(h.max_angular), L1(h.max_angular + 1),
        LL((h.max_angular + 1) * (h.max_angular + 1)) {
    ylm.precompute(L);
  }
};

struct StructView {
  int n;
  const double* pos;
  const int* types;
  const double* cell;
  const int* pbc;
};

StructView view(const orc_batch* b, int s) {
  int64_t a0 = b->atom_offsets[s], a1 = b->atom_offsets[s + 1];
  return StructView{static_cast<int>(a1 - a0), b->positions + 3 * a0,
                    b->types + a0, b->cells + 9 * s, b->pbc + 3 * s};
}

int species_index(const orc_batch* b, int t) {
  for (int k = 0; k < b->n_species; ++k)
    if (b->species[k] == t) return k;
  throw std::runtime_error("atom type not in the species list");
}

// Per-centre expansion (calculator_spherical_expansion.hh:2234-2410).
// coeffs: [S][N][LL] of this centre; grads (optional): [n_pairs_i][3][S][N][LL]
void expand_centre(Calc& C, const orc_batch* b, const StructView& sv, int i,
                   const Pairs& P, double* coeffs, uint64_t* kmask,
                   double* grads) {
  const int S = b->n_species, N = C.N, L = C.L, L1 = C.L1, LL = C.LL;
  const size_t blk = size_t(N) * LL;
  const bool G = grads != nullptr;
  std::fill(coeffs, coeffs + S * blk, 0.0);
  int64_t p0 = P.off[i], p1 = P.off[i + 1];
  size_t np = size_t(p1 - p0);
  if (G) std::fill(grads, grads + np * 3 * S * blk, 0.0);
  int si = species_index(b, sv.types[i]);
  uint64_t mask = 1ull << si;
  const double inv_s4pi = 1.0 / std::sqrt(4.0 * kPi);
  for (int n = 0; n < N; ++n) coeffs[si * blk + n * LL] += C.rad.center[n] * inv_s4pi;
  std::vector<double> cij(blk), pg(N * L1);
  for (int64_t p = p0 + 1; p < p1; ++p) {  // pairs() skips the self pair
    int j = P.j[p];
    int sj = species_index(b, sv.types[j]);
    mask |= 1ull << sj;
    double dist = P.dist[p];
    double dir[3];
    for (int d = 0; d < 3; ++d) dir[d] = P.vec[3 * p + d] / dist;
    C.ylm.calc(dir, G);
    C.rad.neighbour(dist, G);
    double fc = cutoff_fc(C.hp, dist);
    double* cb = coeffs + sj * blk;
    for (int n = 0; n < N; ++n)
      for (int l = 0; l <= L; ++l) {
        double r = C.rad.R[n * L1 + l];
        for (int m = -l; m <= l; ++m) {
          int lm = l * l + l + m;
          double v = r * C.ylm.Y[lm] * fc;
          cij[n * LL + lm] = v;
          cb[n * LL + lm] += v;
        }
      }
    if (G) {
      // :2306-2366
      double dfc = cutoff_dfc(C.hp, dist);
      for (int k = 0; k < N * L1; ++k) pg[k] = C.rad.dR[k] * fc + C.rad.R[k] * dfc;
      bool not_image = j != i;
      double* gself = grads;                       // slot 0 = self pair
      double* gpair = grads + size_t(p - p0) * 3 * S * blk;
      for (int c = 0; c < 3; ++c) {
        const double* dYc = &C.ylm.dY[c * LL];
        for (int n = 0; n < N; ++n)
          for (int l = 0; l <= L; ++l) {
            double a1 = pg[n * L1 + l], r = C.rad.R[n * L1 + l];
            for (int m = -l; m <= l; ++m) {
              int lm = l * l + l + m;
              double v = a1 * C.ylm.Y[lm] * dir[c] + r * dYc[lm] * fc / dist;
              size_t o = (size_t(c) * S + sj) * blk + n * LL + lm;
              if (not_image) gself[o] -= v;
              gpair[o] = v;
            }
          }
      }
    }
  }
  if (kmask) *kmask = mask;
  if (!C.rad.use_spline) {
    // finalize: c <- O^T c per key block (:605-618, property_block_sparse
    // lhs_dot); gradients per Cartesian block (lhs_dot_der)
    auto fin = [&](double* blkp) {
      std::vector<double> tmp(blk);
      for (int n = 0; n < N; ++n)
        for (int lm = 0; lm < LL; ++lm) {
          double s = 0;
          for (int m = 0; m < N; ++m) s += C.rad.O[m * N + n] * blkp[m * LL + lm];
          tmp[n * LL + lm] = s;
        }
      std::copy(tmp.begin(), tmp.end(), blkp);
    };
    for (int s = 0; s < S; ++s) fin(coeffs + s * blk);
    if (G)
      for (size_t q = 0; q < np; ++q)
        for (int c = 0; c < 3; ++c)
          for (int s = 0; s < S; ++s) fin(grads + ((q * 3 + c) * S + s) * blk);
  }
}

int n_keys(int S) { return S * (S + 1) / 2; }

// Power spectrum of one centre (calculator_spherical_invariants.hh:747-794)
// ps: [K][N*N][L1]; returns the norm
double ps_centre(const Calc& C, int S, const double* c, double* ps, bool normalize) {
  const int N = C.N, L1 = C.L1, LL = C.LL;
  const size_t blk = size_t(N) * LL;
  const size_t pblk = size_t(N) * N * L1;
  int key = 0;
  for (int a = 0; a < S; ++a)
    for (int bsp = a; bsp < S; ++bsp, ++key) {
      double* pk = ps + key * pblk;
      const double* ca = c + a * blk;
      const double* cb = c + bsp * blk;
      for (int n1 = 0; n1 < N; ++n1)
        for (int n2 = 0; n2 < N; ++n2)
          for (int l = 0; l < L1; ++l) {
            double s = 0;
            for (int m = -l; m <= l; ++m) {
              int lm = l * l + l + m;
              s += ca[n1 * LL + lm] * cb[n2 * LL + lm];
            }
            double v = s * (1.0 / std::sqrt(2 * l + 1));
            if (a < bsp) v *= kSqrt2;
            pk[(n1 * N + n2) * L1 + l] = v;
          }
    }
  const size_t D = size_t(n_keys(S)) * pblk;
  double nrm = 0;
  for (size_t k = 0; k < D; ++k) nrm += ps[k] * ps[k];
  nrm = std::sqrt(nrm);
  if (normalize && std::fabs(nrm) > 0.)
    for (size_t k = 0; k < D; ++k) ps[k] /= nrm;
  return nrm;
}

// Power-spectrum gradient for one pair slot (:796-921), un-normalised;
// dc: [3][S][N][LL] gradient of c for the slot, c: [S][N][LL]; out [3][D]
void ps_grad_slot(const Calc& C, int S, const double* c, const double* dc,
                  double* out) {
  const int N = C.N, L1 = C.L1, LL = C.LL;
  const size_t blk = size_t(N) * LL, pblk = size_t(N) * N * L1;
  const size_t D = size_t(n_keys(S)) * pblk;
  for (int x = 0; x < 3; ++x) {
    const double* g = dc + size_t(x) * S * blk;
    double* o = out + x * D;
    int key = 0;
    for (int a = 0; a < S; ++a)
      for (int bsp = a; bsp < S; ++bsp, ++key) {
        const double* ca = c + a * blk;
        const double* cb = c + bsp * blk;
        const double* ga = g + a * blk;
        const double* gb = g + bsp * blk;
        for (int n1 = 0; n1 < N; ++n1)
          for (int n2 = 0; n2 < N; ++n2)
            for (int l = 0; l < L1; ++l) {
              double s1 = 0, s2 = 0;
              for (int m = -l; m <= l; ++m) {
                int lm = l * l + l + m;
                s1 += ga[n1 * LL + lm] * cb[n2 * LL + lm];
                s2 += gb[n2 * LL + lm] * ca[n1 * LL + lm];
              }
              double lf = 1.0 / std::sqrt(2 * l + 1);
              double v = s1 * lf + s2 * lf;
              if (a < bsp) v *= kSqrt2;
              o[key * pblk + (n1 * N + n2) * L1 + l] = v;
            }
      }
  }
}

// (:588-652) gradient of the normalised vector
void ps_grad_normalize(size_t D, const double* pn, double inv_norm, double* g) {
  for (int x = 0; x < 3; ++x) {
    double* gx = g + x * D;
    double dot = 0;
    for (size_t k = 0; k < D; ++k) {
      gx[k] *= inv_norm;
      dot += gx[k] * pn[k];
    }
    for (size_t k = 0; k < D; ++k) gx[k] -= dot * pn[k];
  }
}

double pow_zeta(double v, int zeta) {
  // models/kernels.hh:48-63
  if (zeta == 0) return pow_u(v, 0);
  if (zeta == 1) return v;
  if (zeta == 2) return v * v;
  if (zeta == 3) return v * v * v;
  return pow_u(v, size_t(zeta));
}

double volume(const double* C) {
  double a = C[0], b = C[1], c = C[2], d = C[3], e = C[4], f = C[5], g = C[6],
         h = C[7], i = C[8];
  return std::fabs(a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g));
}

template <class F>
int guarded(F&& f) {
  try {
    g_err.clear();
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Per-centre sweep shared by the representation entry points.  For every
// centre it calls fn(s, i_global, i_local, centre_state) with the expansion
// (and its per-slot gradients when want_grad) of that centre.
struct CentreState {
  const Pairs* P;
  std::vector<double> c, dc, ps, dps;
  double norm = 0;
  uint64_t mask = 0;
};

template <class Fn>
void sweep(Calc& C, const orc_batch* b, bool want_ps, bool want_grad,
           int64_t max_centres, Fn fn, int64_t first_centre = 0) {
  const int S = b->n_species;
  const size_t blk = size_t(C.N) * C.LL;
  const size_t D = size_t(n_keys(S)) * C.N * C.N * C.L1;
  if (S > 64) throw std::runtime_error("at most 64 species");
  int64_t done = 0, seen = 0;
  for (int s = 0; s < b->n_struct; ++s) {
    StructView sv = view(b, s);
    check_structure(sv.n, sv.pos, sv.types, sv.cell);
    Pairs P;
    build_pairs(sv.n, sv.pos, sv.cell, sv.pbc, C.hp.cutoff, P);
    CentreState st;
    st.P = &P;
    st.c.resize(S * blk);
    if (want_ps) st.ps.resize(D);
    for (int i = 0; i < sv.n; ++i) {
      if (max_centres >= 0 && done >= max_centres) return;
      if (seen++ < first_centre) continue;
      size_t np = size_t(P.off[i + 1] - P.off[i]);
      if (want_grad) st.dc.resize(np * 3 * S * blk);
      expand_centre(C, b, sv, i, P, st.c.data(), &st.mask,
                    want_grad ? st.dc.data() : nullptr);
      if (want_ps) {
        st.norm = ps_centre(C, S, st.c.data(), st.ps.data(), C.hp.normalize != 0);
        if (want_grad) {
          st.dps.resize(np * 3 * D);
          for (size_t q = 0; q < np; ++q)
            ps_grad_slot(C, S, st.c.data(), st.dc.data() + q * 3 * S * blk,
                         st.dps.data() + q * 3 * D);
          if (C.hp.normalize && std::fabs(st.norm) > 0.)
            for (size_t q = 0; q < np; ++q)
              ps_grad_normalize(D, st.ps.data(), 1.0 / st.norm,
                                st.dps.data() + q * 3 * D);
        }
      }
      fn(s, b->atom_offsets[s] + i, i, sv, st);
      ++done;
    }
  }
}

}  // namespace

/* ======================================================================== */
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

double orc_hyp1f1(double a, double b, double z, int derivative, int mmax,
                  double tol) {
  Hyp1f1 h(a, b, size_t(mmax), tol);
  return h.f(z, derivative != 0);
}

double orc_hyp1f1_switch_point(double a, double b) {
  Hyp1f1 h(a, b, 200, 1e-13);
  return h.z_asympt;
}

int orc_hyp1f1_sph(int N, int L, double r, double alpha, const double* fac_b,
                   int recursion, double tol, double* values, double* derivatives) {
  return guarded([&] {
    Hyp1f1Sph h(recursion != 0, tol, 200);
    h.precompute(N, L);
    h.calc(r, alpha, fac_b, true);
    std::copy(h.vals.begin(), h.vals.end(), values);
    if (derivatives) std::copy(h.ders.begin(), h.ders.end(), derivatives);
  });
}

int orc_ylm(int L, const double* direction, double* harmonics,
            double* derivatives, double* alp) {
  return guarded([&] {
    Ylm y;
    y.precompute(L);
    y.calc(direction, derivatives != nullptr);
    std::copy(y.Y.begin(), y.Y.end(), harmonics);
    if (derivatives) std::copy(y.dY.begin(), y.dY.end(), derivatives);
    if (alp) std::copy(y.alp.begin(), y.alp.end(), alp);
  });
}

double orc_cutoff_fc(const orc_hypers* h, double r) { return cutoff_fc(*h, r); }
double orc_cutoff_dfc(const orc_hypers* h, double r) { return cutoff_dfc(*h, r); }

void* orc_calc_create(const orc_hypers* h) {
  try {
    g_err.clear();
    if (h->max_radial < 1 || h->max_angular < 0)
      throw std::runtime_error("invalid max_radial / max_angular");
    return new Calc(*h);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_calc_destroy(void* calc) { delete static_cast<Calc*>(calc); }

int orc_calc_radial_info(void* calc, double* fac_b, double* ortho_norm,
                         double* center, int* grid, double* x1, double* x2) {
  Calc& C = *static_cast<Calc*>(calc);
  if (fac_b) std::copy(C.rad.fac_b.begin(), C.rad.fac_b.end(), fac_b);
  if (ortho_norm) std::copy(C.rad.O.begin(), C.rad.O.end(), ortho_norm);
  if (center) std::copy(C.rad.center.begin(), C.rad.center.end(), center);
  if (grid) *grid = C.rad.use_spline ? C.rad.spline.n : 0;
  if (x1) *x1 = C.rad.spline.x1;
  if (x2) *x2 = C.rad.spline.x2;
  return 0;
}

int orc_calc_radial(void* calc, double r, double* R, double* dR) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    C.rad.neighbour(r, dR != nullptr);
    if (R) std::copy(C.rad.R.begin(), C.rad.R.end(), R);
    if (dR) std::copy(C.rad.dR.begin(), C.rad.dR.end(), dR);
  });
}

int orc_calc_spline_tables(void* calc, double* y, double* d2) {
  Calc& C = *static_cast<Calc*>(calc);
  if (!C.rad.use_spline) return -1;
  if (y) std::copy(C.rad.spline.y.begin(), C.rad.spline.y.end(), y);
  if (d2) std::copy(C.rad.spline.d2.begin(), C.rad.spline.d2.end(), d2);
  return 0;
}

int64_t orc_nl_count(int n_atoms, const double* pos, const double* cell,
                     const int* pbc, double cutoff) {
  int64_t n = -1;
  guarded([&] {
    check_structure(n_atoms, pos, nullptr, cell);
    Pairs P;
    build_pairs(n_atoms, pos, cell, pbc, cutoff, P);
    n = static_cast<int64_t>(P.j.size());
  });
  return n;
}

int orc_nl_fill(int n_atoms, const double* pos, const double* cell,
                const int* pbc, double cutoff, int64_t* off, int* pj,
                double* pvec, double* pdist) {
  return guarded([&] {
    check_structure(n_atoms, pos, nullptr, cell);
    Pairs P;
    build_pairs(n_atoms, pos, cell, pbc, cutoff, P);
    std::copy(P.off.begin(), P.off.end(), off);
    std::copy(P.j.begin(), P.j.end(), pj);
    if (pvec) std::copy(P.vec.begin(), P.vec.end(), pvec);
    if (pdist) std::copy(P.dist.begin(), P.dist.end(), pdist);
  });
}

int64_t orc_batch_pairs(void* calc, const orc_batch* b) {
  int64_t tot = -1;
  guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    int64_t t = 0;
    for (int s = 0; s < b->n_struct; ++s) {
      StructView sv = view(b, s);
      check_structure(sv.n, sv.pos, sv.types, sv.cell);
      Pairs P;
      build_pairs(sv.n, sv.pos, sv.cell, sv.pbc, C.hp.cutoff, P);
      t += static_cast<int64_t>(P.j.size());
    }
    tot = t;
  });
  return tot;
}

int orc_expansion(void* calc, const orc_batch* b, double* coeffs,
                  uint64_t* key_mask) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    const size_t per = size_t(b->n_species) * C.N * C.LL;
    sweep(C, b, false, false, -1,
          [&](int, int64_t ig, int, const StructView&, CentreState& st) {
            if (coeffs) std::copy(st.c.begin(), st.c.end(), coeffs + ig * per);
            if (key_mask) key_mask[ig] = st.mask;
          });
  });
}

int orc_expansion_gradients(void* calc, const orc_batch* b, double* grads) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    size_t off = 0;
    sweep(C, b, false, true, -1,
          [&](int, int64_t, int, const StructView&, CentreState& st) {
            std::copy(st.dc.begin(), st.dc.end(), grads + off);
            off += st.dc.size();
          });
  });
}

int orc_power_spectrum(void* calc, const orc_batch* b, double* ps,
                       double* inv_norms, int64_t max_centres) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    const size_t D = size_t(n_keys(b->n_species)) * C.N * C.N * C.L1;
    sweep(C, b, true, false, max_centres,
          [&](int, int64_t ig, int, const StructView&, CentreState& st) {
            if (ps) std::copy(st.ps.begin(), st.ps.end(), ps + ig * D);
            if (inv_norms)
              inv_norms[ig] = (C.hp.normalize && st.norm > 0) ? 1.0 / st.norm : 1.0;
          });
  });
}

int orc_power_spectrum_gradients(void* calc, const orc_batch* b, double* grads,
                                 int* info) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    size_t off = 0, row = 0;
    sweep(C, b, true, true, -1,
          [&](int s, int64_t, int i, const StructView& sv, CentreState& st) {
            std::copy(st.dps.begin(), st.dps.end(), grads + off);
            off += st.dps.size();
            for (int64_t p = st.P->off[i]; p < st.P->off[i + 1]; ++p, ++row) {
              if (!info) continue;
              int j = st.P->j[p];
              info[row * 5 + 0] = s;
              info[row * 5 + 1] = i;
              info[row * 5 + 2] = j;
              info[row * 5 + 3] = sv.types[i];
              info[row * 5 + 4] = sv.types[j];
            }
          });
  });
}

int orc_gap_knm(void* calc, const orc_batch* b, const orc_sparse* sp, int target,
                double* knm) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    const size_t D = size_t(n_keys(b->n_species)) * C.N * C.N * C.L1;
    const int M = sp->n_sparse;
    size_t rows = target == 0 ? size_t(b->n_struct) : size_t(b->atom_offsets[b->n_struct]);
    std::fill(knm, knm + rows * M, 0.0);
    sweep(C, b, true, false, -1,
          [&](int s, int64_t ig, int i, const StructView& sv, CentreState& st) {
            double* row = knm + (target == 0 ? size_t(s) : size_t(ig)) * M;
            for (int m = 0; m < M; ++m) {
              if (sp->sp_species[m] != sv.types[i]) continue;
              const double* t = sp->values + size_t(m) * D;
              double d = 0;
              for (size_t k = 0; k < D; ++k) d += st.ps[k] * t[k];
              row[m] += pow_zeta(d, sp->zeta);
            }
          });
  });
}

int orc_gap_kmm(const orc_sparse* sp, int D, double* kmm) {
  return guarded([&] {
    const int M = sp->n_sparse;
    for (int a = 0; a < M; ++a)
      for (int c = 0; c < M; ++c) {
        double v = 0;
        if (sp->sp_species[a] == sp->sp_species[c]) {
          double d = 0;
          for (int k = 0; k < D; ++k)
            d += sp->values[size_t(a) * D + k] * sp->values[size_t(c) * D + k];
          v = pow_zeta(d, sp->zeta);
        }
        kmm[size_t(a) * M + c] = v;
      }
  });
}

// models/sparse_kernels.hh:258-499
int orc_gap_knm_derivative(void* calc, const orc_batch* b, const orc_sparse* sp,
                           int neg_stress, double* out) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    const size_t D = size_t(n_keys(b->n_species)) * C.N * C.N * C.L1;
    const int M = sp->n_sparse;
    const int64_t n_atoms = b->atom_offsets[b->n_struct];
    const size_t rows = size_t(3 * n_atoms) + (neg_stress ? size_t(6 * b->n_struct) : 0);
    std::fill(out, out + rows * M, 0.0);
    const int zeta = sp->zeta;
    std::vector<double> kz(M), tx(3);
    sweep(C, b, true, true, -1,
          [&](int s, int64_t ig, int i, const StructView& sv, CentreState& st) {
            const int64_t a0 = b->atom_offsets[s];
            for (int m = 0; m < M; ++m) {
              kz[m] = 0;
              if (sp->sp_species[m] != sv.types[i]) continue;
              const double* t = sp->values + size_t(m) * D;
              double d = 0;
              for (size_t k = 0; k < D; ++k) d += st.ps[k] * t[k];
              kz[m] = zeta > 1 ? zeta * pow_zeta(d, zeta - 1) : 1.0;
            }
            const Pairs& P = *st.P;
            int64_t p0 = P.off[i];
            for (int64_t p = p0; p < P.off[i + 1]; ++p) {
              const double* g = st.dps.data() + size_t(p - p0) * 3 * D;
              int j = P.j[p];
              for (int m = 0; m < M; ++m) {
                if (sp->sp_species[m] != sv.types[i]) continue;
                const double* t = sp->values + size_t(m) * D;
                for (int x = 0; x < 3; ++x) {
                  double d = 0;
                  for (size_t k = 0; k < D; ++k) d += g[x * D + k] * t[k];
                  tx[x] = d * kz[m];
                  out[size_t(3 * (a0 + j) + x) * M + m] += tx[x];
                }
                if (neg_stress) {
                  // r_ji = r_i - r_j(image) = -pair vector
                  double rji[3] = {-P.vec[3 * p], -P.vec[3 * p + 1], -P.vec[3 * p + 2]};
                  size_t r0 = size_t(3 * n_atoms) + size_t(6 * s);
                  const int vo[3][2] = {{4, 2}, {5, 0}, {3, 1}};
                  for (int x = 0; x < 3; ++x) {
                    out[(r0 + x) * M + m] += rji[x] * tx[x];
                    out[(r0 + vo[x][0]) * M + m] += rji[vo[x][1]] * tx[x];
                  }
                }
              }
            }
            (void)ig;
          });
    if (neg_stress)
      for (int s = 0; s < b->n_struct; ++s) {
        double vol = volume(b->cells + 9 * s);
        size_t r0 = size_t(3 * n_atoms) + size_t(6 * s);
        for (int r = 0; r < 6; ++r)
          for (int m = 0; m < M; ++m) out[(r0 + r) * M + m] /= vol;
      }
  });
}

// models/sparse_kernel_predict.hh:70-377
int orc_gap_predict(void* calc, const orc_batch* b, const orc_sparse* sp,
                    const double* w, double* energy, double* gradient,
                    double* neg_stress, int64_t max_centres, int64_t first_centre) {
  return guarded([&] {
    Calc& C = *static_cast<Calc*>(calc);
    const size_t D = size_t(n_keys(b->n_species)) * C.N * C.N * C.L1;
    const int M = sp->n_sparse;
    const int64_t n_atoms = b->atom_offsets[b->n_struct];
    const int zeta = sp->zeta;
    if (energy) std::fill(energy, energy + b->n_struct, 0.0);
    if (gradient) std::fill(gradient, gradient + 3 * n_atoms, 0.0);
    if (neg_stress) std::fill(neg_stress, neg_stress + 6 * b->n_struct, 0.0);
    const bool want_grad = gradient || neg_stress;
    std::vector<double> wt(M), tt(D);
    sweep(C, b, true, want_grad, max_centres,
          [&](int s, int64_t, int i, const StructView& sv, CentreState& st) {
            const int64_t a0 = b->atom_offsets[s];
            std::fill(tt.begin(), tt.end(), 0.0);
            for (int m = 0; m < M; ++m) {
              if (sp->sp_species[m] != sv.types[i]) continue;
              const double* t = sp->values + size_t(m) * D;
              double d = 0;
              for (size_t k = 0; k < D; ++k) d += st.ps[k] * t[k];
              if (energy) energy[s] += pow_zeta(d, zeta) * w[m];
              if (!want_grad) continue;
              // weights scaled by zeta (X.T)^(zeta-1), then T~ = sum w~ T
              double ws = w[m] * (zeta * pow_zeta(d, zeta - 1));
              for (size_t k = 0; k < D; ++k) tt[k] += ws * t[k];
            }
            if (!want_grad) return;
            const Pairs& P = *st.P;
            int64_t p0 = P.off[i];
            for (int64_t p = p0; p < P.off[i + 1]; ++p) {
              const double* g = st.dps.data() + size_t(p - p0) * 3 * D;
              double f[3];
              for (int x = 0; x < 3; ++x) {
                double d = 0;
                for (size_t k = 0; k < D; ++k) d += g[x * D + k] * tt[k];
                f[x] = d;
              }
              int j = P.j[p];
              if (gradient)
                for (int x = 0; x < 3; ++x) gradient[3 * (a0 + j) + x] += f[x];
              if (neg_stress) {
                double rji[3] = {-P.vec[3 * p], -P.vec[3 * p + 1], -P.vec[3 * p + 2]};
                const int vo[3][2] = {{4, 2}, {5, 0}, {3, 1}};
                for (int x = 0; x < 3; ++x) {
                  neg_stress[6 * s + x] += rji[x] * f[x];
                  neg_stress[6 * s + vo[x][0]] += rji[vo[x][1]] * f[x];
                }
              }
            }
          },
          first_centre);
    if (neg_stress)
      for (int s = 0; s < b->n_struct; ++s) {
        double vol = volume(b->cells + 9 * s);
        for (int r = 0; r < 6; ++r) neg_stress[6 * s + r] /= vol;
      }
  });
}

}  // extern "C"
