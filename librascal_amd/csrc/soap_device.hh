// Device-side building blocks of the SOAP expansion: cutoff functions,
// GTO radial integrals (direct Hyp1f1 in FP64 or cubic-spline lookup), and
// real spherical harmonics with their gradients.
//
// The formulas restate representations/cutoff_functions.hh:69-283,
// math/hyp1f1.cc:41-491, representations/calculator_spherical_expansion.hh
// :518-584 and math/spherical_harmonics.cc:97-363.
#pragma once

#include <hip/hip_runtime.h>

#include "radial.hh"

namespace rgpu {

template <class T>
__device__ __forceinline__ T dev_cos(T x);
template <>
__device__ __forceinline__ double dev_cos<double>(double x) { return cos(x); }
template <>
__device__ __forceinline__ float dev_cos<float>(float x) { return cosf(x); }
template <class T>
__device__ __forceinline__ T dev_sin(T x);
template <>
__device__ __forceinline__ double dev_sin<double>(double x) { return sin(x); }
template <>
__device__ __forceinline__ float dev_sin<float>(float x) { return sinf(x); }

// ShiftedCosine switching function and its derivative (cutoff_functions.hh:69-107)
// times the optional RadialScaling factor (:211-283).  Evaluated in FP64: it
// is one value per pair.
__device__ inline void cutoff_eval(const RadialDev& R, double r, double* fc, double* dfc) {
  double f, df;
  if (r <= R.rc - R.sw) {
    f = 1.0;
    df = 0.0;
  } else if (r > R.rc) {
    f = 0.0;
    df = 0.0;
  } else {
    double rs = kPi * (r - R.rc + R.sw) / R.sw;
    f = 0.5 * (1. + cos(rs));
    df = -0.5 * kPi / R.sw * sin(rs);
  }
  if (R.cutoff_type == 1) {
    double u, du;
    if (fabs(R.rs_rate) <= kDblFtol) {
      double q = pow(r / R.rs_scale, R.rs_exponent);
      u = 1. / q;
      du = -R.rs_exponent / r / q;
    } else if (R.rs_exponent == 0) {
      u = 1.;
      du = 0.;
    } else {
      double ff = pow(r / R.rs_scale, R.rs_exponent);
      u = R.rs_rate / (R.rs_rate + ff);
      du = -R.rs_rate * R.rs_exponent * ff / r / ((R.rs_rate + ff) * (R.rs_rate + ff));
    }
    double fs = f;
    f = u * fs;
    df = du * fs + u * df;
  }
  *fc = f;
  if (dfc) *dfc = df;
}

// --- Hyp1f1 (FP64) -------------------------------------------------------
__device__ inline double pow_u_dev(double x, unsigned n) {
  double v = 1.0;
  do {
    if (n & 1) v *= x;
    n >>= 1;
    x *= x;
  } while (n);
  return v;
}

// adaptive power series, four terms per step (hyp1f1.cc:123-153)
__device__ inline double hyp_series_sum(const double* __restrict__ c, double z, double tol) {
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

// asymptotic 2F0 series (hyp1f1.cc:248-270)
__device__ inline double hyp_asym_sum(const double* __restrict__ c, double z, double tol) {
  double iz = 1.0 / z, res = 1.0, izpow = 1.0;
  for (int i = 0; i < kHypTerms; ++i) {
    izpow *= iz;
    double s = c[i] * izpow;
    if (res > 0 && fabs(s) < tol * res) break;
    res += s;
  }
  return res;
}

// G(a,b,z) = Gamma(a)/Gamma(b) exp(z2) 1F1(a,b,z), or its z-derivative
// form (hyp1f1.cc:72-82, :197-208, :345-354)
__device__ inline double hyp_g(const RadialDev& R, int idx, double z, double z2, double ez2,
                               bool der) {
  const HypCoef hc = R.hyp[idx];
  const double tol = 1e-13;
  if (z > hc.z_switch) {
    if (hc.is_exp) return exp(z + z2) / hc.prefac_a;
    const double* c = R.ha + (size_t(idx) * 2 + (der ? 1 : 0)) * kHypTerms;
    double zab;
    if (hc.is_n_and_l) {
      int e = int(2 * (hc.a - hc.b));
      zab = sqrt(e < 0 ? pow_u_dev(1.0 / z, unsigned(-e)) : pow_u_dev(z, unsigned(e)));
    } else {
      zab = pow(z, hc.a - hc.b);
    }
    return hyp_asym_sum(c, z, tol) * exp(z + z2) * zab;
  }
  if (hc.is_exp) return hc.prefac_s * exp(z + z2);
  const double* c = R.hc + (size_t(idx) * 2 + (der ? 1 : 0)) * kHypTerms;
  double f = hyp_series_sum(c, z, tol);
  if (der) f = f * hc.a / hc.b;
  return hc.prefac_s * f * ez2;
}

// Plain 1F1(a, b, z) from table entry idx (series below the switching
// point, asymptotic above; hyp1f1.hh:185-195)
__device__ inline double hyp_plain(const RadialDev& R, int idx, double z) {
  const HypCoef hc = R.hyp[idx];
  const double tol = 1e-13;
  if (z > hc.z_switch) {
    const double* c = R.ha + size_t(idx) * 2 * kHypTerms;
    const int e = int(2 * (hc.a - hc.b));
    const double zab = hc.is_n_and_l ? sqrt(e < 0 ? pow_u_dev(1.0 / z, unsigned(-e)) : pow_u_dev(z, unsigned(e)))
                                     : pow(z, hc.a - hc.b);
    return hc.prefac_a * exp(z) * zab * hyp_asym_sum(c, z, tol);
  }
  return hyp_series_sum(R.hc + size_t(idx) * 2 * kHypTerms, z, tol);
}

// DVR radial functions of Gauss-Legendre point n at distance r
// (calculator_spherical_expansion.hh:809-900 with math/bessel.cc:30-190):
// out[l] = sqrt(w_n) x_n exp(-a (r^2 + x_n^2)) i_l(2 a r x_n), i_l the
// modified spherical Bessel function, from a downward recursion started by
// 1F1 at the top two orders (argument <= 50) or an upward one from the
// closed forms of i_0, i_1 (above); values below 1e-100 are zero.  dout[l]
// its r-derivative (the recursion for the gradients, one order above L).
__device__ inline void dvr_radial_n(const RadialDev& R, int n, double r, double* out,
                                    double* dout) {
  const int L = R.L, omax = L + 2;
  const double a = R.fac_a, x = R.dvr_x[n], f = R.dvr_f[n];
  double v[kMaxL1 + 1];
  if (r < 1e-5) {
    // zeroth-order approximation (bessel.cc:140-151)
    const double v0 = exp(-x * x * a);
    for (int l = 0; l <= L; ++l) {
      out[l] = l == 0 && v0 >= 1e-100 ? f * v0 : 0.0;
      if (dout) dout[l] = 0.0;
    }
    return;
  }
  const double z = 2.0 * a * r * x, iz = 1.0 / z;
  if (z > 50) {
    // upward (bessel.cc:60-82): i_0 = sinh(z) / z, i_1 = cosh(z) / z - i_0 / z
    const double em = exp(-a * (x - r) * (x - r)), ep = exp(-a * (x + r) * (x + r));
    v[0] = (em - ep) * 0.5 * iz;
    v[1] = (em + ep) * 0.5 * iz - v[0] * iz;
    for (int o = 2; o < omax; ++o) v[o] = v[o - 2] - v[o - 1] * (2.0 * o - 1.0) * iz;
  } else {
    // downward (bessel.cc:84-110)
    const double efac = exp(-a * r * r) * exp(-a * x * x), ez = exp(-z);
    for (int k = 0; k < 2; ++k) {
      const int o = omax - 2 + k;
      const double ig = 1.0 / tgamma(1.5 + o);
      v[o] = ez * ig * pow_u_dev(z * 0.5, unsigned(o)) * 0.5 * 1.77245385090551602730 *
             hyp_plain(R, k, 2.0 * z) * efac;
    }
    for (int o = omax - 3; o >= 0; --o) v[o] = v[o + 2] + v[o + 1] * (2.0 * o + 3.0) * iz;
  }
  for (int o = 0; o < omax; ++o)
    if (v[o] < 1e-100) v[o] = 0.0;
  for (int l = 0; l <= L; ++l) out[l] = f * v[l];
  if (dout) {
    // bessel.cc:112-130
    const double g = 2.0 * a * x;
    for (int l = 0; l <= L; ++l) {
      double d = -2.0 * a * r * v[l];
      d += l == 0 ? g * v[1] : g * (l * v[l - 1] + (l + 1) * v[l + 1]) / (2 * l + 1);
      dout[l] = f * d;
    }
  }
}

// Raw GTO radial integral R[l] (and dR/dr) for one n at distance r
// (calculator_spherical_expansion.hh:518-584 with hyp1f1.cc:395-491).
// out/dout: L1 values each (dout may be null).
__device__ inline void gto_radial_n(const RadialDev& R, int n, double r, double* out,
                                    double* dout) {
  const int L = R.L, L1 = R.L1;
  const double fa = R.fac_a, fb = R.fac_b[n];
  const double arij = fa * r, z2 = -r * arij, ez2 = exp(z2);
  const double zz = (arij * fa) * (1.0 / (fa + fb));
  const double z = zz * r, dzdr = zz * 2;
  double G[kMaxL1], dG[kMaxL1];
  if (R.recursion) {
    int l = L;
    double M1p2p = hyp_g(R, n * L1 + l, z, z2, ez2, false);
    double M2p3p = hyp_g(R, n * L1 + l, z, z2, ez2, true);
    G[l] = M1p2p;
    dG[l] = M2p3p;
    double MP1p2p = hyp_g(R, n * L1 + l - 1, z, z2, ez2, false);
    double MP2p3p = hyp_g(R, n * L1 + l - 1, z, z2, ez2, true);
    G[l - 1] = MP1p2p;
    dG[l - 1] = MP2p3p;
    for (l -= 2; l > 0; l -= 2) {
      double a = 0.5 * (n + l + 3), b = l + 1.5;
      double M1p1p = z * M2p3p + M1p2p * (b + 1);
      double Moo = (z * (a - b) * M1p2p + M1p1p * b) / a;
      G[l] = Moo;
      dG[l] = M1p1p;
      M2p3p = M1p1p;
      M1p2p = Moo;
      a = 0.5 * (n + l - 1 + 3);
      b = l - 1 + 1.5;
      double MP1p1p = z * MP2p3p + MP1p2p * (b + 1);
      double MPoo = (z * (a - b) * MP1p2p + MP1p1p * b) / a;
      G[l - 1] = MPoo;
      dG[l - 1] = MP1p1p;
      MP2p3p = MP1p1p;
      MP1p2p = MPoo;
    }
    if (L % 2 == 0) {
      double a = 0.5 * (n + 3), b = 1.5;
      double M1p1p = z * M2p3p + M1p2p * (b + 1);
      double Moo = (z * (a - b) * M1p2p + M1p1p * b) / a;
      G[0] = Moo;
      dG[0] = M1p1p;
    }
  } else {
    for (int l = 0; l <= L; ++l) {
      G[l] = hyp_g(R, n * L1 + l, z, z2, ez2, false);
      if (dout) dG[l] = hyp_g(R, n * L1 + l, z, z2, ez2, true);
    }
  }
  // (a + b_n)^{-(3+n+l)/2} and (r a)^l
  const double ab = 1.0 / sqrt(fa + fb);
  double abl = pow_u_dev(ab, unsigned(3 + n));
  double dfal = 1.0;
  const double dfa = r * fa;
  for (int l = 0; l <= L; ++l) {
    if (l > 0) {
      abl *= ab;
      dfal *= dfa;
    }
    double v = abl * G[l] * dfal;
    out[l] = v;
    if (dout) {
      double dg = dG[l] * dzdr - (2 * arij) * G[l];
      dout[l] = abl * dg * dfal + v * (double(l) / r);
    }
  }
}

// Raw radial functions of basis function n at distance r (GTO: before the
// orthonormalisation, applied by the caller; DVR: final)
__device__ inline void radial_n(const RadialDev& R, int n, double r, double* out, double* dout) {
  if (R.basis == 1)
    dvr_radial_n(R, n, r, out, dout);
  else
    gto_radial_n(R, n, r, out, dout);
}

// Cubic spline lookup of all F = N*L1 functions at r for function f
// (math/interpolator.cc:177-230): the cell left of the nearest grid point.
template <class T>
struct SplineCell {
  int klo;
  T a, b, inv_h;
};

template <class T>
__device__ inline SplineCell<T> spline_cell(const RadialDev& R, double r) {
  SplineCell<T> c;
  int k = int((r - R.x1) * R.per_unit) - 1;
  k = k < 0 ? 0 : (k > R.search_size ? R.search_size : k);
  c.klo = k;
  double xhi = R.grid[k + 1];
  double a = (xhi - r) / R.h;
  c.a = T(a);
  c.b = T(1 - a);
  c.inv_h = T(1.0 / R.h);
  return c;
}

template <class T>
struct Spl4;
template <>
struct Spl4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ const type* table(const RadialDev& R) {
    return reinterpret_cast<const type*>(R.spl32);
  }
};
template <>
struct Spl4<double> {
  typedef double type __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ const type* table(const RadialDev& R) {
    return reinterpret_cast<const type*>(R.spl64);
  }
};

// cubic of cell c.klo for function f at the point the cell was made for
// (math/interpolator.hh:166-186)
template <class T>
__device__ __forceinline__ void spline_eval(const RadialDev& R, const SplineCell<T>& c, int f,
                                            T* val, T* der) {
  const int F = R.N * R.L1;
  const typename Spl4<T>::type v = Spl4<T>::table(R)[size_t(c.klo) * F + f];
  const T ylo = v.x, dlo = v.y, yhi = v.z, dhi = v.w;
  const T a = c.a, b = c.b;
  *val = a * (ylo + (a * a - T(1)) * dlo) + b * (yhi + (b * b - T(1)) * dhi);
  if (der)
    *der = (yhi - ylo - (T(3) * a * a - T(1)) * dlo + (T(3) * b * b - T(1)) * dhi) * c.inv_h;
}

// --- real spherical harmonics --------------------------------------------
// Y[(L+1)^2] at unit direction (x,y,z); when dY != null also the gradient on
// the unit sphere dY[3][(L+1)^2] (no 1/r factor).  ylm_ab holds the recursion
// coefficients: a[l][m] at [l*(L+1)+m], b at [(L+1)^2 + l*(L+1)+m].
template <class T>
__device__ inline void ylm_eval(int L, const T* __restrict__ ylm_ab, T x, T y, T z, T* Y,
                                T* dY) {
  const int L1 = L + 1, LL = L1 * L1;
  T n2 = x * x + y * y + z * z;
  if (fabs(double(n2) - 1.0) > (sizeof(T) == 8 ? kDblFtol : 1e-6)) {
    T inv = T(1) / sqrt(n2);
    x *= inv;
    y *= inv;
    z *= inv;
  }
  const T cos_t = z;
  const T sxy = sqrt(x * x + y * y);
  T cos_p = T(1), sin_p = T(0);
  if (double(sxy) >= kDblFtol) {
    cos_p = x / sxy;
    sin_p = y / sxy;
  }
  const T sin_t = sqrt(T(1) - cos_t * cos_t);
  // associated Legendre functions, normalisation folded in (:97-133)
  T P[(kMaxL1) * (kMaxL1 + 1)];
  const int W = L + 2;
  for (int k = 0; k < L1 * W; ++k) P[k] = T(0);
  const T s2pi = T(0.39894228040143267794);  // sqrt(1/(2 pi))
  T acc = s2pi;
  P[0] = s2pi;
  if (L > 0) {
    P[W] = cos_t * T(1.73205080756887729353) * s2pi;
    acc = acc * -T(1.22474487139158904910) * sin_t;
    P[W + 1] = acc;
  }
  for (int l = 2; l <= L; ++l) {
    for (int m = 0; m < l - 1; ++m) {
      T t = cos_t * P[(l - 1) * W + m];
      t += ylm_ab[LL + l * L1 + m] * P[(l - 2) * W + m];
      t *= ylm_ab[l * L1 + m];
      P[l * W + m] = t;
    }
    P[l * W + l - 1] = acc * cos_t * sqrt(T(2 * l + 1));
    acc = acc * sin_t * (-sqrt(T(1) + T(0.5) / T(l)));
    P[l * W + l] = acc;
  }
  // cos(m phi), sin(m phi) by Chebyshev recurrence
  T cm[kMaxL1], sm[kMaxL1];
  cm[0] = T(1);
  sm[0] = T(0);
  if (L > 0) {
    cm[1] = cos_p;
    sm[1] = sin_p;
  }
  for (int m = 2; m <= L; ++m) {
    cm[m] = T(2) * cos_p * cm[m - 1] - cm[m - 2];
    sm[m] = T(2) * cos_p * sm[m - 1] - sm[m - 2];
  }
  const T isq2 = T(0.70710678118654752440);
  for (int l = 0; l <= L; ++l) {
    const int base = l * l + l;
    for (int m = 1; m <= l; ++m) {
      Y[base + m] = P[l * W + m] * cm[m];
      Y[base - m] = P[l * W + m] * sm[m];
    }
    Y[base] = P[l * W] * isq2;
  }
  if (!dY) return;
  // gradient on the sphere (:227-363); sin(theta) taken as sqrt(x^2+y^2)
  const T st = sxy;
  T* dx = dY;
  T* dyv = dY + LL;
  T* dz = dY + 2 * LL;
  dx[0] = dyv[0] = dz[0] = T(0);
  for (int l = 1; l <= L; ++l) {
    const int base = l * l + l;
    const T pl0 = sqrt(T(l) * T(l + 1));
    const T q = pl0 * isq2 * P[l * W + 1];
    dx[base] = cos_t * cos_p * q;
    dyv[base] = cos_t * sin_p * q;
    dz[base] = -st * q;
    for (int k = 0; k < l; ++k) {
      const int m = k + 1;
      const T fk = sqrt(T(l - k) * T(l + k + 1));
      const T fk1 = sqrt(T(l - k - 1) * T(l + k + 2));
      const T pk2 = (k + 2 <= l) ? P[l * W + k + 2] : T(0);
      const T lpd = fk * P[l * W + k] - fk1 * pk2;
      T pdf;
      if (st > T(0.1))
        pdf = T(k + 1) / st * P[l * W + k + 1];
      else
        pdf = T(-0.5) / cos_t * (fk * P[l * W + k] + fk1 * pk2);
      const T hc = T(-0.5) * cos_t;
      dx[base + m] = sin_p * pdf * sm[m] + hc * cos_p * cm[m] * lpd;
      dyv[base + m] = -cos_p * pdf * sm[m] + hc * sin_p * cm[m] * lpd;
      dz[base + m] = T(0.5) * st * cm[m] * lpd;
      dx[base - m] = -sin_p * pdf * cm[m] + hc * cos_p * sm[m] * lpd;
      dyv[base - m] = cos_p * pdf * cm[m] + hc * sin_p * sm[m] * lpd;
      dz[base - m] = T(0.5) * st * sm[m] * lpd;
    }
  }
}

}  // namespace rgpu
