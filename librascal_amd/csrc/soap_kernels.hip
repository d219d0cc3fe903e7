// SOAP expansion, power spectrum and their adjoint on gfx950.
//
// One workgroup (256 threads, 4 wavefronts) owns one centre.  Pairs are
// processed in batches: per pair the radial integrals come from the spline
// table (coalesced row reads) or from FP64 Hyp1f1, the real spherical
// harmonics from a wavefront-level recursion (one lane per m, lockstep over
// l, neighbouring-m values by lane shuffles), and the per-centre coefficients
// accumulate in registers, each thread owning a fixed set of (species, n, lm)
// entries.  The result is identical to the reference's per-pair outer-product
// accumulation (calculator_spherical_expansion.hh:2234-2410) up to rounding.
#include <hip/hip_runtime.h>

#include "soap_device.hh"
#include "soap_kernels.hh"

namespace rgpu {

namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// Real spherical harmonics of nq pairs, one wavefront segment of (L+1) lanes
// per pair.  dirs[q][3] (unit vectors), writes Y[q][LL] and optionally
// dY[q][3][LL].  Every lane of the wavefront must call this (shuffles).
template <class T>
__device__ void ylm_wave(int L, const T* dirs, int q0, int nq, T* Y, T* dY, int ystride,
                         int dystride) {
  const int L1 = L + 1;
  const int lane = threadIdx.x & 63;
  const int seg = lane / L1;
  const int m = lane - seg * L1;
  const int per_wave = 64 / L1;
  const int q = q0 + seg;
  const bool active = seg < per_wave && q < nq;
  T x = T(0), y = T(0), z = T(1);
  if (active) {
    x = dirs[3 * q];
    y = dirs[3 * q + 1];
    z = dirs[3 * q + 2];
  }
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
  // cos(k phi), sin(k phi) for k = m and m+1 (Chebyshev recurrence)
  T c_m = T(1), s_m = T(0), c_m1 = cos_p, s_m1 = sin_p;
  for (int k = 1; k <= m; ++k) {
    const T cn = T(2) * cos_p * c_m1 - c_m;
    const T sn = T(2) * cos_p * s_m1 - s_m;
    c_m = c_m1;
    s_m = s_m1;
    c_m1 = cn;
    s_m1 = sn;
  }
  // diagonal P(m, m) with the sqrt(1/(2 pi)) normalisation folded in
  const T s2pi = T(0.39894228040143267794);
  T diag = s2pi;
  for (int k = 1; k <= m; ++k) {
    if (k == 1)
      diag = diag * -T(1.22474487139158904910) * sin_t;
    else
      diag = diag * sin_t * (-sqrt(T(1) + T(0.5) / T(k)));
  }
  const T isq2 = T(0.70710678118654752440);
  T P1 = T(0), P2 = T(0);
  T* Yq = Y + q * ystride;
  T* dYq = dY ? dY + q * dystride : nullptr;
  const int LL = L1 * L1;
  for (int l = 0; l <= L; ++l) {
    T P = T(0);
    if (m == l) {
      P = diag;
    } else if (m == l - 1) {
      P = (l == 1) ? cos_t * T(1.73205080756887729353) * s2pi : diag * cos_t * sqrt(T(2 * l + 1));
    } else if (m < l - 1) {
      const T ll = T(l) * T(l), mm = T(m) * T(m), l1 = T(l - 1) * T(l - 1);
      const T a = sqrt((T(4) * ll - T(1)) / (ll - mm));
      const T b = -sqrt((l1 - mm) / (T(4) * l1 - T(1)));
      T t = cos_t * P1;
      t += b * P2;
      t *= a;
      P = t;
    }
    const int base = l * l + l;
    if (active && m <= l) {
      if (m == 0) {
        Yq[base] = P * isq2;
      } else {
        Yq[base + m] = P * c_m;
        Yq[base - m] = P * s_m;
      }
    }
    if (dYq) {
      const T Pn1 = __shfl(P, lane + 1, 64);
      const T Pn2 = __shfl(P, lane + 2, 64);
      if (active && l >= 1 && m <= l) {
        if (m == 0) {
          const T qv = sqrt(T(l) * T(l + 1)) * isq2 * Pn1;
          dYq[base] = cos_t * cos_p * qv;
          dYq[LL + base] = cos_t * sin_p * qv;
          dYq[2 * LL + base] = -sxy * qv;
        }
        if (m < l) {
          const int k = m, M = m + 1;
          const T fk = sqrt(T(l - k) * T(l + k + 1));
          const T fk1 = sqrt(T(l - k - 1) * T(l + k + 2));
          const T pk2 = (k + 2 <= l) ? Pn2 : T(0);
          const T lpd = fk * P - fk1 * pk2;
          T pdf;
          if (sxy > T(0.1))
            pdf = T(k + 1) / sxy * Pn1;
          else
            pdf = T(-0.5) / cos_t * (fk * P + fk1 * pk2);
          const T hc = T(-0.5) * cos_t;
          dYq[base + M] = sin_p * pdf * s_m1 + hc * cos_p * c_m1 * lpd;
          dYq[LL + base + M] = -cos_p * pdf * s_m1 + hc * sin_p * c_m1 * lpd;
          dYq[2 * LL + base + M] = T(0.5) * sxy * c_m1 * lpd;
          dYq[base - M] = -sin_p * pdf * c_m1 + hc * cos_p * s_m1 * lpd;
          dYq[LL + base - M] = cos_p * pdf * c_m1 + hc * sin_p * s_m1 * lpd;
          dYq[2 * LL + base - M] = T(0.5) * sxy * s_m1 * lpd;
        }
      }
    }
    P2 = P1;
    P1 = P;
  }
  if (dYq && active && m == 0) {
    dYq[0] = T(0);
    dYq[LL] = T(0);
    dYq[2 * LL] = T(0);
  }
}

// Radial values (and derivatives) of the batch: R[q][n*L1+l].
template <class T>
__device__ void radial_batch(const RadialDev& rad, const double* dist, int nb, T* R, T* dR) {
  const int F = rad.N * rad.L1;
  if (rad.use_spline) {
    for (int it = threadIdx.x; it < nb * F; it += blockDim.x) {
      const int q = it / F, f = it - q * F;
      SplineCell<T> cell = spline_cell<T>(rad, dist[q]);
      T v, d;
      spline_eval<T>(rad, cell, f, &v, dR ? &d : nullptr);
      R[q * F + f] = v;
      if (dR) dR[q * F + f] = d;
    }
  } else {
    for (int it = threadIdx.x; it < nb * rad.N; it += blockDim.x) {
      const int q = it / rad.N, n = it - q * rad.N;
      double v[kMaxL1], d[kMaxL1];
      gto_radial_n(rad, n, dist[q], v, dR ? d : nullptr);
      for (int l = 0; l < rad.L1; ++l) {
        R[q * F + n * rad.L1 + l] = T(v[l]);
        if (dR) dR[q * F + n * rad.L1 + l] = T(d[l]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Expansion forward.
constexpr int kPB = 16;  // pairs per batch

template <class T, int KPT>
__global__ __launch_bounds__(kThreads) void expansion_kernel(RadialDev rad, int S, PairView pv,
                                                             T* __restrict__ cout,
                                                             unsigned long long* __restrict__ keymask) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, L = rad.L, L1 = rad.L1, LL = rad.LL;
  const int F = N * L1;
  const int B = N * LL;          // entries per species block
  const int E = S * B;           // entries per centre
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Ysh = reinterpret_cast<T*>(smem_raw);               // [kPB][LL]
  T* Rsh = Ysh + kPB * LL;                               // [kPB][F]
  T* dirsh = Rsh + kPB * F;                              // [kPB][3]
  T* fcsh = dirsh + 3 * kPB;                             // [kPB]
  double* distsh = reinterpret_cast<double*>(fcsh + kPB);  // [kPB]
  int* spsh = reinterpret_cast<int*>(distsh + kPB);        // [kPB]
  __shared__ unsigned long long kmask;

  const int64_t p0 = pv.centre_off[i], p1 = pv.centre_off[i + 1];
  const int np = int(p1 - p0) - 1;  // without the self pair
  const int si = pv.species_of[i];
  if (threadIdx.x == 0) kmask = 1ull << si;

  // thread-owned entries
  T acc[KPT];
  int ent_s[KPT], ent_nl[KPT], ent_lm[KPT];
  const double inv_s4pi = 0.28209479177387814347;  // 1/sqrt(4 pi)
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int e = threadIdx.x + kThreads * k;
    acc[k] = T(0);
    ent_s[k] = -1;
    ent_nl[k] = 0;
    ent_lm[k] = 0;
    if (e < E) {
      const int s = e / B, r = e - s * B, n = r / LL, lm = r - n * LL;
      int l = 0;
      while ((l + 1) * (l + 1) <= lm) ++l;
      ent_s[k] = s;
      ent_nl[k] = n * L1 + l;
      ent_lm[k] = lm;
      // centre contribution first (expansion.hh:2244-2247)
      if (s == si && lm == 0) acc[k] = T(rad.center[n] * inv_s4pi);
    }
  }
  __syncthreads();

  for (int q0 = 0; q0 < np; q0 += kPB) {
    const int nb = min(kPB, np - q0);
    if (threadIdx.x < nb) {
      const int q = threadIdx.x;
      const int64_t p = p0 + 1 + q0 + q;
      const double d = pv.pair_dist[p];
      const int j = pv.pair_j[p];
      distsh[q] = d;
      dirsh[3 * q] = T(pv.pair_vec[3 * p] / d);
      dirsh[3 * q + 1] = T(pv.pair_vec[3 * p + 1] / d);
      dirsh[3 * q + 2] = T(pv.pair_vec[3 * p + 2] / d);
      double fc;
      cutoff_eval(rad, d, &fc, nullptr);
      fcsh[q] = T(fc);
      const int sj = pv.species_of[j];
      spsh[q] = sj;
      atomicOr(&kmask, 1ull << sj);
    }
    __syncthreads();
    {
      const int per_wave = 64 / L1;
      const int wave = threadIdx.x >> 6;
      for (int qb = wave * per_wave; qb < nb; qb += 4 * per_wave)
        ylm_wave<T>(L, dirsh, qb, nb, Ysh, nullptr, LL, 0);
    }
    radial_batch<T>(rad, distsh, nb, Rsh, nullptr);
    __syncthreads();
    for (int q = 0; q < nb; ++q) {
      const int s = spsh[q];
      const T fq = fcsh[q];
      const T* Yq = Ysh + q * LL;
      const T* Rq = Rsh + q * F;
#pragma unroll
      for (int k = 0; k < KPT; ++k)
        if (ent_s[k] == s) acc[k] += (Rq[ent_nl[k]] * Yq[ent_lm[k]]) * fq;
    }
    __syncthreads();
  }

  T* ci = cout + i * int64_t(E);
  if (rad.use_spline) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + kThreads * k;
      if (e < E) ci[e] = acc[k];
    }
  } else {
    // c <- O^T c per species block (expansion.hh:605-618)
    T* blk = Ysh;  // reuse: needs N*LL entries
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const int e = threadIdx.x + kThreads * k;
        if (ent_s[k] == s) blk[e - s * B] = acc[k];
      }
      __syncthreads();
      for (int r = threadIdx.x; r < B; r += kThreads) {
        const int n = r / LL, lm = r - n * LL;
        double v = 0;
        for (int mm = 0; mm < N; ++mm) v += rad.O[mm * N + n] * double(blk[mm * LL + lm]);
        ci[s * B + r] = T(v);
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) keymask[i] = kmask;
}

// ---------------------------------------------------------------------------
// Power spectrum: p^{ab}[n1 N + n2][l] = (2l+1)^{-1/2} (sqrt 2 if a<b)
// sum_m c_a[n1,lm] c_b[n2,lm] (calculator_spherical_invariants.hh:747-788),
// 4x4 register tiles over (n1, n2) per (key, l).  Un-normalised; |p|^2 is
// returned per centre.
template <class T>
__global__ __launch_bounds__(kThreads) void power_spectrum_kernel(int N, int L, int S,
                                                                  const T* __restrict__ c,
                                                                  T* __restrict__ p,
                                                                  double* __restrict__ norm2) {
  const int64_t i = blockIdx.x;
  const int L1 = L + 1, LL = L1 * L1, B = N * LL;
  const int K = S * (S + 1) / 2;
  const int NT = (N + 3) / 4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* cs = reinterpret_cast<T*>(smem_raw);  // [S][N][LL]
  __shared__ double red[kThreads / 64];
  const T* ci = c + i * int64_t(S) * B;
  for (int e = threadIdx.x; e < S * B; e += kThreads) cs[e] = ci[e];
  __syncthreads();
  const int64_t D = int64_t(K) * N * N * L1;
  T* pi = p + i * D;
  double ss = 0;
  const int items = K * L1 * NT * NT;
  for (int it = threadIdx.x; it < items; it += kThreads) {
    int r = it;
    const int t2 = r % NT;
    r /= NT;
    const int t1 = r % NT;
    r /= NT;
    const int l = r % L1;
    const int key = r / L1;
    int a = 0, kk = key;
    while (kk >= S - a) {
      kk -= S - a;
      ++a;
    }
    const int b = a + kk;
    const T* ca = cs + a * B;
    const T* cb = cs + b * B;
    const int n1 = 4 * t1, n2 = 4 * t2;
    T acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = T(0);
    for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm) {
      T va[4], vb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        va[u] = (n1 + u < N) ? ca[(n1 + u) * LL + lm] : T(0);
        vb[u] = (n2 + u < N) ? cb[(n2 + u) * LL + lm] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += va[u] * vb[v];
    }
    T f = T(1) / sqrt(T(2 * l + 1));
    if (a < b) f *= T(1.41421356237309504880);
    T* pk = pi + int64_t(key) * N * N * L1;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (n1 + u < N && n2 + v < N) {
          const T val = acc[u][v] * f;
          pk[((n1 + u) * N + n2 + v) * L1 + l] = val;
          ss += double(val) * double(val);
        }
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    norm2[i] = t;
  }
}

template <class T>
__global__ void normalise_kernel(int64_t n, int64_t D, const T* __restrict__ p,
                                 const double* __restrict__ norm2, int normalize,
                                 double* __restrict__ out) {
  const int64_t total = n * D;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e / D;
    double v = double(p[e]);
    if (normalize) {
      const double nr = sqrt(norm2[i]);
      if (nr > 0) v /= nr;
    }
    out[e] = v;
  }
}

// ---------------------------------------------------------------------------
// Adjoint, step 1: dL/dX -> dL/dc_raw per centre.
//   dL/dp = (G - X (X.G)) / |p| when normalised (calculator_spherical_
//   invariants.hh:588-652 is the forward form of this projection);
//   dL/dc_a[k,lm] = sum over keys containing a of the transposed contraction
//   of the power-spectrum product; finally dL/dc_raw = O dL/dc for the direct
//   (un-splined) radial basis.
template <class T, int KPT>
__global__ __launch_bounds__(kThreads) void backward_dc_kernel(RadialDev rad, int S,
                                                               const T* __restrict__ c,
                                                               const T* __restrict__ p,
                                                               const double* __restrict__ norm2,
                                                               int normalize, const T* __restrict__ G,
                                                               T* __restrict__ W) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, B = N * LL, E = S * B;
  const int K = S * (S + 1) / 2;
  const int64_t D = int64_t(K) * N * N * L1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* cs = reinterpret_cast<T*>(smem_raw);  // [S][N][LL]
  __shared__ double red[kThreads / 64];
  __shared__ double sh_alpha, sh_beta;
  const T* ci = c + i * int64_t(E);
  const T* pi = p + i * D;
  const T* gi = G + i * D;
  for (int e = threadIdx.x; e < E; e += kThreads) cs[e] = ci[e];
  // alpha = 1/|p|, beta = (p.G)/|p|^3 so that dL/dp = alpha G - beta p
  double dot = 0;
  if (normalize)
    for (int64_t o = threadIdx.x; o < D; o += kThreads) dot += double(pi[o]) * double(gi[o]);
  for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    const double n2 = norm2[i];
    if (normalize && n2 > 0) {
      const double inv = 1.0 / sqrt(n2);
      sh_alpha = inv;
      sh_beta = t * inv * inv * inv;
    } else {
      sh_alpha = 1.0;
      sh_beta = 0.0;
    }
  }
  __syncthreads();
  const T alpha = T(sh_alpha), beta = T(sh_beta);
  T out[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int e = threadIdx.x + kThreads * k;
    out[k] = T(0);
    if (e >= E) continue;
    const int a = e / B, r = e - a * B, kn = r / LL, lm = r - kn * LL;
    int l = 0;
    while ((l + 1) * (l + 1) <= lm) ++l;
    const T lf = T(1) / sqrt(T(2 * l + 1));
    T sum = T(0);
    for (int b = 0; b < S; ++b) {
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      const int key = lo * S - lo * (lo - 1) / 2 + (hi - lo);
      const T w = (lo < hi) ? T(1.41421356237309504880) : T(1);
      const int64_t kb = int64_t(key) * N * N * L1;
      const T* cb = cs + b * B;
      T s1 = T(0);
      for (int n = 0; n < N; ++n) {
        // element (a's index k, b's index n) of the key block
        int64_t o1 = (a <= b) ? kb + (int64_t(kn) * N + n) * L1 + l : kb + (int64_t(n) * N + kn) * L1 + l;
        T g1 = alpha * gi[o1] - beta * pi[o1];
        T term = g1;
        if (a == b) {
          int64_t o2 = kb + (int64_t(n) * N + kn) * L1 + l;
          term += alpha * gi[o2] - beta * pi[o2];
        }
        s1 += term * cb[n * LL + lm];
      }
      sum += w * lf * s1;
    }
    out[k] = sum;
  }
  T* Wi = W + i * int64_t(E);
  if (rad.use_spline) {
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + kThreads * k;
      if (e < E) Wi[e] = out[k];
    }
  } else {
    __syncthreads();
    // reuse cs as the staging block for dL/dc -> O dL/dc
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int e = threadIdx.x + kThreads * k;
      if (e < E) cs[e] = out[k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < E; e += kThreads) {
      const int s = e / B, r = e - s * B, mrow = r / LL, lm = r - mrow * LL;
      double v = 0;
      for (int n = 0; n < N; ++n) v += rad.O[mrow * N + n] * double(cs[s * B + n * LL + lm]);
      Wi[e] = T(v);
    }
  }
}

// Adjoint, step 2: per pair g = dL/dr_j for the pair (i, j):
//   g = sum_{n,lm} W[s_j][n][lm] [ (R' fc + R fc') Y r_hat + R fc grad_Y / r ]
// (the pair-gradient formula of calculator_spherical_expansion.hh:2306-2366
// contracted with W), scattered to grad[j] += g and grad[i] -= g unless j is
// a periodic image of i; the virial sum r_ji (x) g is reduced per centre.
constexpr int kPBB = 4;  // pairs per batch in the adjoint (one per wavefront)

template <class T>
__global__ __launch_bounds__(kThreads) void backward_pairs_kernel(RadialDev rad, int S, PairView pv,
                                                                  const T* __restrict__ W,
                                                                  double* __restrict__ grad,
                                                                  double* __restrict__ centre_virial) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, L = rad.L, L1 = rad.L1, LL = rad.LL;
  const int F = N * L1, B = N * LL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Ysh = reinterpret_cast<T*>(smem_raw);  // [kPBB][LL]
  T* dYsh = Ysh + kPBB * LL;                // [kPBB][3][LL]
  T* Rsh = dYsh + kPBB * 3 * LL;            // [kPBB][F]
  T* dRsh = Rsh + kPBB * F;                 // [kPBB][F]
  T* dirsh = dRsh + kPBB * F;               // [kPBB][3]
  double* distsh = reinterpret_cast<double*>(dirsh + 3 * kPBB);  // [kPBB]
  double* fcsh = distsh + kPBB;                                  // [kPBB]
  double* dfcsh = fcsh + kPBB;                                   // [kPBB]
  int* spsh = reinterpret_cast<int*>(dfcsh + kPBB);              // [kPBB]
  __shared__ double vir[kThreads / 64][6];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t p0 = pv.centre_off[i], p1 = pv.centre_off[i + 1];
  const int np = int(p1 - p0) - 1;
  const T* Wi = W + i * int64_t(S) * B;
  double v6[6] = {0, 0, 0, 0, 0, 0};
  for (int q0 = 0; q0 < np; q0 += kPBB) {
    const int nb = min(kPBB, np - q0);
    if (threadIdx.x < nb) {
      const int q = threadIdx.x;
      const int64_t p = p0 + 1 + q0 + q;
      const double d = pv.pair_dist[p];
      distsh[q] = d;
      dirsh[3 * q] = T(pv.pair_vec[3 * p] / d);
      dirsh[3 * q + 1] = T(pv.pair_vec[3 * p + 1] / d);
      dirsh[3 * q + 2] = T(pv.pair_vec[3 * p + 2] / d);
      double fc, dfc;
      cutoff_eval(rad, d, &fc, &dfc);
      fcsh[q] = fc;
      dfcsh[q] = dfc;
      spsh[q] = pv.species_of[pv.pair_j[p]];
    }
    __syncthreads();
    {
      const int per_wave = 64 / L1;
      for (int qb = wave * per_wave; qb < nb; qb += 4 * per_wave)
        ylm_wave<T>(L, dirsh, qb, nb, Ysh, dYsh, LL, 3 * LL);
    }
    radial_batch<T>(rad, distsh, nb, Rsh, dRsh);
    __syncthreads();
    if (wave < nb) {
      const int q = wave;
      const T fc = T(fcsh[q]), dfc = T(dfcsh[q]);
      const T* Ws = Wi + spsh[q] * B;
      const T* Yq = Ysh + q * LL;
      const T* dYq = dYsh + q * 3 * LL;
      const T* Rq = Rsh + q * F;
      const T* dRq = dRsh + q * F;
      T A = T(0), Bx = T(0), By = T(0), Bz = T(0);
      for (int e = lane; e < B; e += 64) {
        const int n = e / LL, lm = e - n * LL;
        int l = 0;
        while ((l + 1) * (l + 1) <= lm) ++l;
        const T w = Ws[e];
        const T r = Rq[n * L1 + l];
        const T dr = dRq[n * L1 + l];
        A += w * (dr * fc + r * dfc) * Yq[lm];
        const T wr = w * r;
        Bx += wr * dYq[lm];
        By += wr * dYq[LL + lm];
        Bz += wr * dYq[2 * LL + lm];
      }
      for (int off = 32; off > 0; off >>= 1) {
        A += __shfl_down(A, off, 64);
        Bx += __shfl_down(Bx, off, 64);
        By += __shfl_down(By, off, 64);
        Bz += __shfl_down(Bz, off, 64);
      }
      if (lane == 0) {
        const double d = distsh[q];
        const double sc = double(fc) / d;
        const double g[3] = {double(A) * double(dirsh[3 * q]) + sc * double(Bx),
                             double(A) * double(dirsh[3 * q + 1]) + sc * double(By),
                             double(A) * double(dirsh[3 * q + 2]) + sc * double(Bz)};
        const int64_t p = p0 + 1 + q0 + q;
        const int j = pv.pair_j[p];
        atomicAdd(grad + 3 * j, g[0]);
        atomicAdd(grad + 3 * j + 1, g[1]);
        atomicAdd(grad + 3 * j + 2, g[2]);
        if (j != int(i)) {
          atomicAdd(grad + 3 * i, -g[0]);
          atomicAdd(grad + 3 * i + 1, -g[1]);
          atomicAdd(grad + 3 * i + 2, -g[2]);
        }
        // r_ji = r_i - r_j(image) = -pair vector (sparse_kernel_predict.hh:354-364)
        const double rji[3] = {-pv.pair_vec[3 * p], -pv.pair_vec[3 * p + 1], -pv.pair_vec[3 * p + 2]};
        v6[0] += rji[0] * g[0];
        v6[4] += rji[2] * g[0];
        v6[1] += rji[1] * g[1];
        v6[5] += rji[0] * g[1];
        v6[2] += rji[2] * g[2];
        v6[3] += rji[1] * g[2];
      }
    }
    __syncthreads();
  }
  if (lane == 0)
    for (int k = 0; k < 6; ++k) vir[wave][k] = v6[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    double t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += vir[w][threadIdx.x];
    centre_virial[6 * i + threadIdx.x] = t;
  }
}

// per-structure (segment) sums of per-centre rows of width w <= 8 (+=)
__global__ void segment_sum_kernel(const double* __restrict__ in, int w,
                                   const int64_t* __restrict__ seg_off, int n_seg,
                                   double* __restrict__ out) {
  const int s = blockIdx.x;
  if (s >= n_seg) return;
  __shared__ double red[kThreads / 64][8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t a = seg_off[s] + threadIdx.x; a < seg_off[s + 1]; a += kThreads)
    for (int k = 0; k < w; ++k) v[k] += in[int64_t(w) * a + k];
  for (int k = 0; k < w; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < w; ++k) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < w) {
    double t = 0;
    for (int x = 0; x < kThreads / 64; ++x) t += red[x][threadIdx.x];
    out[int64_t(w) * s + threadIdx.x] += t;
  }
}

// raw GTO radial integrals (FP64 Hyp1f1) at points r: out[pt][n][l]
__global__ void radial_direct_kernel(RadialDev rad, const double* __restrict__ r, int64_t n_pts,
                                     double* __restrict__ out, double* __restrict__ dout) {
  const int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (it >= n_pts * rad.N) return;
  const int64_t pt = it / rad.N;
  const int n = int(it - pt * rad.N);
  double v[kMaxL1], d[kMaxL1];
  gto_radial_n(rad, n, r[pt], v, dout ? d : nullptr);
  const int64_t o = (pt * rad.N + n) * rad.L1;
  for (int l = 0; l < rad.L1; ++l) {
    out[o + l] = v[l];
    if (dout) dout[o + l] = d[l];
  }
}

// out[pt][n][l] = sum_m O[m][n] in[pt][m][l]
__global__ void radial_finalise_kernel(RadialDev rad, const double* __restrict__ in, int64_t n_pts,
                                       double* __restrict__ out) {
  const int F = rad.N * rad.L1;
  const int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (it >= n_pts * F) return;
  const int64_t pt = it / F;
  const int f = int(it - pt * F), n = f / rad.L1, l = f - n * rad.L1;
  double s = 0;
  for (int m = 0; m < rad.N; ++m) s += rad.O[m * rad.N + n] * in[(pt * rad.N + m) * rad.L1 + l];
  out[it] = s;
}

// neighbour radial contribution as used by the expansion (spline or raw)
template <class T>
__global__ void radial_eval_kernel(RadialDev rad, const double* __restrict__ r, int64_t n_pts,
                                   double* __restrict__ R, double* __restrict__ dR) {
  const int F = rad.N * rad.L1;
  const int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (it >= n_pts * F) return;
  const int64_t pt = it / F;
  const int f = int(it - pt * F);
  SplineCell<T> cell = spline_cell<T>(rad, r[pt]);
  T v, d;
  spline_eval<T>(rad, cell, f, &v, dR ? &d : nullptr);
  R[it] = double(v);
  if (dR) dR[it] = double(d);
}

template <class T>
__global__ __launch_bounds__(kThreads) void ylm_eval_kernel(int L, const double* __restrict__ dirs,
                                                            int64_t n, double* __restrict__ Y,
                                                            double* __restrict__ dY) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L1 = L + 1, LL = L1 * L1;
  const int per_wave = 64 / L1, per_block = 4 * per_wave;
  T* dsh = reinterpret_cast<T*>(smem_raw);
  T* Ysh = dsh + 3 * per_block;
  T* dYsh = Ysh + per_block * LL;
  const int64_t q0 = int64_t(blockIdx.x) * per_block;
  const int nb = int(min<int64_t>(per_block, n - q0));
  for (int t = threadIdx.x; t < 3 * nb; t += kThreads) dsh[t] = T(dirs[3 * q0 + t]);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  ylm_wave<T>(L, dsh, wave * per_wave, min(nb, (wave + 1) * per_wave), Ysh, dY ? dYsh : nullptr, LL,
              3 * LL);
  __syncthreads();
  for (int t = threadIdx.x; t < nb * LL; t += kThreads) Y[q0 * LL + t] = double(Ysh[t]);
  if (dY)
    for (int t = threadIdx.x; t < nb * 3 * LL; t += kThreads) dY[q0 * 3 * LL + t] = double(dYsh[t]);
}

template <class T, int KPT>
void launch_expansion_kpt(const RadialDev& rad, int S, const PairView& pv, T* c,
                          unsigned long long* keymask, hipStream_t stream) {
  const int F = rad.N * rad.L1;
  size_t smem = sizeof(T) * (kPB * rad.LL + kPB * F + 4 * kPB) + sizeof(double) * kPB + sizeof(int) * kPB;
  size_t fin = sizeof(T) * rad.N * rad.LL;
  if (fin > smem) smem = fin;
  hipLaunchKernelGGL((expansion_kernel<T, KPT>), dim3(unsigned(pv.n_centres)), dim3(kThreads), smem,
                     stream, rad, S, pv, c, keymask);
  RGPU_LAUNCH_CHECK();
}

template <class T, int KPT>
void launch_backward_dc_kpt(const RadialDev& rad, int S, int64_t n, const T* c, const T* p,
                            const double* norm2, bool normalize, const T* G, T* W, hipStream_t stream) {
  size_t smem = sizeof(T) * S * rad.N * rad.LL;
  hipLaunchKernelGGL((backward_dc_kernel<T, KPT>), dim3(unsigned(n)), dim3(kThreads), smem, stream, rad,
                     S, c, p, norm2, int(normalize), G, W);
  RGPU_LAUNCH_CHECK();
}

}  // namespace

template <class T>
void launch_expansion(const RadialDev& rad, int S, const PairView& pv, T* c,
                      unsigned long long* keymask, hipStream_t stream) {
  if (pv.n_centres == 0) return;
  const int E = S * rad.N * rad.LL;
  const int kpt = (E + kThreads - 1) / kThreads;
  if (kpt <= 4)
    launch_expansion_kpt<T, 4>(rad, S, pv, c, keymask, stream);
  else if (kpt <= 8)
    launch_expansion_kpt<T, 8>(rad, S, pv, c, keymask, stream);
  else if (kpt <= 16)
    launch_expansion_kpt<T, 16>(rad, S, pv, c, keymask, stream);
  else if (kpt <= 24)
    launch_expansion_kpt<T, 24>(rad, S, pv, c, keymask, stream);
  else if (kpt <= 32)
    launch_expansion_kpt<T, 32>(rad, S, pv, c, keymask, stream);
  else if (kpt <= 64)
    launch_expansion_kpt<T, 64>(rad, S, pv, c, keymask, stream);
  else
    throw InvalidArgument("species x max_radial x (max_angular+1)^2 too large (> 16384)");
}

template <class T>
void launch_power_spectrum(const RadialDev& rad, int S, int64_t n, const T* c, T* p, double* norm2,
                           hipStream_t stream) {
  if (n == 0) return;
  size_t smem = sizeof(T) * S * rad.N * rad.LL;
  if (smem > 64 * 1024) throw InvalidArgument("expansion block too large for the power spectrum");
  hipLaunchKernelGGL(power_spectrum_kernel<T>, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad.N,
                     rad.L, S, c, p, norm2);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_normalise(int64_t n, int64_t D, const T* p, const double* norm2, bool normalize, double* out,
                      hipStream_t stream) {
  if (n == 0) return;
  int64_t total = n * D;
  unsigned blocks = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(normalise_kernel<T>, dim3(blocks), dim3(256), 0, stream, n, D, p, norm2,
                     int(normalize), out);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_backward(const RadialDev& rad, int S, const PairView& pv, const int* /*atom_struct*/,
                     const T* c, const T* p, const double* norm2, bool normalize, const T* G, T* W,
                     double* grad, double* centre_virial, hipStream_t stream) {
  const int64_t n = pv.n_centres;
  if (n == 0) return;
  const int E = S * rad.N * rad.LL;
  const int kpt = (E + kThreads - 1) / kThreads;
  if (kpt <= 4)
    launch_backward_dc_kpt<T, 4>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else if (kpt <= 8)
    launch_backward_dc_kpt<T, 8>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else if (kpt <= 16)
    launch_backward_dc_kpt<T, 16>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else if (kpt <= 24)
    launch_backward_dc_kpt<T, 24>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else if (kpt <= 32)
    launch_backward_dc_kpt<T, 32>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else if (kpt <= 64)
    launch_backward_dc_kpt<T, 64>(rad, S, n, c, p, norm2, normalize, G, W, stream);
  else
    throw InvalidArgument("species x max_radial x (max_angular+1)^2 too large (> 16384)");
  const int F = rad.N * rad.L1;
  size_t smem = sizeof(T) * (kPBB * rad.LL * 4 + 2 * kPBB * F + 3 * kPBB) + sizeof(double) * 3 * kPBB +
                sizeof(int) * kPBB;
  hipLaunchKernelGGL(backward_pairs_kernel<T>, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad, S, pv,
                     W, grad, centre_virial);
  RGPU_LAUNCH_CHECK();
}

void launch_segment_sum(const double* in, int w, const int64_t* seg_off, int n_seg, double* out,
                        hipStream_t stream) {
  if (n_seg == 0) return;
  hipLaunchKernelGGL(segment_sum_kernel, dim3(unsigned(n_seg)), dim3(kThreads), 0, stream, in, w, seg_off,
                     n_seg, out);
  RGPU_LAUNCH_CHECK();
}

void launch_radial_direct(const RadialDev& rad, const double* r, int64_t n_pts, double* raw, double* draw,
                          hipStream_t stream) {
  if (n_pts == 0) return;
  int64_t items = n_pts * rad.N;
  hipLaunchKernelGGL(radial_direct_kernel, dim3(unsigned((items + 127) / 128)), dim3(128), 0, stream, rad, r,
                     n_pts, raw, draw);
  RGPU_LAUNCH_CHECK();
}

void launch_radial_finalise(const RadialDev& rad, const double* raw, int64_t n_pts, double* out,
                            hipStream_t stream) {
  if (n_pts == 0) return;
  int64_t items = n_pts * rad.N * rad.L1;
  hipLaunchKernelGGL(radial_finalise_kernel, dim3(unsigned((items + 255) / 256)), dim3(256), 0, stream,
                     rad, raw, n_pts, out);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_radial_spline_eval(const RadialDev& rad, const double* r, int64_t n_pts, double* R, double* dR,
                               hipStream_t stream) {
  if (n_pts == 0) return;
  int64_t items = n_pts * rad.N * rad.L1;
  hipLaunchKernelGGL(radial_eval_kernel<T>, dim3(unsigned((items + 255) / 256)), dim3(256), 0, stream,
                     rad, r, n_pts, R, dR);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_ylm_eval(int L, const double* dirs, int64_t n, double* Y, double* dY, hipStream_t stream) {
  if (n == 0) return;
  const int L1 = L + 1, LL = L1 * L1, per_block = 4 * (64 / L1);
  size_t smem = sizeof(T) * (3 * per_block + per_block * LL + per_block * 3 * LL);
  hipLaunchKernelGGL(ylm_eval_kernel<T>, dim3(unsigned((n + per_block - 1) / per_block)), dim3(kThreads),
                     smem, stream, L, dirs, n, Y, dY);
  RGPU_LAUNCH_CHECK();
}

template void launch_radial_spline_eval<double>(const RadialDev&, const double*, int64_t, double*, double*,
                                                hipStream_t);
template void launch_radial_spline_eval<float>(const RadialDev&, const double*, int64_t, double*, double*,
                                               hipStream_t);
template void launch_ylm_eval<double>(int, const double*, int64_t, double*, double*, hipStream_t);
template void launch_ylm_eval<float>(int, const double*, int64_t, double*, double*, hipStream_t);

template void launch_expansion<double>(const RadialDev&, int, const PairView&, double*,
                                       unsigned long long*, hipStream_t);
template void launch_expansion<float>(const RadialDev&, int, const PairView&, float*,
                                      unsigned long long*, hipStream_t);
template void launch_power_spectrum<double>(const RadialDev&, int, int64_t, const double*, double*,
                                            double*, hipStream_t);
template void launch_power_spectrum<float>(const RadialDev&, int, int64_t, const float*, float*,
                                           double*, hipStream_t);
template void launch_normalise<double>(int64_t, int64_t, const double*, const double*, bool, double*,
                                       hipStream_t);
template void launch_normalise<float>(int64_t, int64_t, const float*, const double*, bool, double*,
                                      hipStream_t);
template void launch_backward<double>(const RadialDev&, int, const PairView&, const int*, const double*,
                                      const double*, const double*, bool, const double*, double*,
                                      double*, double*, hipStream_t);
template void launch_backward<float>(const RadialDev&, int, const PairView&, const int*, const float*,
                                     const float*, const double*, bool, const float*, float*, double*,
                                     double*, hipStream_t);

}  // namespace rgpu
