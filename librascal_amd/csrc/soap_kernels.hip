// SOAP expansion, power spectrum and their adjoint on gfx950.
//
// One workgroup (256 threads, 4 wavefronts) owns one centre.  Its pairs are
// staged in LDS in batches: per pair the real spherical harmonics come from
// a wavefront-level recursion (one lane per m, lockstep over l,
// neighbouring-m values by lane shuffles, recursion coefficients from a
// table in LDS) and the radial functions from the spline table (or FP64
// Hyp1f1 followed by the O^T orthonormalisation).  Pairs of a batch are
// grouped by neighbour species with wavefront ballots, so the accumulation
// walks one species block at a time and each thread keeps only a few
// (n, lm) accumulators in registers.  The result equals the reference's
// per-pair outer-product accumulation (calculator_spherical_expansion.hh
// :2234-2410) up to rounding.
//
// The adjoint keeps dL/dc of the centre in LDS: it is formed key block by
// key block from dL/dp (power-spectrum transpose, register-tiled 4 x 1,
// next key block prefetched into registers), then the same workgroup
// recomputes Y_lm, dY_lm, R and dR per pair and contracts them with dL/dc,
// so the per-pair gradient tensor is never materialised.
//
// Index arithmetic avoids integer division by runtime values (tens of VALU
// instructions each on this GPU): quotients come from an FP32 reciprocal
// with one correction step, l of a packed lm from a square root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>
#include <utility>
#include <vector>

#include "soap_device.hh"
#include "packing.hh"
#include "scan.hh"
#include "soap_kernels.hh"

namespace rgpu {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};

// n / d for 0 <= n < 2^22: FP32 reciprocal, then one correction step
struct FastDiv {
  int d;
  float inv;
};
__device__ __forceinline__ FastDiv fast_div(int d) { return FastDiv{d, 1.0f / float(d)}; }
__device__ __forceinline__ int operator/(int n, const FastDiv& f) {
  int q = int(float(n) * f.inv);
  const int r = n - q * f.d;
  q += (r >= f.d) - (r < 0);
  return q;
}

// l of the packed index lm = l*l + l + m (exact for lm < 2^20)
__device__ __forceinline__ int l_of_lm(int lm) { return int(sqrtf(float(lm) + 0.5f)); }

// index (a, b), a <= b, of a species-pair key in the order (0,0), (0,1), ...
__device__ __forceinline__ void key_species(int key, int S, int* a, int* b) {
  int aa = 0, kk = key;
  while (kk >= S - aa) {
    kk -= S - aa;
    ++aa;
  }
  *a = aa;
  *b = aa + kk;
}

__host__ __device__ inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Y_lm coefficient table in LDS (copied from the context's device table)
template <class T>
__host__ __device__ inline size_t ylm_table_bytes(int L1) {
  return align_up(sizeof(T) * size_t(ylm_table_size(L1)), 64);
}
template <class T>
__device__ __forceinline__ const T* ylm_table_src(const RadialDev& R);
template <class T>
__device__ __forceinline__ const YGradTerm<T>* ygrad_src(const RadialDev& R);
template <>
__device__ __forceinline__ const YGradTerm<double>* ygrad_src<double>(const RadialDev& R) {
  return R.ygrad64;
}
template <>
__device__ __forceinline__ const YGradTerm<float>* ygrad_src<float>(const RadialDev& R) {
  return R.ygrad32;
}
template <>
__device__ __forceinline__ const double* ylm_table_src<double>(const RadialDev& R) {
  return R.ylm64;
}
template <>
__device__ __forceinline__ const float* ylm_table_src<float>(const RadialDev& R) {
  return R.ylm32;
}
template <class T>
__device__ void load_ylm_table(const RadialDev& R, T* dst) {
  const T* src = ylm_table_src<T>(R);
  const int n = ylm_table_size(R.L1);
  for (int e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
}

// ---------------------------------------------------------------------------
// Real spherical harmonics of nq pairs, one wavefront segment of (L+1) lanes
// per pair (lane = m >= 0 of its segment; it writes m and -m).  dirs[q][3]
// (unit vectors), writes Y[q][LL] and optionally dY[q][3][LL] (gradient on
// the unit sphere, no 1/r).  Every lane of the wavefront must call this
// (shuffles).  Follows math/spherical_harmonics.cc:135-363: associated
// Legendre recursion with the sqrt(1/(2 pi)) normalisation folded in, m = 0
// scaled by 1/sqrt(2), cos/sin(m phi) by Chebyshev recurrence, phi = 0 on
// the z axis, and the pole/equator branch of the gradient at sin(theta) 0.1.
template <class T, int LC = -1>
__device__ void ylm_wave(int L_, const T* __restrict__ yc, const T* dirs, int q0, int nq, T* Y, T* dY,
                         int ystride, int dystride) {
  typedef typename Vec4<T>::type V4;
  const int L = LC >= 0 ? LC : L_;  // LC: l_max fixed at compile time
  const int L1 = L + 1;
  const int lane = threadIdx.x & 63;
  const int seg = lane / L1;
  const int m = lane - seg * L1;
  const int per_wave = 64 / L1;
  const int q = q0 + seg;
  const bool active = seg < per_wave && q < nq;
  const T* dfac = yc + 4 * L1 * L1;
  const T* s2l1 = dfac + L1;
  const T* sll1 = s2l1 + L1;
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
  for (int k = 1; k <= m; ++k) diag = diag * sin_t * dfac[k];
  const T isq2 = T(0.70710678118654752440);
  const bool dy_on = dY != nullptr;
  const T inv_sxy = sxy > T(0.1) ? T(1) / sxy : T(0);
  const T inv_ct = sxy > T(0.1) ? T(0) : T(-0.5) / cos_t;
  T P1 = T(0), P2 = T(0);
  T* Yq = Y + q * ystride;
  T* dYq = dy_on ? dY + q * dystride : nullptr;
  const int LL = L1 * L1;
  for (int l = 0; l <= L; ++l) {
    const V4 c4 = reinterpret_cast<const V4*>(yc)[l * L1 + m];  // a, b, f_k, f_k+1
    T P = T(0);
    if (m == l) {
      P = diag;
    } else if (m == l - 1) {
      P = (l == 1) ? cos_t * T(1.73205080756887729353) * s2pi : diag * cos_t * s2l1[l];
    } else if (m < l - 1) {
      T t = cos_t * P1;
      t += c4.y * P2;
      t *= c4.x;
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
    if (dy_on) {
      const T Pn1 = __shfl(P, lane + 1, 64);
      const T Pn2 = __shfl(P, lane + 2, 64);
      if (active && l >= 1 && m <= l) {
        if (m == 0) {
          const T qv = sll1[l] * isq2 * Pn1;
          dYq[base] = cos_t * cos_p * qv;
          dYq[LL + base] = cos_t * sin_p * qv;
          dYq[2 * LL + base] = -sxy * qv;
        }
        if (m < l) {
          const int M = m + 1;
          const T fk = c4.z, fk1 = c4.w;
          const T pk2 = (m + 2 <= l) ? Pn2 : T(0);
          const T lpd = fk * P - fk1 * pk2;
          const T pdf = sxy > T(0.1) ? T(m + 1) * inv_sxy * Pn1 : inv_ct * (fk * P + fk1 * pk2);
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
  if (dy_on && active && m == 0) {
    dYq[0] = T(0);
    dYq[LL] = T(0);
    dYq[2 * LL] = T(0);
  }
}

// ---------------------------------------------------------------------------
// LDS layout of one batch of PB pairs (FP64 values first for alignment).
template <class T>
struct PairBatch {
  double* dist;  // [PB]
  double* fc;    // [PB] cutoff function
  double* dfc;   // [PB] its derivative (adjoint)
  double* vec;   // [PB][3] pair vectors (adjoint)
  double* raw;   // [PB][F] raw radial integrals (direct Hyp1f1)
  double* draw;  // [PB][F] their derivatives (direct, adjoint)
  YGradTerm<T>* gt;  // [LL][3] gradient terms of Y_lm (adjoint)
  T* Y;          // [PB][LL]
  T* dY;         // [PB][3][LL] (not staged: the adjoint uses gt)
  T* R;          // [PB][F] finalised radial functions (x fc in the forward)
  T* dR;         // [PB][F] (adjoint)
  T* dir;        // [PB][3]
  T* ca;         // [PB] spline weights of the pair's cell
  T* cb;         // [PB]
  T* ih;         // [PB] inverse grid spacing of the pair's table
  T* res;        // [PB][4] per-pair contractions (adjoint)
  int* sp;       // [PB] neighbour species
  int* lst;      // [PB] batch positions grouped by species (forward)
  int* klo;      // [PB] spline cell
  int* pj;       // [PB] neighbour atom (adjoint)
};

__host__ __device__ inline size_t pair_batch_bytes(int PB, int LL, int F, int tsize, bool direct,
                                                   bool adjoint) {
  // the adjoint's gradient-term table {T c0, c1; int i0, i1}[LL][3] first
  size_t b = adjoint ? align_up(size_t(LL) * 3 * (2 * size_t(tsize) + 8), 64) : 0;
  b += 8 * size_t(PB) * (adjoint ? 6 : 2);
  if (direct) b += 8 * size_t(PB) * F * (adjoint ? 2 : 1);
  b += size_t(tsize) * PB * (size_t(LL) + size_t(F) * (adjoint ? 2 : 1) + 6 + (adjoint ? 4 : 0));
  b = align_up(b, 8);
  b += 4 * size_t(PB) * (adjoint ? 4 : 3);
  return align_up(b, 64);
}

template <class T, bool DIRECT, bool ADJ>
__device__ PairBatch<T> carve_batch(unsigned char* base, int PB, int LL, int F) {
  PairBatch<T> pb;
  pb.gt = nullptr;
  if (ADJ) {
    pb.gt = reinterpret_cast<YGradTerm<T>*>(base);
    base += align_up(size_t(LL) * 3 * sizeof(YGradTerm<T>), 64);
  }
  double* d = reinterpret_cast<double*>(base);
  pb.dist = d;
  d += PB;
  pb.fc = d;
  d += PB;
  pb.dfc = pb.vec = nullptr;
  if (ADJ) {
    pb.dfc = d;
    d += PB;
    pb.vec = d;
    d += 3 * PB;
  }
  pb.raw = pb.draw = nullptr;
  if (DIRECT) {
    pb.raw = d;
    d += size_t(PB) * F;
    if (ADJ) {
      pb.draw = d;
      d += size_t(PB) * F;
    }
  }
  T* t = reinterpret_cast<T*>(d);
  pb.Y = t;
  t += PB * LL;
  pb.dY = nullptr;
  pb.R = t;
  t += PB * F;
  pb.dR = nullptr;
  if (ADJ) {
    pb.dR = t;
    t += PB * F;
  }
  pb.dir = t;
  t += 3 * PB;
  pb.ca = t;
  t += PB;
  pb.cb = t;
  t += PB;
  pb.ih = t;
  t += PB;
  pb.res = nullptr;
  if (ADJ) {
    pb.res = t;
    t += 4 * PB;
  }
  const size_t off = align_up(size_t(reinterpret_cast<unsigned char*>(t) - base), 8);
  pb.sp = reinterpret_cast<int*>(base + off);
  pb.lst = pb.sp + PB;
  pb.klo = pb.lst + PB;
  pb.pj = ADJ ? pb.klo + PB : nullptr;
  return pb;
}

// Inputs of one batch of pairs held in registers (thread t holds pair t of
// the batch), fetched a batch ahead so that their global latency overlaps
// the work on the previous batch.
struct PairPre {
  double d, vx, vy, vz, fc, dfc, xhi, h;
  int j, sp, klo;
};

template <bool DIRECT, bool ADJ>
__device__ __forceinline__ void pair_fetch(const RadialDev& rad, const PairView& pv, int64_t pbase,
                                           int nb, PairPre& f) {
  if (int(threadIdx.x) < nb) {
    const int64_t p = pbase + threadIdx.x;
    f.d = pv.pair_dist[p];
    f.vx = pv.pair_vec[3 * p];
    f.vy = pv.pair_vec[3 * p + 1];
    f.vz = pv.pair_vec[3 * p + 2];
    f.fc = pv.pair_fc[p];
    if (ADJ) f.dfc = pv.pair_dfc[p];
    f.j = pv.pair_j[p];
    f.sp = pv.species_of[f.j];
    if (!DIRECT) {
      // spline cell as in spline_cell(); with one table per species, on
      // the neighbour species' own grid, its cells numbered on from the
      // tables before it
      if (rad.sp_tab) {
        const int t = rad.sp_tab[f.sp], last = rad.tab_n[t] - 2;
        int k = int((f.d - rad.x1) * rad.tab_pu[t]) - 1;
        k = k < 0 ? 0 : (k > last ? last : k);
        f.xhi = rad.tab_grid[rad.tab_grid0[t] + k + 1];
        f.h = rad.tab_h[t];
        f.klo = rad.tab_cell0[t] + k;
      } else {
        int k = int((f.d - rad.x1) * rad.per_unit) - 1;
        k = k < 0 ? 0 : (k > rad.search_size ? rad.search_size : k);
        f.xhi = rad.grid[k + 1];
        f.h = rad.h;
        f.klo = k;
      }
    }
  }
}

// Stages a batch of nb pairs (inputs fetched by pair_fetch): distances, directions,
// cutoff values, neighbour species, Y_lm (and gradients) and the finalised
// radial functions (and derivatives); in the forward R is multiplied by fc.
// When lists are requested (forward), wavefront 0 groups the batch by
// species: s_count[s] pairs listed from lst[s_start[s]].  Wavefronts with
// no Y_lm work take the radial functions.  All threads call; returns after
// a barrier.
template <class T, bool DIRECT, bool ADJ, int NC = -1, int LC = -1>
__device__ void stage_pairs(const RadialDev& rad, const T* yc, const PairPre& f, int nb,
                            PairBatch<T>& b, int* s_start, int* s_count,
                            unsigned long long* kmask) {
  // NC, LC: n_max and l_max fixed at compile time (or -1)
  const int N = NC >= 0 ? NC : rad.N, L = LC >= 0 ? LC : rad.L, L1 = L + 1, LL = L1 * L1, F = N * L1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < nb) {
    const double d = f.d;
    b.dist[tid] = d;
    b.dir[3 * tid] = T(f.vx / d);
    b.dir[3 * tid + 1] = T(f.vy / d);
    b.dir[3 * tid + 2] = T(f.vz / d);
    b.fc[tid] = f.fc;
    if (ADJ) {
      b.dfc[tid] = f.dfc;
      b.vec[3 * tid] = f.vx;
      b.vec[3 * tid + 1] = f.vy;
      b.vec[3 * tid + 2] = f.vz;
      b.pj[tid] = f.j;
    }
    b.sp[tid] = f.sp;
    if (!DIRECT) {
      const double a = (f.xhi - d) / f.h;
      b.klo[tid] = f.klo;
      b.ca[tid] = T(a);
      b.cb[tid] = T(1 - a);
      if (ADJ && rad.sp_tab) b.ih[tid] = T(1.0 / f.h);
    }
  }
  __syncthreads();
  if (s_count && wave == 0) {
    const bool act = lane < nb;
    const int s = act ? b.sp[lane] : -1;
    unsigned long long rem = __ballot(act);
    int base = 0;
    while (rem) {
      const int leader = __ffsll((long long)rem) - 1;
      const int sl = __shfl(s, leader, 64);
      const unsigned long long m = __ballot(act && s == sl);
      const int cnt = __popcll(m);
      if (act && s == sl) b.lst[base + __popcll(m & ((1ull << lane) - 1ull))] = lane;
      if (lane == 0) {
        s_start[sl] = base;
        s_count[sl] = cnt;
        *kmask |= 1ull << sl;
      }
      base += cnt;
      rem &= ~m;
    }
  }
  const int per_wave = 64 / L1;
  for (int qb = wave * per_wave; qb < nb; qb += kWaves * per_wave)
    ylm_wave<T, LC>(L, yc, b.dir, qb, nb, b.Y, ADJ ? b.dY : nullptr, LL, 3 * LL);
  // radial functions on the wavefronts left idle by the Y_lm pass
  const int ywaves = (nb + per_wave - 1) / per_wave;
  const int r0 = ywaves >= kWaves ? 0 : ywaves * 64;
  const int rstride = kThreads - r0;
  if (!DIRECT) {
    if (tid >= r0) {
      const FastDiv fF = fast_div(F), fL1 = fast_div(L1);
      // one table: a single grid step (per-species tables keep theirs per pair)
      const T inv_h = T(1.0 / rad.h);
      const bool per_table = rad.sp_tab != nullptr;
#pragma unroll 4
      for (int it = tid - r0; it < nb * F; it += rstride) {
        const int q = it / fF, f = it - q * F;
        SplineCell<T> cell;
        cell.klo = b.klo[q];
        cell.a = b.ca[q];
        cell.b = b.cb[q];
        cell.inv_h = per_table ? b.ih[q] : inv_h;
        T v, dv;
        spline_eval<T>(rad, cell, f, &v, ADJ ? &dv : nullptr);
        if (ADJ) {
          // the adjoint keeps R[q][l][n]: one lane's radial channels are
          // contiguous
          const int n = f / fL1, o = q * F + (f - n * L1) * N + n;
          b.R[o] = v;
          b.dR[o] = dv;
        } else {
          b.R[it] = v * T(b.fc[q]);
        }
      }
    }
  } else {
    for (int it = tid; it < nb * N; it += kThreads) {
      const int q = it / N, n = it - q * N;
      double v[kMaxL1], dv[kMaxL1];
      radial_n(rad, n, b.dist[q], v, ADJ ? dv : nullptr);
      for (int l = 0; l < L1; ++l) {
        b.raw[q * F + n * L1 + l] = v[l];
        if (ADJ) b.draw[q * F + n * L1 + l] = dv[l];
      }
    }
    __syncthreads();
    // O^T R per pair (calculator_spherical_expansion.hh:605-618 applied
    // before the accumulation, which is linear)
    for (int it = tid; it < nb * F; it += kThreads) {
      const int q = it / F, f = it - q * F, n = f / L1, l = f - n * L1;
      double s = 0, ds = 0;
      for (int m = 0; m < N; ++m) {
        const double o = rad.O[m * N + n];
        s += o * b.raw[q * F + m * L1 + l];
        if (ADJ) ds += o * b.draw[q * F + m * L1 + l];
      }
      if (ADJ) {
        b.R[q * F + l * N + n] = T(s);
        b.dR[q * F + l * N + n] = T(ds);
      } else {
        b.R[it] = T(s * b.fc[q]);
      }
    }
  }
  __syncthreads();
}

// register budget of the forward: small-KPT spline kernels are held to
// 4-5 waves per SIMD (their LDS allows that many workgroups per CU)
template <class T, int KPT, bool DIRECT>
constexpr int expansion_min_waves() {
  return (!DIRECT && KPT <= 8) ? (sizeof(T) == 4 ? 5 : 4) : 1;
}

// ---------------------------------------------------------------------------
// Expansion forward: c[i][s][n][lm] = [s == s_i][lm == 0] R0[n] / sqrt(4 pi)
//   + sum over pairs (i, j), s_j = s, of fc(r) R_nl(r) Y_lm(r_hat)
// KPT = entries of one species block per thread; NC, LC: n_max and l_max
// fixed at compile time (or -1).
template <class T, int KPT, bool DIRECT, int NC = -1, int LC = -1>
__global__ __launch_bounds__(kThreads, (expansion_min_waves<T, KPT, DIRECT>())) void expansion_kernel(
    RadialDev rad, int S, int PB, PairView pv, T* __restrict__ cout,
    unsigned long long* __restrict__ keymask) {
  const int64_t i = blockIdx.x;
  const int N = NC >= 0 ? NC : rad.N, L1 = LC >= 0 ? LC + 1 : rad.L1, LL = L1 * L1, F = N * L1, B = N * LL;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  PairBatch<T> pb = carve_batch<T, DIRECT, false>(smem_raw + ylm_table_bytes<T>(L1), PB, LL, F);
  __shared__ int s_start[kMaxSpecies], s_count[kMaxSpecies];
  __shared__ unsigned long long kmask;
  const int tid = threadIdx.x;
  const int64_t p0 = pv.centre_off[i] + 1;  // the self pair contributes nothing here
  const int np = int(pv.centre_off[i + 1] - p0);
  const int si = pv.species_of[i];
  const double inv_s4pi = 0.28209479177387814347;  // 1/sqrt(4 pi)
  load_ylm_table<T>(rad, yc);

  const FastDiv fLL = fast_div(LL);
  // entry k of this thread: e = tid + 256 k < B, (n, lm) = (e / LL, e % LL);
  // its R and Y offsets within a pair's rows (entries past B read the
  // first ones, so that the pair loop has no branch, and are never stored)
  int offR[KPT], offY[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int e = tid + kThreads * k;
    offR[k] = 0;
    offY[k] = 0;
    if (e < B) {
      const int n = e / fLL, lm = e - n * LL;
      offR[k] = n * L1 + l_of_lm(lm);
      offY[k] = lm;
    }
  }
  if (tid == 0) kmask = 1ull << si;
  T* ci = cout + i * int64_t(S) * B;

  PairPre pre;
  pair_fetch<DIRECT, false>(rad, pv, p0, max(0, min(PB, np)), pre);
  for (int q0 = 0; q0 == 0 || q0 < np; q0 += PB) {
    const int nb = max(0, min(PB, np - q0));
    if (tid < S) s_count[tid] = 0;
    stage_pairs<T, DIRECT, false, NC, LC>(rad, yc, pre, nb, pb, s_start, s_count, &kmask);
    // next batch's inputs in flight during this batch's accumulation
    if (q0 + PB < np) pair_fetch<DIRECT, false>(rad, pv, p0 + q0 + PB, min(PB, np - q0 - PB), pre);
    for (int s = 0; s < S; ++s) {
      const int cnt = s_count[s];
      if (q0 > 0 && cnt == 0) continue;
      const int st = s_start[s];
      T* cs = ci + int64_t(s) * B;
      T acc[KPT];
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const int e = tid + kThreads * k;
        acc[k] = T(0);
        if (e < B) {
          if (q0 > 0)
            acc[k] = cs[e];
          else if (s == si && offY[k] == 0)
            acc[k] = T(rad.center[(rad.sp_tab ? rad.sp_tab[si] * rad.N : 0) + e / fLL] * inv_s4pi);
        }
      }
      for (int t = 0; t < cnt; ++t) {
        const int q = pb.lst[st + t];
        const T* Yq = pb.Y + q * LL;
        const T* Rq = pb.R + q * F;
#pragma unroll
        for (int k = 0; k < KPT; ++k) acc[k] += Rq[offR[k]] * Yq[offY[k]];
      }
#pragma unroll
      for (int k = 0; k < KPT; ++k)
        if (tid + kThreads * k < B) cs[tid + kThreads * k] = acc[k];
    }
    __syncthreads();
  }
  if (tid == 0) keymask[i] = kmask;
}

// ---------------------------------------------------------------------------
// Power spectrum: p^{ab}[n1 N + n2][l] = (2l+1)^{-1/2} (sqrt 2 if a<b)
// sum_m c_a[n1,lm] c_b[n2,lm] (calculator_spherical_invariants.hh:747-788),
// 4x4 register tiles over (n1, n2) per (key, l); KC keys at a time are
// staged in LDS and written out contiguously.  The expansion rows sit in
// LDS with an odd stride (no bank conflicts between the tile rows), followed
// by three zero rows, so that a tile reads its four rows unconditionally
// (rows past N of a species read the next species' rows: those products
// are never stored).  Work items run l-major, so the lanes of a wavefront
// share one or two l and loop over the same number of m.
// Un-normalised; |p|^2 is returned per centre.
// key_ab[k] = a << 8 | b of key k; tri[t] = n1 << 8 | n2 of the packed
// diagonal pair t (packed planes only)
template <bool SPLIT>
__device__ void ps_tables(int N, int S, int* key_ab, int* tri) {
  const int K = S * (S + 1) / 2;
  for (int k = threadIdx.x; k < K; k += kThreads) {
    int a, b;
    key_species(k, S, &a, &b);
    key_ab[k] = (a << 8) | b;
  }
  if (SPLIT)
    for (int t = threadIdx.x; t < N * (N + 1) / 2; t += kThreads) {
      int n1, n2;
      packed_pair(t, N, &n1, &n2);
      tri[t] = (n1 << 8) | n2;
    }
}

// The power spectrum of centre i from its expansion rows in LDS (cs: S N
// rows of LS = (L+1)^2 | 1 values, then three zero rows; key_ab and tri
// filled), in KC-key passes over the stage buffer; writes the FP64/FP32
// rows or the BF16 operand planes and norm2[i].  Every thread calls; ends
// after a barrier.
template <class T, bool SPLIT, int NC = -1, int LC = -1>
__device__ __forceinline__ void ps_body(int N_, int L_, int S, int KC, int64_t Dq, int64_t i, const T* cs, T* stage,
                        const int* key_ab, const int* tri, T* __restrict__ p,
                        unsigned short* __restrict__ p_hi, unsigned short* __restrict__ p_lo,
                        double* __restrict__ norm2) {
  // NC, LC: N and L fixed at compile time (the LDS offsets of a tile's rows
  // and stage entries become immediates), or -1 for the runtime values
  const int N = NC >= 0 ? NC : N_, L = LC >= 0 ? LC : L_;
  const int L1 = L + 1, LS = (L1 * L1) | 1;
  const int K = S * (S + 1) / 2;
  const int NT = (N + 3) / 4;
  const int KB = N * N * L1;  // one key block
  __shared__ double red[kWaves];
  const int64_t pbase = i * int64_t(K) * KB;
  const FastDiv fNT = fast_div(NT), fL1 = fast_div(L1), fT = fast_div(NT * NT);
  double ss = 0;
  // packed diagonal value o of a key block: its offset in the full block,
  // bit 30 set for an n1 < n2 entry (the same for every diagonal key; the
  // four of this thread's first quad are kept in registers)
  auto diag_src = [&](int o) {
    const int t = o / fL1, l = o - t * L1;
    const int n1 = tri[t] >> 8, n2 = tri[t] & 255;
    return ((n1 * N + n2) * L1 + l) | (n1 < n2 ? 1 << 30 : 0);
  };
  int dsrc[4] = {0, 0, 0, 0};
  if (SPLIT) {
    const int KDp = packed_diag_block(N, L1);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (4 * int(threadIdx.x) + u < KDp) dsrc[u] = diag_src(4 * int(threadIdx.x) + u);
  }
  for (int k0 = 0; k0 < K; k0 += KC) {
    const int kc = min(KC, K - k0);
    const int per_l = kc * NT * NT;
    const FastDiv fPL = fast_div(per_l);
    const int items = L1 * per_l;
    for (int it = threadIdx.x; it < items; it += kThreads) {
      const int l = it / fPL;
      const int r = it - l * per_l;
      const int kk = r / fT;
      const int t = r - kk * (NT * NT);
      const int t1 = t / fNT;
      const int t2 = t - t1 * NT;
      const int ab = key_ab[k0 + kk];
      const int a = ab >> 8, b = ab & 255;
      // the packed output of a diagonal key needs n1 <= n2 only
      if (SPLIT && a == b && t1 > t2) continue;
      const T* ca = cs + a * N * LS;
      const T* cb = cs + b * N * LS;
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
          va[u] = ca[(n1 + u) * LS + lm];
          vb[u] = cb[(n2 + u) * LS + lm];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] += va[u] * vb[v];
      }
      T f = T(1) / sqrt(T(2 * l + 1));
      if (a < b) f *= T(1.41421356237309504880);
      T* sk = stage + kk * KB;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (n1 + u < N && n2 + v < N) sk[((n1 + u) * N + n2 + v) * L1 + l] = acc[u][v] * f;
    }
    __syncthreads();
    if (SPLIT) {
      // packed row (off-diagonal keys whole, diagonal keys n1 <= n2 with the
      // n1 < n2 entries times sqrt 2) as the matrix-core operand planes
      // v = hi + lo (RNE BF16 each); the norm from the full layout
      const int KD = packed_diag_block(N, L1);
      // four values per step where every key block keeps 16-byte stage
      // reads and 8-byte plane stores aligned
      const bool vec4 = ((KB | KD) & 3) == 0;
      for (int kk = 0; kk < kc && vec4; ++kk) {
        const int ab = key_ab[k0 + kk];
        const int a = ab >> 8, b = ab & 255;
        const T* sk = stage + kk * KB;
        const int64_t ob = i * Dq + packed_key_offset(a, b, S, N, L1);
        uint2* h2 = reinterpret_cast<uint2*>(p_hi + ob);
        uint2* l2 = reinterpret_cast<uint2*>(p_lo + ob);
        const int nq = (a < b ? KB : KD) >> 2;
        for (int q = threadIdx.x; q < nq; q += kThreads) {
          float v[4];
          if (a < b) {
            const typename Vec4<T>::type s4 = reinterpret_cast<const typename Vec4<T>::type*>(sk)[q];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              v[u] = float(s4[u]);
              ss += double(v[u]) * double(v[u]);
            }
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              // source of packed value 4 q + u in the key block (bit 30:
              // an n1 < n2 entry); this thread's first quad from registers
              int d = q == int(threadIdx.x) ? dsrc[u] : diag_src(4 * q + u);
              const bool off = d >> 30;
              v[u] = float(sk[d & ((1 << 30) - 1)]);
              ss += (off ? 2.0 : 1.0) * double(v[u]) * double(v[u]);
              if (off) v[u] *= 1.41421356237309504880f;
            }
          }
          unsigned hb[4], lb[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const __bf16 h = (__bf16)v[u];
            hb[u] = __builtin_bit_cast(unsigned short, h);
            lb[u] = __builtin_bit_cast(unsigned short, (__bf16)(v[u] - (float)h));
          }
          h2[q] = make_uint2(hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16));
          l2[q] = make_uint2(lb[0] | (lb[1] << 16), lb[2] | (lb[3] << 16));
        }
      }
      for (int kk = 0; kk < kc && !vec4; ++kk) {
        const int ab = key_ab[k0 + kk];
        const int a = ab >> 8, b = ab & 255;
        const T* sk = stage + kk * KB;
        const int64_t ob = i * Dq + packed_key_offset(a, b, S, N, L1);
        const int nk = a < b ? KB : KD;
        for (int o = threadIdx.x; o < nk; o += kThreads) {
          T v;
          if (a < b) {
            v = sk[o];
            ss += double(v) * double(v);
          } else {
            const int t = o / fL1, l = o - t * L1;
            const int n1 = tri[t] >> 8, n2 = tri[t] & 255;
            v = sk[(n1 * N + n2) * L1 + l];
            ss += (n1 < n2 ? 2.0 : 1.0) * double(v) * double(v);
            if (n1 < n2) v *= T(1.41421356237309504880);
          }
          const __bf16 h = (__bf16)float(v);
          p_hi[ob + o] = __builtin_bit_cast(unsigned short, h);
          p_lo[ob + o] = __builtin_bit_cast(unsigned short, (__bf16)(float(v) - (float)h));
        }
      }
    } else {
      const int64_t ob = pbase + int64_t(k0) * KB;
      for (int o = threadIdx.x; o < kc * KB; o += kThreads) {
        const T v = stage[o];
        p[ob + o] = v;
        ss += double(v) * double(v);
      }
    }
    __syncthreads();
  }
  if (SPLIT) {
    // zero pad columns of the packed row
    const int64_t Dp = packed_feature_size(S, N, L1);
    for (int64_t o = Dp + threadIdx.x; o < Dq; o += kThreads) {
      p_hi[i * Dq + o] = 0;
      p_lo[i * Dq + o] = 0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kWaves; ++w) t += red[w];
    norm2[i] = t;
  }
  __syncthreads();  // red and the stage may be reused
}

// out row r = row rows[r] of p (rows null: row r), divided by its norm when
// normalising

template <class T, bool SPLIT, int NC = -1, int LC = -1>
__global__ __launch_bounds__(kThreads) void power_spectrum_kernel(int N_, int L_, int S, int KC, int64_t Dq,
                                                                  const T* __restrict__ c,
                                                                  T* __restrict__ p,
                                                                  unsigned short* __restrict__ p_hi,
                                                                  unsigned short* __restrict__ p_lo,
                                                                  double* __restrict__ norm2) {
  const int N = NC >= 0 ? NC : N_, L = LC >= 0 ? LC : L_;
  const int64_t i = blockIdx.x;
  const int L1 = L + 1, LL = L1 * L1, B = N * LL, LS = LL | 1;
  const int K = S * (S + 1) / 2;
  const int KB = N * N * L1;  // one key block
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* cs = reinterpret_cast<T*>(smem_raw);  // [S][N][LS], then 3 zero rows
  T* stage = cs + align_up(size_t(S * N + 3) * LS, 4);  // [KC][N][N][L1], 16-byte aligned
  int* key_ab = reinterpret_cast<int*>(stage + KC * KB);  // [K] a << 8 | b
  int* tri = key_ab + K;  // [N(N+1)/2] n1 << 8 | n2 of the packed diagonal pairs
  const T* ci = c + i * int64_t(S) * B;
  const FastDiv fLL = fast_div(LL);
  for (int e = threadIdx.x; e < S * B; e += kThreads) {
    const int row = e / fLL;
    cs[row * LS + (e - row * LL)] = ci[e];
  }
  for (int e = threadIdx.x; e < 3 * LS; e += kThreads) cs[S * N * LS + e] = T(0);
  ps_tables<SPLIT>(N, S, key_ab, tri);
  __syncthreads();
  ps_body<T, SPLIT, NC, LC>(N, L, S, KC, Dq, i, cs, stage, key_ab, tri, p, p_hi, p_lo, norm2);
}

// ---------------------------------------------------------------------------
// Persistent fused expansion -> power spectrum.  Each workgroup loops over
// centres i = blockIdx.x, blockIdx.x + gridDim.x, ...  For a centre it
// accumulates the expansion straight into LDS (the rows the power spectrum
// reads: S N rows of LS values, the three zero rows after them set once),
// writes c once to HBM for the adjoint (coalesced), then forms the power
// spectrum from LDS as power_spectrum_kernel does (ps_body).  The pair
// batches and the key stage share one LDS region; the first pair batch of
// the next centre is fetched into registers during the power spectrum.
// Same arithmetic as expansion_kernel + power_spectrum_kernel.  With PS
// false only the expansion runs: c accumulated in LDS over all pair
// batches and written once (expansion_kernel keeps one species block in
// registers per batch and so stores and reloads c between batches).
template <class T, int KPT, bool DIRECT, bool SPLIT, bool PS>
__global__ __launch_bounds__(kThreads, 4) void expansion_ps_kernel(
    RadialDev rad, int S, int PB, int KC, int64_t Dq, PairView pv, T* __restrict__ cout,
    unsigned long long* __restrict__ keymask, T* __restrict__ p, unsigned short* __restrict__ p_hi,
    unsigned short* __restrict__ p_lo, double* __restrict__ norm2) {
  const int N = rad.N, L = rad.L, L1 = rad.L1, LL = rad.LL, F = N * L1, B = N * LL, LS = LL | 1;
  const int K = S * (S + 1) / 2;
  const int tid = threadIdx.x;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  T* cs = reinterpret_cast<T*>(smem_raw + ylm_table_bytes<T>(L1));
  int* key_ab = reinterpret_cast<int*>(cs + (S * N + 3) * LS);
  int* tri = key_ab + K;
  unsigned char* region =
      smem_raw + align_up(size_t(reinterpret_cast<unsigned char*>(tri + N * (N + 1) / 2) - smem_raw), 64);
  PairBatch<T> pb = carve_batch<T, DIRECT, false>(region, PB, LL, F);
  T* stage = reinterpret_cast<T*>(region);
  __shared__ int s_start[kMaxSpecies], s_count[kMaxSpecies];
  __shared__ unsigned long long kmask;
  const double inv_s4pi = 0.28209479177387814347;  // 1/sqrt(4 pi)
  load_ylm_table<T>(rad, yc);
  if (PS) {
    for (int e = tid; e < 3 * LS; e += kThreads) cs[S * N * LS + e] = T(0);
    ps_tables<SPLIT>(N, S, key_ab, tri);
  }

  const FastDiv fLL = fast_div(LL);
  int ent[KPT], off[KPT];  // (n*L1 + l) << 16 | lm, and n*LS + lm; -1 past the block
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int e = tid + kThreads * k;
    ent[k] = -1;
    off[k] = 0;
    if (e < B) {
      const int n = e / fLL, lm = e - n * LL;
      ent[k] = ((n * L1 + l_of_lm(lm)) << 16) | lm;
      off[k] = n * LS + lm;
    }
  }
  PairPre pre;
  int64_t i = blockIdx.x;
  if (i < pv.n_centres) {
    const int64_t p0 = pv.centre_off[i] + 1;
    pair_fetch<DIRECT, false>(rad, pv, p0, max(0, min(PB, int(pv.centre_off[i + 1] - p0))), pre);
  }
  for (; i < pv.n_centres; i += gridDim.x) {
    const int64_t p0 = pv.centre_off[i] + 1;  // the self pair contributes nothing here
    const int np = int(pv.centre_off[i + 1] - p0);
    const int si = pv.species_of[i];
    if (tid == 0) kmask = 1ull << si;
    for (int q0 = 0; q0 == 0 || q0 < np; q0 += PB) {
      const int nb = max(0, min(PB, np - q0));
      if (tid < S) s_count[tid] = 0;
      stage_pairs<T, DIRECT, false>(rad, yc, pre, nb, pb, s_start, s_count, &kmask);
      if (q0 + PB < np) pair_fetch<DIRECT, false>(rad, pv, p0 + q0 + PB, min(PB, np - q0 - PB), pre);
      for (int s = 0; s < S; ++s) {
        const int cnt = s_count[s];
        if (q0 > 0 && cnt == 0) continue;
        const int st = s_start[s];
        T* cb = cs + s * N * LS;
        T acc[KPT];
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
          acc[k] = T(0);
          if (ent[k] >= 0) {
            if (q0 > 0)
              acc[k] = cb[off[k]];
            else if (s == si && (ent[k] & 0xffff) == 0)
              acc[k] = T(rad.center[(rad.sp_tab ? rad.sp_tab[si] * rad.N : 0) +
                                    (tid + kThreads * k) / fLL] * inv_s4pi);
          }
        }
        for (int t = 0; t < cnt; ++t) {
          const int q = pb.lst[st + t];
          const T* Yq = pb.Y + q * LL;
          const T* Rq = pb.R + q * F;
#pragma unroll
          for (int k = 0; k < KPT; ++k)
            if (ent[k] >= 0) acc[k] += Rq[ent[k] >> 16] * Yq[ent[k] & 0xffff];
        }
#pragma unroll
        for (int k = 0; k < KPT; ++k)
          if (ent[k] >= 0) cb[off[k]] = acc[k];
      }
      __syncthreads();
    }
    // c for the adjoint, written once
    T* ci = cout + i * int64_t(S) * B;
    for (int e = tid; e < S * B; e += kThreads) {
      const int row = e / fLL;
      ci[e] = cs[row * LS + (e - row * LL)];
    }
    if (tid == 0) keymask[i] = kmask;
    if (PS)
      ps_body<T, SPLIT>(N, L, S, KC, Dq, i, cs, stage, key_ab, tri, p, p_hi, p_lo, norm2);
    else
      __syncthreads();  // c is read out before the next centre overwrites it
    // the next centre's first pair batch
    const int64_t in = i + gridDim.x;
    if (in < pv.n_centres) {
      const int64_t q0n = pv.centre_off[in] + 1;
      pair_fetch<DIRECT, false>(rad, pv, q0n, max(0, min(PB, int(pv.centre_off[in + 1] - q0n))), pre);
    }
  }
}

template <class T>
__global__ void normalise_kernel(int64_t n, int64_t D, const T* __restrict__ p,
                                 const double* __restrict__ norm2, int normalize,
                                 const int64_t* __restrict__ rows, double* __restrict__ out) {
  const int64_t total = n * D;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e / D;
    const int64_t i = rows ? rows[r] : r;
    double v = double(p[i * D + (e - r * D)]);
    if (normalize) {
      const double nr = sqrt(norm2[i]);
      if (nr > 0) v /= nr;
    }
    out[e] = v;
  }
}

// Step 3 of the adjoint for centre i: per-pair contraction of dL/dc (in
// LDS, [species][lm][n]) with the pair gradients, force atomics and the
// centre virial.  The pair batches are staged in `region` (radial functions
// as R[pair][l][n]).  Each pair is contracted by a group
// of 16, 32 or 64 lanes (reduced within the group), sized so that a batch
// spreads over the workgroup; the per-pair FP64 tail (force, atomics,
// virial) then runs one lane per pair on wavefront 0 (PB <= 64).
template <class T, bool DIRECT>
__device__ __forceinline__ void adjoint_pairs(const RadialDev& rad, int S, int PB, const PairView& pv,
                                              const T* yc, const T* dc, unsigned char* region,
                                              int64_t i, double* __restrict__ centre_virial) {
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, F = N * L1, B = N * LL;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // lanes per pair: the largest power of two in [16, 64] that still gives
  // every pair of a full batch its own group
  const int GS = PB <= 4 ? 64 : (PB <= 8 ? 32 : 16);
  const int gpw = 64 / GS;  // groups per wavefront
  const int grp = lane / GS, gl = lane & (GS - 1);
  PairBatch<T> pb = carve_batch<T, DIRECT, true>(region, PB, LL, F);
  const int64_t p0 = pv.centre_off[i] + 1;
  const int np = int(pv.centre_off[i + 1] - p0);
  // dL/dc rows W[lm][n] and the pairs' R[l][n], dR[l][n] are read four
  // radial channels at a time where N allows
  const bool vec4 = (N & 3) == 0;
  typedef typename Vec4<T>::type V4;
  double fi[3] = {0, 0, 0}, v6[6] = {0, 0, 0, 0, 0, 0};
  // gradient terms of Y_lm (ordered before their first use by the barriers
  // of the first batch's staging)
  {
    const YGradTerm<T>* src = ygrad_src<T>(rad);
    for (int e = tid; e < 3 * LL; e += kThreads) pb.gt[e] = src[e];
  }
  // the inputs of each batch are fetched while the previous one is
  // contracted
  PairPre pre;
  pair_fetch<DIRECT, true>(rad, pv, p0, max(0, min(PB, np)), pre);
  for (int q0 = 0; q0 < np; q0 += PB) {
    const int nb = min(PB, np - q0);
    stage_pairs<T, DIRECT, true>(rad, yc, pre, nb, pb, nullptr, nullptr, nullptr);
    if (q0 + PB < np) pair_fetch<DIRECT, true>(rad, pv, p0 + q0 + PB, min(PB, np - q0 - PB), pre);
    for (int qb = gpw * wave; qb < nb; qb += gpw * kWaves) {
      const int q = qb + grp;
      const bool act = q < nb;
      // B_c = sum_lm D[lm] dY_c[lm] with dY_c[lm] = sum of the gradient
      // terms of Y_{l-1} - l r_c Y[lm]: the last part is summed once as
      // Sl = sum_lm l D[lm] Y[lm]
      T A = T(0), Bx = T(0), By = T(0), Bz = T(0), Sl = T(0);
      if (act) {
        const T* W = dc + pb.sp[q] * B;
        const T* Yq = pb.Y + q * LL;
        const T* Rq = pb.R + q * F;
        const T* dRq = pb.dR + q * F;
        const T fc = T(pb.fc[q]), dfc = T(pb.dfc[q]);
        for (int lm = gl; lm < LL; lm += GS) {
          const int l = l_of_lm(lm);
          const T* w = W + lm * N;
          const T* r = Rq + l * N;
          const T* dr = dRq + l * N;
          T Dn = T(0), Ddn = T(0);
          if (vec4) {
            for (int n = 0; n < N; n += 4) {
              const V4 wv = *reinterpret_cast<const V4*>(w + n);
              const V4 rv = *reinterpret_cast<const V4*>(r + n);
              const V4 dv = *reinterpret_cast<const V4*>(dr + n);
              Dn += wv[0] * rv[0] + wv[1] * rv[1] + wv[2] * rv[2] + wv[3] * rv[3];
              Ddn += wv[0] * dv[0] + wv[1] * dv[1] + wv[2] * dv[2] + wv[3] * dv[3];
            }
          } else {
            for (int n = 0; n < N; ++n) {
              Dn += w[n] * r[n];
              Ddn += w[n] * dr[n];
            }
          }
          const T y = Yq[lm];
          A += (Ddn * fc + Dn * dfc) * y;
          Sl += T(l) * Dn * y;
          const YGradTerm<T>* g = pb.gt + 3 * lm;
          Bx += Dn * (g[0].c0 * Yq[g[0].i0] + g[0].c1 * Yq[g[0].i1]);
          By += Dn * (g[1].c0 * Yq[g[1].i0] + g[1].c1 * Yq[g[1].i1]);
          Bz += Dn * (g[2].c0 * Yq[g[2].i0] + g[2].c1 * Yq[g[2].i1]);
        }
      }
      for (int off = GS >> 1; off > 0; off >>= 1) {
        A += __shfl_xor(A, off, GS);
        Bx += __shfl_xor(Bx, off, GS);
        By += __shfl_xor(By, off, GS);
        Bz += __shfl_xor(Bz, off, GS);
        Sl += __shfl_xor(Sl, off, GS);
      }
      if (act && gl == 0) {
        pb.res[4 * q] = A;
        pb.res[4 * q + 1] = Bx - pb.dir[3 * q] * Sl;
        pb.res[4 * q + 2] = By - pb.dir[3 * q + 1] * Sl;
        pb.res[4 * q + 3] = Bz - pb.dir[3 * q + 2] * Sl;
      }
    }
    __syncthreads();
    if (tid < nb) {
      const int q = tid;
      const double d = pb.dist[q];
      const double sc = pb.fc[q] / d;
      const double vx = pb.vec[3 * q], vy = pb.vec[3 * q + 1], vz = pb.vec[3 * q + 2];
      const double A = double(pb.res[4 * q]);
      const double g[3] = {A * vx / d + sc * double(pb.res[4 * q + 1]),
                           A * vy / d + sc * double(pb.res[4 * q + 2]),
                           A * vz / d + sc * double(pb.res[4 * q + 3])};
      const int j = pb.pj[q];
      double* pf = pv.pair_force + 3 * (p0 + q0 + q);
      pf[0] = g[0];
      pf[1] = g[1];
      pf[2] = g[2];
      if (j != int(i)) {
        fi[0] -= g[0];
        fi[1] -= g[1];
        fi[2] -= g[2];
      }
      // r_ji = r_i - r_j(image) = -pair vector
      v6[0] -= vx * g[0];
      v6[1] -= vy * g[1];
      v6[2] -= vz * g[2];
      v6[3] -= vy * g[2];
      v6[4] -= vz * g[0];
      v6[5] -= vx * g[1];
    }
    __syncthreads();  // the next batch overwrites this one
  }
  // only wavefront 0 holds partial sums
  if (wave == 0) {
    double r[9] = {fi[0], fi[1], fi[2], v6[0], v6[1], v6[2], v6[3], v6[4], v6[5]};
#pragma unroll
    for (int k = 0; k < 9; ++k)
      for (int off = 32; off > 0; off >>= 1) r[k] += __shfl_down(r[k], off, 64);
    if (lane == 0) {
      for (int k = 0; k < 3; ++k) pv.centre_force[3 * i + k] = r[k];
      for (int k = 0; k < 6; ++k) centre_virial[6 * i + k] = r[3 + k];
    }
  }
}

// ---------------------------------------------------------------------------
// Fused adjoint.  Given G = dL/dX (X = p/|p| the exported features):
//  1. dL/dp = G/|p| - (X.G) p/|p|^2 (the transpose of the normalisation
//     derivative, calculator_spherical_invariants.hh:588-652); X.G is either
//     passed per centre (for the GAP energy X.G = zeta E_i, the energy being
//     homogeneous of degree zeta in X) or reduced here;
//  2. dL/dc_a[n,lm] = sum over keys (a,b) of (2l+1)^{-1/2} (sqrt 2 if a<b)
//     sum_n2 dL/dp^{ab}[n,n2,l] c_b[n2,lm] (both sides of each key; the
//     diagonal keys symmetrised), kept in LDS;
//  3. per pair (i, j): g = sum_{n,lm} dL/dc[s_j][n][lm]
//       [ (R' fc + R fc') Y r_hat + R fc grad_Y / r ]
//     (the pair-gradient formula of calculator_spherical_expansion.hh
//     :2306-2366 contracted with dL/dc) is stored per pair, and the centre
//     keeps -sum g over the pairs whose j is not a periodic image of i;
//     force_gather_kernel then sums, per atom, its centre term and the g of
//     the pairs naming it, in ascending pair order (no atomics: the forces
//     are the same on every run); the virial sum r_ji (x) g is reduced per
//     centre (sparse_kernel_predict.hh:354-364).
constexpr int kAdjPre = 8;  // key-block elements per thread held for the prefetch

template <class T, bool DIRECT>
__global__ __launch_bounds__(kThreads) void adjoint_kernel(
    RadialDev rad, int S, int PB, int c_in_lds, PairView pv, const T* __restrict__ c,
    const T* __restrict__ p, const double* __restrict__ norm2, int normalize,
    const double* __restrict__ xg, double xg_scale, const T* __restrict__ G,
    const T* __restrict__ Grow, const double* __restrict__ Gscale,
    double* __restrict__ centre_virial) {
  typedef typename Vec4<T>::type V4;
  const int64_t i = blockIdx.x;
  // rank-one dL/dX (Grow != null): G_i = Gscale[i] Grow; a centre with a
  // zero scale contributes nothing
  if (Grow && Gscale[i] == 0.0) {
    if (threadIdx.x < 6) centre_virial[6 * i + threadIdx.x] = 0.0;
    if (threadIdx.x < 3) pv.centre_force[3 * i + threadIdx.x] = 0.0;
    const int64_t q0 = 3 * (pv.centre_off[i] + 1), q1 = 3 * pv.centre_off[i + 1];
    for (int64_t q = q0 + threadIdx.x; q < q1; q += kThreads) pv.pair_force[q] = 0.0;
    return;
  }
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, B = N * LL, E = S * B;
  const int N4 = (N + 3) & ~3;
  const int K = S * (S + 1) / 2;
  const int KB = N * N * L1;
  const int64_t D = int64_t(K) * KB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  T* dc = reinterpret_cast<T*>(smem_raw + ylm_table_bytes<T>(L1));  // dL/dc [S][LL][N]
  unsigned char* region = reinterpret_cast<unsigned char*>(dc) + align_up(sizeof(T) * E, 64);
  load_ylm_table<T>(rad, yc);
  const T* cs;
  T* gA;
  if (c_in_lds) {
    T* cl = reinterpret_cast<T*>(region);
    for (int e = tid; e < E; e += kThreads) cl[e] = c[i * int64_t(E) + e];
    cs = cl;
    gA = reinterpret_cast<T*>(region + align_up(sizeof(T) * E, 64));
  } else {
    cs = c + i * int64_t(E);
    gA = reinterpret_cast<T*>(region);
  }
  // gA[l][x][y] = g[y][x][l], gB[l][x][y] = g[x][y][l]; the l planes LS
  // apart, two values more than a plane, so that the transposing stores of
  // consecutive l (one per thread) fall in different LDS banks
  const int LS = N * N4 + 2;
  T* gB = gA + L1 * LS;
  __shared__ double red[kWaves][9];
  __shared__ double sh_ab[2];

  const T* Gi = Grow ? nullptr : G + i * D;
  const T gs = Grow ? T(Gscale[i]) : T(1);
  auto gval = [&](int64_t o) -> T { return Grow ? gs * Grow[o] : Gi[o]; };
  auto pval = [&](int64_t o) -> T { return p[i * D + o]; };
  double alpha = 1.0, beta = 0.0;
  if (normalize) {
    const double n2 = norm2[i];
    double dotxg = 0;
    if (xg) {
      dotxg = xg_scale * xg[i];
    } else {
      double dot = 0;
      for (int64_t o = tid; o < D; o += kThreads) dot += double(pval(o)) * double(gval(o));
      for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
      if (lane == 0) red[wave][0] = dot;
      __syncthreads();
      if (tid == 0) {
        double t = 0;
        for (int w = 0; w < kWaves; ++w) t += red[w][0];
        sh_ab[0] = t;
      }
      __syncthreads();
      dotxg = n2 > 0 ? sh_ab[0] / sqrt(n2) : 0.0;
    }
    if (n2 > 0) {
      alpha = 1.0 / sqrt(n2);
      beta = dotxg / n2;
    }
  }
  const T ta = T(alpha), tb = T(beta);
  for (int e = tid; e < E; e += kThreads) dc[e] = T(0);
  for (int e = tid; e < 2 * L1 * LS; e += kThreads) gA[e] = T(0);  // zero pads

  // key blocks: element o = (x*N + y)*L1 + l; thread holds o = tid + 256 r
  const FastDiv fNL1 = fast_div(N * L1), fL1 = fast_div(L1), fLL = fast_div(LL);
  const bool prefetch = KB <= kAdjPre * kThreads;
  T pre[kAdjPre];
  auto fetch = [&](int key) {
    if (!prefetch) return;
    const int64_t pk = int64_t(key) * KB;
#pragma unroll
    for (int r = 0; r < kAdjPre; ++r) {
      const int o = tid + kThreads * r;
      pre[r] = o < KB ? ta * gval(pk + o) - tb * pval(pk + o) : T(0);
    }
  };
  fetch(0);
  __syncthreads();
  const int units_a = (N4 / 4) * LL;
  for (int key = 0; key < K; ++key) {
    int a, b;
    key_species(key, S, &a, &b);
    if (prefetch) {
#pragma unroll
      for (int r = 0; r < kAdjPre; ++r) {
        const int o = tid + kThreads * r;
        if (o < KB) {
          const int x = o / fNL1, rr = o - x * N * L1, y = rr / fL1, l = rr - y * L1;
          gA[l * LS + y * N4 + x] = pre[r];
          gB[l * LS + x * N4 + y] = pre[r];
        }
      }
    } else {
      const int64_t pk = int64_t(key) * KB;
      for (int o = tid; o < KB; o += kThreads) {
        const T g = ta * gval(pk + o) - tb * pval(pk + o);
        const int x = o / fNL1, rr = o - x * N * L1, y = rr / fL1, l = rr - y * L1;
        gA[l * LS + y * N4 + x] = g;
        gB[l * LS + x * N4 + y] = g;
      }
    }
    __syncthreads();
    if (a == b) {
      // symmetrise: gA[l][x][y] = g[y][x] + g[x][y]
      for (int e = tid; e < L1 * LS; e += kThreads) gA[e] += gB[e];
      __syncthreads();
    }
    if (key + 1 < K) fetch(key + 1);
    const T w = (a < b) ? T(1.41421356237309504880) : T(1);
    const int units = units_a * (a < b ? 2 : 1);
    for (int u = tid; u < units; u += kThreads) {
      const bool side_b = u >= units_a;
      const int uu = side_b ? u - units_a : u;
      const int nt = uu / fLL, lm = uu - nt * LL;
      const int l = l_of_lm(lm);
      const T* gt = (side_b ? gB : gA) + l * LS + 4 * nt;
      const T* cx = cs + (side_b ? a : b) * B + lm;
      V4 s4 = {T(0), T(0), T(0), T(0)};
      for (int x = 0; x < N; ++x) {
        const T cv = cx[x * LL];
        const V4 g4 = *reinterpret_cast<const V4*>(gt + x * N4);
        s4 += g4 * cv;
      }
      const T f = w / sqrt(T(2 * l + 1));
      T* out = dc + (side_b ? b : a) * B + lm * N;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = 4 * nt + v;
        if (n < N) out[n] += f * s4[v];
      }
    }
    __syncthreads();
  }

  // per-pair contraction; the batch reuses the region after dL/dc
  adjoint_pairs<T, DIRECT>(rad, S, PB, pv, yc, dc, region, i, centre_virial);
}

// Adjoint of the mixed path: dL/dX (G) and p in the packed layout (p as its
// BF16 operand planes), the dL/dp blocks staged in LDS in the full layout
// KCH key blocks at a time beside the expansion coefficients.  Same
// mathematics as adjoint_kernel; the dL/dc contraction keeps its outputs in
// registers across the key blocks instead of one LDS update per key:
//  1. g = alpha G - beta p of keys [k0, k0 + KCH) in the layout
//     [key][x][y][l]; a diagonal key is unpacked symmetrised, g + g^T
//     (sqrt 2 (alpha G' - beta p') off the diagonal, 2 (alpha G' - beta p')
//     on it, from the packed values G', p');
//  2. thread unit (a, 4 radial channels n, lm):
//       dL/dc_a[n][lm] += (2l+1)^{-1/2} w_ab sum_y g^(ab)[n][y][l] c_b[y][lm]
//     over the species b whose key is in the chunk (g^(ab) read transposed
//     for a > b, w_ab = sqrt 2 for a != b); after the last chunk written
//     over c;
//  3. the pair phase, its batches over the g region.
// UPT = units per thread (ceil(S * ceil(N/4) * LL / 256)).
template <bool DIRECT, int UPT>
__global__ __launch_bounds__(kThreads) void adjoint_staged_kernel(
    RadialDev rad, int S, int PB, int KCH, PairView pv, const float* __restrict__ c,
    const unsigned short* __restrict__ p_hi, const unsigned short* __restrict__ p_lo,
    const double* __restrict__ norm2, int normalize, const double* __restrict__ xg, double xg_scale,
    const float* __restrict__ G, double* __restrict__ centre_virial) {
  typedef float T;
  typedef typename Vec4<T>::type V4;
  const int64_t i = blockIdx.x;
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, B = N * LL, E = S * B;
  const int K = S * (S + 1) / 2;
  const int KB = N * N * L1, KD = packed_diag_block(N, L1);
  const int64_t Dp = packed_feature_size(S, N, L1), Dq = packed_row_stride(S, N, L1);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  // c [S][N][LL], then dL/dc [S][LL][N]
  T* cd = reinterpret_cast<T*>(smem_raw + ylm_table_bytes<T>(L1));
  unsigned char* region = reinterpret_cast<unsigned char*>(cd) + align_up(sizeof(T) * E, 64);
  T* gl = reinterpret_cast<T*>(region);
  __shared__ double red[kWaves];
  __shared__ double sh_ab[2];
  __shared__ int tri[kMaxN * (kMaxN + 1) / 2];  // n1 << 8 | n2 of the packed diagonal pairs
  load_ylm_table<T>(rad, yc);
  for (int e = tid; e < E; e += kThreads) cd[e] = c[i * int64_t(E) + e];
  for (int t = tid; t < N * (N + 1) / 2; t += kThreads) {
    int n1, n2;
    packed_pair(t, N, &n1, &n2);
    tri[t] = (n1 << 8) | n2;
  }
  // the first chunk's staging reads tri, and the X . G reduction below is
  // skipped when xg is given or the rows are not normalised
  __syncthreads();

  const T* Gi = G + i * Dq;
  const unsigned short* Hi = p_hi + i * Dq;
  const unsigned short* Lo = p_lo + i * Dq;
  auto pval = [&](int64_t o) -> T {
    return __builtin_bit_cast(float, unsigned(Hi[o]) << 16) +
           __builtin_bit_cast(float, unsigned(Lo[o]) << 16);
  };
  double alpha = 1.0, beta = 0.0;
  if (normalize) {
    const double n2 = norm2[i];
    double dotxg = 0;
    if (xg) {
      dotxg = xg_scale * xg[i];
    } else {
      // X . G of the full rows equals the packed dot product
      double dot = 0;
      for (int64_t o = tid; o < Dp; o += kThreads) dot += double(pval(o)) * double(Gi[o]);
      for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
      if (lane == 0) red[wave] = dot;
      __syncthreads();
      if (tid == 0) {
        double t = 0;
        for (int w = 0; w < kWaves; ++w) t += red[w];
        sh_ab[0] = t;
      }
      __syncthreads();
      dotxg = n2 > 0 ? sh_ab[0] / sqrt(n2) : 0.0;
    }
    if (n2 > 0) {
      alpha = 1.0 / sqrt(n2);
      beta = dotxg / n2;
    }
  }
  const T ta = T(alpha), tb = T(beta);
  const FastDiv fL1 = fast_div(L1);
  // four elements per step where a key block keeps 16-byte alignment
  const bool vec4 = ((KB | KD) & 3) == 0;

  // dL/dc units: thread unit (a, radial channels n0..n0+3, lm)
  const int NQ = (N + 3) >> 2;
  const int units = S * NQ * LL;
  const int sx = N * L1;  // stride of the first radial index of a key block
  T acc[UPT][4];
#pragma unroll
  for (int k = 0; k < UPT; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = T(0);

  for (int k0 = 0; k0 < K; k0 += KCH) {
    const int k1 = min(K, k0 + KCH);
    if (k0 > 0) __syncthreads();  // every unit has read the previous chunk
    // 1. stage g of keys [k0, k1)
    for (int k = k0; k < k1; ++k) {
      int a, b;
      key_pair(k, S, &a, &b);
      T* gk = gl + (k - k0) * KB;
      const int64_t src = packed_key_offset(a, b, S, N, L1);
      if (a < b) {
        if (vec4) {
          const V4* G4 = reinterpret_cast<const V4*>(Gi + src);
          const uint2* H2 = reinterpret_cast<const uint2*>(Hi + src);
          const uint2* L2 = reinterpret_cast<const uint2*>(Lo + src);
          V4* g4 = reinterpret_cast<V4*>(gk);
          for (int q = tid; q < (KB >> 2); q += kThreads) {
            const V4 gv = G4[q];
            const uint2 h = H2[q], lo = L2[q];
            V4 pv4;
            pv4[0] = __builtin_bit_cast(float, h.x << 16) + __builtin_bit_cast(float, lo.x << 16);
            pv4[1] = __builtin_bit_cast(float, h.x & 0xffff0000u) +
                     __builtin_bit_cast(float, lo.x & 0xffff0000u);
            pv4[2] = __builtin_bit_cast(float, h.y << 16) + __builtin_bit_cast(float, lo.y << 16);
            pv4[3] = __builtin_bit_cast(float, h.y & 0xffff0000u) +
                     __builtin_bit_cast(float, lo.y & 0xffff0000u);
            g4[q] = ta * gv - tb * pv4;
          }
        } else {
          for (int o = tid; o < KB; o += kThreads) gk[o] = ta * Gi[src + o] - tb * pval(src + o);
        }
      } else {
        for (int o = tid; o < KD; o += kThreads) {
          const int t = o / fL1, l = o - t * L1;
          const int n1 = tri[t] >> 8, n2 = tri[t] & 255;
          const T v = ta * Gi[src + o] - tb * pval(src + o);
          const T w = n1 < n2 ? T(1.41421356237309504880) * v : T(2) * v;
          gk[(n1 * N + n2) * L1 + l] = w;
          gk[(n2 * N + n1) * L1 + l] = w;
        }
      }
    }
    __syncthreads();

    // 2. contraction of the chunk's keys into the register accumulators
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
      const int u = tid + kThreads * k;
      if (u < units) {
        const int anq = u / LL, lm = u - anq * LL;
        const int a = anq / NQ, n0 = 4 * (anq - a * NQ);
        const int l = l_of_lm(lm);
        int na[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) na[v] = min(n0 + v, N - 1);
        for (int b = 0; b < S; ++b) {
          const int lo = min(a, b), hi = max(a, b);
          const int key = lo * S - lo * (lo - 1) / 2 + (hi - lo);
          if (key < k0 || key >= k1) continue;
          const T* gk = gl + (key - k0) * KB + l;
          // g^(ab)[n][y] for a <= b, g^(ba)[y][n] for a > b
          const int sn = a <= b ? sx : L1, sy = a <= b ? L1 : sx;
          const T* cb = cd + b * B + lm;
          T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
          const T* g0 = gk + na[0] * sn;
          const T* g1 = gk + na[1] * sn;
          const T* g2 = gk + na[2] * sn;
          const T* g3 = gk + na[3] * sn;
          for (int y = 0; y < N; ++y) {
            const T cv = cb[y * LL];
            const int oy = y * sy;
            s0 += g0[oy] * cv;
            s1 += g1[oy] * cv;
            s2 += g2[oy] * cv;
            s3 += g3[oy] * cv;
          }
          const T w = a == b ? T(1) : T(1.41421356237309504880);
          acc[k][0] += w * s0;
          acc[k][1] += w * s1;
          acc[k][2] += w * s2;
          acc[k][3] += w * s3;
        }
      }
    }
  }
  __syncthreads();  // every unit has read c
#pragma unroll
  for (int k = 0; k < UPT; ++k) {
    const int u = tid + kThreads * k;
    if (u < units) {
      const int anq = u / LL, lm = u - anq * LL;
      const int a = anq / NQ, n0 = 4 * (anq - a * NQ);
      const T f = T(1) / sqrt(T(2 * l_of_lm(lm) + 1));
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (n0 + v < N) cd[a * B + lm * N + n0 + v] = f * acc[k][v];
    }
  }
  __syncthreads();

  // 3. pairs; the batches reuse the g region
  adjoint_pairs<T, DIRECT>(rad, S, PB, pv, yc, cd, region, i, centre_virial);
}

// ---------------------------------------------------------------------------
// Adjoint of the mixed path on the matrix cores (v_mfma_f32_16x16x4_f32).
//
// Per angular channel l, with C_l[(a,n)][m] = c_a[n][l^2 + m] (S N rows,
// 2l+1 columns) and f_l = (2l+1)^{-1/2}, the transpose of the power
// spectrum p^{ab}[n1 n2 l] = f_l w_ab sum_m c_a[n1 lm] c_b[n2 lm]
// (calculator_spherical_invariants.hh:747-788) applied to
// dL/dp = alpha G - beta p is
//   dL/dC_l = alpha f_l W_l C_l - 2 beta f_l^2 C_l (C_l^T C_l),
// where W_l is the symmetric S N x S N matrix with blocks
// W^{ab} = sqrt2 G^{ab} (a < b), its transpose (a > b) and G^{aa} + G^{aa T}
// (the last term because |p|^2 = sum_l f_l^2 |C_l C_l^T|_F^2 =
// sum_l f_l^2 |C_l^T C_l|_F^2, whose gradient is 4 f_l^2 C_l C_l^T C_l).
// So the invariants themselves are never read: only G (packed layout, as
// written by the force GEMM) and c.  X . G, when not given, follows from the
// same products: p . G = 1/2 sum_l f_l <W_l C_l, C_l>.
//
// Tiles: one 16-row tile per species (N <= 16 rows used) and 16-column
// tiles over m for each l; K runs over the 4-row steps of one species
// block.  A key block (a, b) is staged in LDS as W^{ab} [l][n][y] (rows of
// NP = 2 mod 4 floats, so that the 16 rows x 2 K columns of a half
// wavefront's A fragment fall in distinct banks), the sqrt 2 of an
// off-diagonal block folded in and a diagonal block unpacked symmetrised
// (sqrt2 G' off the diagonal, 2 G' on it, G' packed).  It feeds the tiles
// of rows a with K over block b, and (a < b) the tiles of rows b with K
// over block a (read transposed).
// Wavefront w owns the output tiles t = w, w + 4, ... (MAXT at most), its
// accumulators in registers across all key chunks.  The result is written
// over c in LDS as dL/dc [species][lm][n]; the pair phase follows.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <bool DIRECT, int MAXT>
__global__ __launch_bounds__(kThreads, 2) void adjoint_mfma_kernel(
    RadialDev rad, int S, int PB, int KCH, PairView pv, const float* __restrict__ c,
    const double* __restrict__ norm2, int normalize, const double* __restrict__ xg, double xg_scale,
    const float* __restrict__ G, double* __restrict__ centre_virial) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, L = rad.L, L1 = rad.L1, LL = rad.LL, B = N * LL, E = S * B;
  // row stride of c in LDS, 16 mod 32: the four 16-lane groups of a B
  // fragment (rows y..y+3, 16 consecutive m) cover the banks twice
  const int LS = (LL + 15) / 32 * 32 + 16;
  const int NP = 4 * ((N + 1) / 4) + 2;  // row stride of a staged W block (>= N, 2 mod 4)
  const int WB = L1 * N * NP;           // one staged key block
  const int K = S * (S + 1) / 2;
  const int64_t Dq = packed_row_stride(S, N, L1);
  const int tid = threadIdx.x, lane = tid & 63;
  // wavefront-uniform (scalar registers for the tile bookkeeping)
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int i16 = lane & 15, k4 = lane >> 4;
  const int NKS = (N + 3) >> 2;  // K steps over one species block
  extern __shared__ __align__(64) unsigned char smem_raw[];
  float* yc = reinterpret_cast<float*>(smem_raw);
  float* cs = reinterpret_cast<float*>(smem_raw + ylm_table_bytes<float>(L1));  // [S N][LS]
  unsigned char* region = reinterpret_cast<unsigned char*>(cs) + align_up(sizeof(float) * S * N * LS, 64);
  float* gl = reinterpret_cast<float*>(region);
  __shared__ double red[kWaves];
  __shared__ double sh_ab[2];
  load_ylm_table<float>(rad, yc);
  const float* ci = c + i * int64_t(E);
  for (int e = tid; e < E; e += kThreads) {
    const int row = e / LL;
    cs[row * LS + (e - row * LL)] = ci[e];
  }

  // output tiles of this wavefront: (species a, l, column tile ct)
  int T_col = 0;
  for (int l = 0; l <= L; ++l) T_col += (2 * l + 16) >> 4;
  const int n_tiles = S * T_col;
  int t_a[MAXT], t_l[MAXT], t_ct[MAXT];
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) {
    const int t = wave + kWaves * tt;
    t_a[tt] = -1;
    t_l[tt] = 0;
    t_ct[tt] = 0;
    if (t < n_tiles) {
      t_a[tt] = t / T_col;
      int u = t - t_a[tt] * T_col, l = 0;
      while (u >= ((2 * l + 16) >> 4)) {
        u -= (2 * l + 16) >> 4;
        ++l;
      }
      t_l[tt] = l;
      t_ct[tt] = u;
    }
  }
  // per-lane A offsets (l-independent) of the K steps: rows n = i16, K
  // index y = 4 ks + k4 of a block, in the orientations ab / ba
  int off_ab[4], off_ba[4], boff[4];
  bool ok[4];  // K index inside the block (only tested when N is not a multiple of 4)
  const bool npad = (N & 3) != 0;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int n = i16, y = 4 * ks + k4;
    ok[ks] = y < N;
    off_ab[ks] = n * NP + y;
    off_ba[ks] = y * NP + n;
    boff[ks] = y * LS;
  }
  f32x4 acc[MAXT];
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) acc[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int KB = N * N * L1, KD = packed_diag_block(N, L1);
  const FastDiv fL1 = fast_div(L1), fN = fast_div(N);
  __shared__ int tri[kMaxN * (kMaxN + 1) / 2];  // n1 << 8 | n2 of the packed diagonal pairs
  for (int t = tid; t < N * (N + 1) / 2; t += kThreads) {
    int n1, n2;
    packed_pair(t, N, &n1, &n2);
    tri[t] = (n1 << 8) | n2;
  }
  __syncthreads();  // tri and c in place

  const float* Gi = G + i * Dq;
  const float rt2 = 1.41421356237309504880f;
  for (int k0 = 0; k0 < K; k0 += KCH) {
    const int k1 = min(K, k0 + KCH);
    if (k0 > 0) __syncthreads();  // every wavefront is done with the previous chunk
    // stage W of the chunk's keys: the chunk's packed range read with
    // independent 16-byte loads, each value scattered to its (l, n1, n2)
    // slot (both slots for a diagonal block)
    {
      int a0, b0, a1, b1;
      key_pair(k0, S, &a0, &b0);
      const int g0 = int(packed_key_offset(a0, b0, S, N, L1));
      int g1 = int(packed_feature_size(S, N, L1));
      if (k1 < K) {
        key_pair(k1, S, &a1, &b1);
        g1 = int(packed_key_offset(a1, b1, S, N, L1));
      }
      const int cnt = g1 - g0;
      auto put = [&](int o, float v) {
        // key of packed column g0 + o (the chunk has at most K keys)
        int key = k0, a = a0, b = b0;
        int start = 0, len = a == b ? KD : KB;
        while (o >= start + len) {
          start += len;
          ++key;
          key_pair(key, S, &a, &b);
          len = a == b ? KD : KB;
        }
        const int rel = o - start, t = rel / fL1, l = rel - t * L1;
        float* w = gl + (key - k0) * WB + l * N * NP;
        if (a < b) {
          const int n1 = t / fN, n2 = t - n1 * N;
          w[n1 * NP + n2] = rt2 * v;
        } else {
          const int n1 = tri[t] >> 8, n2 = tri[t] & 255;
          const float x = (n1 < n2 ? rt2 : 2.0f) * v;
          w[n1 * NP + n2] = x;
          w[n2 * NP + n1] = x;
        }
      };
      if (((g0 | cnt) & 3) == 0) {
        const f32x4* src = reinterpret_cast<const f32x4*>(Gi + g0);
#pragma unroll 4
        for (int q = tid; q < (cnt >> 2); q += kThreads) {
          const f32x4 v = src[q];
          put(4 * q, v[0]);
          put(4 * q + 1, v[1]);
          put(4 * q + 2, v[2]);
          put(4 * q + 3, v[3]);
        }
      } else {
#pragma unroll 4
        for (int q = tid; q < cnt; q += kThreads) put(q, Gi[g0 + q]);
      }
    }
    __syncthreads();
    // each tile of rows a takes K over the blocks b whose key is staged:
    // W^{ab} = block (a, b) as stored for a <= b, block (b, a) read
    // transposed for a > b.  Rows n >= N and columns m >= 2l+1 of a tile
    // read neighbouring LDS values: they only reach outputs that are never
    // used.  K indices y >= N (N not a multiple of 4) are zeroed.
#pragma unroll
    for (int tt = 0; tt < MAXT; ++tt) {
      const int a = t_a[tt];
      if (a < 0) continue;
      const int l = t_l[tt];
      const float* cl = cs + l * l + 16 * t_ct[tt] + i16;
      for (int b = 0; b < S; ++b) {
        const int lo = min(a, b), hi = max(a, b);
        const int key = lo * S - lo * (lo - 1) / 2 + (hi - lo);
        if (key < k0 || key >= k1) continue;
        const float* gk = gl + (key - k0) * WB + l * N * NP;
        const float* cb = cl + b * N * LS;
        if (npad) {
#pragma unroll
          for (int ks = 0; ks < 4; ++ks) {
            if (ks >= NKS) break;
            const float av = ok[ks] ? gk[a <= b ? off_ab[ks] : off_ba[ks]] : 0.f;
            const float bv = ok[ks] ? cb[boff[ks]] : 0.f;
            acc[tt] = mfma4(av, bv, acc[tt]);
          }
        } else {
#pragma unroll
          for (int ks = 0; ks < 4; ++ks) {
            if (ks >= NKS) break;
            acc[tt] = mfma4(gk[a <= b ? off_ab[ks] : off_ba[ks]], cb[boff[ks]], acc[tt]);
          }
        }
      }
    }
  }

  // normalisation: alpha = 1/|p|, beta = (X . G)/|p|^2
  double alpha = 1.0, beta = 0.0;
  if (normalize) {
    const double n2 = norm2[i];
    double dotxg;
    if (xg) {
      dotxg = xg_scale * xg[i];
    } else {
      // p . G = 1/2 sum_l f_l <W_l C_l, C_l> from the accumulators
      double d = 0;
#pragma unroll
      for (int tt = 0; tt < MAXT; ++tt) {
        if (t_a[tt] < 0) continue;
        const int l = t_l[tt], m = 16 * t_ct[tt] + i16;
        if (m >= 2 * l + 1) continue;
        const float fl = 1.0f / sqrtf(float(2 * l + 1));
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int n = 4 * k4 + v;
          if (n < N) d += double(fl) * acc[tt][v] * cs[(t_a[tt] * N + n) * LS + l * l + m];
        }
      }
      d *= 0.5;
      for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, 64);
      if (lane == 0) red[wave] = d;
      __syncthreads();
      if (tid == 0) {
        double t = 0;
        for (int w = 0; w < kWaves; ++w) t += red[w];
        sh_ab[0] = t;
      }
      __syncthreads();
      dotxg = n2 > 0 ? sh_ab[0] / sqrt(n2) : 0.0;
    }
    if (n2 > 0) {
      alpha = 1.0 / sqrt(n2);
      beta = dotxg / n2;
    }
  }
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) {
    const float s = float(alpha) / sqrtf(float(2 * t_l[tt] + 1));
    acc[tt] *= s;
  }
  if (beta != 0.0) {
    // M_l = -2 beta f_l^2 C_l^T C_l, [16 nct][16 nct] per l, in the G region
    __shared__ int moff[kMaxL1];  // first entry of M_l
    if (tid == 0) {
      int o = 0;
      for (int l = 0; l <= L; ++l) {
        moff[l] = o;
        const int w = 16 * ((2 * l + 16) >> 4);
        o += w * w;
      }
    }
    __syncthreads();  // the G chunk is no longer read
    float* Ml = gl;
    const int rows = S * N, RKS = (rows + 3) >> 2;
    // tiles (l, ct, ct2) dealt to the wavefronts in turn
    int tix = 0;
    for (int l = 0; l <= L; ++l) {
      const int nct = (2 * l + 16) >> 4, w = 16 * nct;
      const float sc = float(-2.0 * beta) / float(2 * l + 1);
      for (int ct = 0; ct < nct; ++ct)
        for (int ct2 = 0; ct2 < nct; ++ct2, ++tix) {
          if ((tix & (kWaves - 1)) != wave) continue;
          const int m1 = 16 * ct + i16, m2 = 16 * ct2 + i16;
          const bool ok1 = m1 < 2 * l + 1, ok2 = m2 < 2 * l + 1;
          f32x4 mt = f32x4{0.f, 0.f, 0.f, 0.f};
          for (int ks = 0; ks < RKS; ++ks) {
            const int r = 4 * ks + k4;
            float av = 0.f, bv = 0.f;
            if (r < rows) {
              if (ok1) av = cs[r * LS + l * l + m1];
              if (ok2) bv = cs[r * LS + l * l + m2];
            }
            mt = mfma4(av, bv, mt);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v)
            Ml[moff[l] + (16 * ct + 4 * k4 + v) * w + 16 * ct2 + i16] = sc * mt[v];
        }
    }
    __syncthreads();
    // dL/dC_l += C_l M_l
#pragma unroll
    for (int tt = 0; tt < MAXT; ++tt) {
      if (t_a[tt] < 0) continue;
      const int l = t_l[tt], w = 16 * ((2 * l + 16) >> 4);
      const int m = 16 * t_ct[tt] + i16;
      const int nks = (2 * l + 4) >> 2;  // ceil((2l+1)/4)
      const float* ca = cs + (t_a[tt] * N + i16) * LS + l * l;
      const float* mb = Ml + moff[l] + m;
      for (int ks = 0; ks < nks; ++ks) {
        const int mp = 4 * ks + k4;
        const float av = (i16 < N && mp < 2 * l + 1) ? ca[mp] : 0.f;
        const float bv = mp < 2 * l + 1 ? mb[mp * w] : 0.f;
        acc[tt] = mfma4(av, bv, acc[tt]);
      }
    }
  }
  __syncthreads();  // every read of c is done: dL/dc goes over it
  float* dc = cs;   // [S][LL][N]
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) {
    if (t_a[tt] < 0) continue;
    const int l = t_l[tt], m = 16 * t_ct[tt] + i16;
    if (m >= 2 * l + 1) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int n = 4 * k4 + v;
      if (n < N) dc[t_a[tt] * B + (l * l + m) * N + n] = acc[tt][v];
    }
  }
  __syncthreads();
  adjoint_pairs<float, DIRECT>(rad, S, PB, pv, yc, dc, region, i, centre_virial);
}

// ---------------------------------------------------------------------------
// Adjoint of the mixed path on register tiles, for compile-time N and L.
//
// Same mathematics as adjoint_mfma_kernel: per angular channel l,
//   dL/dC_l = alpha f_l W_l C_l - 2 beta f_l^2 C_l (C_l^T C_l),
// with W_l assembled from the packed dL/dX blocks; the invariants are not
// read.  Work split: wavefront w takes the rows of species a = w, w + 4, ...;
// lane (tile, l, quad) owns an NR x 4 register tile of dL/dC_l (radial rows
// tile*NR .. tile*NR+NR-1, columns m = 4 quad .. 4 quad + 3 of channel l), so
// one (b, n2) step of W_l C_l is a 16-byte read of c_b[n2][l, m..m+3], NR
// reads of W and 4 NR FMAs, every LDS offset of the step a compile-time
// immediate.  The key blocks pass through LDS in chunks of whole rounds
// (AdjRounds): each species has at most one block per round, so the
// wavefronts get equal work in every chunk.  The packed values of the next
// chunk are loaded into registers while the current chunk is contracted.
//
// LDS: Y_lm table | c [S][N][MT] (the m of each l padded to whole quads, pads
// zero) | C_l^T C_l as [l][m'][4 ceil((2l+1)/4)] | staged blocks in the full
// [n1][n2][l] layout (a diagonal block unpacked symmetric, its diagonal
// entries times sqrt 2; the common sqrt 2 of all blocks applied to the
// result), later reused by the pair batches.  dL/dc [S][LL][N] goes over c.

// quads of a padded c row before channel l: sum_{l' < l} ceil((2l'+1)/4)
__host__ __device__ constexpr int adj_quads_before(int l) { return l + (l - 1) * (l - 1) / 4; }
// entries of the C_l^T C_l table before channel l
__host__ __device__ constexpr int adj_gram_before(int l) {
  int s = 0;
  for (int t = 0; t < l; ++t) s += (2 * t + 1) * 4 * ((t >> 1) + 1);
  return s;
}


// Schedule of the key blocks (a, b), a <= b, in rounds (circle method).
// Round 0 holds the diagonal blocks, block k = (k, k).  With n = S rounded up
// to even (species n-1 a bye when S is odd) and r' = r - 1, round r >= 1
// pairs (r', n-1) as its block k = 0 and ((r'+k) mod (n-1), (r'-k) mod (n-1))
// as block k for 1 <= k < n/2; a bye's block is dropped, so the blocks of a
// round are k = k0 .. n/2 - 1, k0 = S mod 2.  Every unordered pair of species
// meets in exactly one round.
struct AdjRounds {
  int S, n, k0, per;  // per: blocks of a round r >= 1
  __host__ __device__ explicit AdjRounds(int s) : S(s), n(s + (s & 1)), k0(s & 1), per((s + (s & 1)) / 2 - (s & 1)) {}
  __host__ __device__ int rounds() const { return n; }
  __host__ __device__ int blocks(int r) const { return r == 0 ? S : per; }
  __host__ __device__ void pair(int r, int k, int* a, int* b) const {
    const int m = n - 1, rr = r - 1;
    int x = rr, y = m;
    if (k > 0) {
      x = rr + k;
      if (x >= m) x -= m;
      y = rr - k;
      if (y < 0) y += m;
    }
    *a = x < y ? x : y;
    *b = x < y ? y : x;
  }
  // block index of species a in round r (-1: a bye)
  __host__ __device__ int block_of(int r, int a) const {
    if (r == 0) return a;
    const int m = n - 1, rr = r - 1;
    int k = 0;
    if (a != m && a != rr) {
      int d = a - rr;
      if (d < 0) d += m;
      k = d < m - d ? d : m - d;
    }
    return k < k0 ? -1 : k - k0;
  }
  __host__ __device__ int partner(int r, int a) const {
    if (r == 0) return a;
    const int m = n - 1, rr = r - 1;
    int b;
    if (a == m)
      b = rr;
    else if (a == rr)
      b = m;
    else {
      b = (2 * rr - a) % m;
      if (b < 0) b += m;
    }
    return b < S ? b : -1;
  }
  // end (exclusive) of the chunk of rounds that starts at r0: rounds are
  // added while the chunk holds at most max_blocks blocks and at most
  // pre * kThreads 16-byte values (one round at least)
  __host__ __device__ int chunk_end(int r0, int max_blocks, int KD4, int KB4, int pre) const {
    int r = r0, nb = 0, n4 = 0;
    while (r < rounds()) {
      const int rb = blocks(r), r4 = rb * (r == 0 ? KD4 : KB4);
      if (r > r0 && (nb + rb > max_blocks || n4 + r4 > pre * kThreads)) break;
      nb += rb;
      n4 += r4;
      ++r;
    }
    return r;
  }
};

// One (b, n2) sweep of a register tile: acc[r] += W[row r][n2] c_b[n2][m..m+3]
// for the NR rows of the tile (row stride RS, K stride KS in the staged block)
template <int N, int NR, int MT, int RS, int KS>
__device__ __forceinline__ void adj_tile_sweep(const float* __restrict__ g, const float* __restrict__ cb,
                                               f32x4 (&acc)[NR]) {
#pragma unroll
  for (int n2 = 0; n2 < N; ++n2) {
    const f32x4 cv = *reinterpret_cast<const f32x4*>(cb + n2 * MT);
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] += g[r * RS + n2 * KS] * cv;
  }
}

// PRE: 16-byte packed values per thread of one staged chunk.  PAIRS: the
// pair phase follows in the same workgroup (dL/dc in LDS); otherwise dL/dc
// [S][LL][N] of centre i is written to dcg + i S LL N for adjoint_pair_kernel
// and the workgroup needs no Y_lm table or pair batch (three per CU).
template <int N, int L, int NR, int ROLES, int PRE, bool DIRECT, bool PAIRS>
__global__ __launch_bounds__(kThreads, (PAIRS || ROLES > 1) ? 2 : 3) void adjoint_tile_kernel(
    RadialDev rad, int S, int PB, int max_blocks, PairView pv, const float* __restrict__ c,
    const double* __restrict__ norm2, int normalize, const double* __restrict__ xg, double xg_scale,
    const float* __restrict__ G, double* __restrict__ centre_virial, float* __restrict__ dcg) {
  constexpr int L1 = L + 1, LL = L1 * L1;
  constexpr int KB = N * N * L1, KD = N * (N + 1) / 2 * L1, KB4 = KB / 4, KD4 = KD / 4;
  constexpr int Q = adj_quads_before(L1), MT = 4 * Q, TILES = N / NR;
  constexpr int GRAM = adj_gram_before(L1);
  constexpr int CPT = (kWaves * ROLES * N * Q + kThreads - 1) / kThreads;  // c quads per thread
  static_assert(N % NR == 0 && TILES * Q <= 64, "one lane per register tile");
  static_assert(KB % 4 == 0 && KD % 4 == 0, "key blocks in whole 16-byte values");
  const float rt2 = 1.41421356237309504880f;
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  extern __shared__ __align__(64) unsigned char smem_raw[];
  float* yc = reinterpret_cast<float*>(smem_raw);
  float* c4 = reinterpret_cast<float*>(smem_raw + (PAIRS ? ylm_table_bytes<float>(L1) : 0));
  float* gram = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(c4) +
                                         align_up(sizeof(float) * S * N * MT, 64));
  unsigned char* region = reinterpret_cast<unsigned char*>(gram) + align_up(sizeof(float) * GRAM, 64);
  float* gl = reinterpret_cast<float*>(region);
  __shared__ int tri[N * (N + 1) / 2];  // n1 << 8 | n2 of the packed diagonal pairs
  __shared__ double red[kWaves];
  __shared__ double sh_dot;

  // this lane's register tile
  const bool lane_ok = lane < TILES * Q;
  const int ln = lane_ok ? lane : 0;
  const int tile = ln / Q, qq = ln - tile * Q;
  int l = 0;
#pragma unroll
  for (int t = 1; t <= L; ++t) l += qq >= adj_quads_before(t) ? 1 : 0;
  const int mq = qq - adj_quads_before(l);
  const int cq = 4 * qq;                 // first column of the quad in a padded c row
  const int gp = 4 * ((l >> 1) + 1);     // row length of C_l^T C_l
  int go = 0;                            // its first entry
#pragma unroll
  for (int t = 0; t < L; ++t) go += t < l ? (2 * t + 1) * 4 * ((t >> 1) + 1) : 0;

  const AdjRounds R(S);
  const FastDiv fper = fast_div(R.per > 0 ? R.per : 1);
  const int64_t Dq = packed_row_stride(S, N, L1);
  const float* Gi = G + i * Dq;

  // packed values of the chunk of rounds [r0, r1) held in registers.  The
  // blocks are numbered in round order (block j < S: the diagonal block
  // (j, j); then round r >= 1 from S + (r-1) per), and a chunk stages its
  // blocks in that order, so block j of a chunk starting at block jb takes
  // slot j - jb.  blk_src[j]: first packed column of block j.
  __shared__ int blk_src[2 * kWaves * (2 * kWaves + 1) / 2];
  auto first_block = [&](int r) { return r == 0 ? 0 : S + (r - 1) * R.per; };
  f32x4 pre[PRE];
  int info[PRE];  // slot * KB4 + q (bit 30: diagonal block), -1 none
  auto fetch = [&](int r0, int r1) {
    const int jb = first_block(r0);
    const int nd = r0 == 0 ? S : 0, nd4 = nd * KD4;
    const int tot = nd4 + (first_block(r1) - jb - nd) * KB4;
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
      const int u = tid + kThreads * k;
      info[k] = -1;
      if (u < tot) {
        int j, q;
        if (u < nd4) {
          j = u / KD4;
          q = u - j * KD4;
        } else {
          const int v = u - nd4;
          j = v / KB4;
          q = v - j * KB4;
          j += nd;
        }
        info[k] = (u < nd4 ? (1 << 30) : 0) | (j * KB4 + q);
        pre[k] = *reinterpret_cast<const f32x4*>(Gi + blk_src[jb + j] + 4 * q);
      }
    }
  };
  auto put = [&]() {
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
      const int d = info[k];
      if (d < 0) continue;
      if (!(d & (1 << 30))) {
        *reinterpret_cast<f32x4*>(gl + 4 * d) = pre[k];
      } else {
        const int e = d & ((1 << 30) - 1), slot = e / KB4, q = e - slot * KB4;
        float* w = gl + slot * KB;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int o = 4 * q + v, t = o / L1, lo = o - t * L1;
          const int n1 = tri[t] >> 8, n2 = tri[t] & 255;
          if (n1 == n2) {
            w[(n1 * N + n1) * L1 + lo] = rt2 * pre[k][v];
          } else {
            w[(n1 * N + n2) * L1 + lo] = pre[k][v];
            w[(n2 * N + n1) * L1 + lo] = pre[k][v];
          }
        }
      }
    }
  };

  // c into the padded rows, one quad (row, l, 4 m) per step; the loads of c
  // and of the first chunk are in flight together
  const int tot_q = S * N * Q;
  const float* ci = c + i * int64_t(S) * N * LL;
  f32x4 cv[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int u = tid + kThreads * k;
    cv[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (u < tot_q) {
      const int row = u / Q, qd = u - row * Q;
      int lc = 0;
#pragma unroll
      for (int t = 1; t <= L; ++t) lc += qd >= adj_quads_before(t) ? 1 : 0;
      const int m0 = 4 * (qd - adj_quads_before(lc));
      const float* src = ci + row * LL + lc * lc + m0;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (m0 + v < 2 * lc + 1) cv[k][v] = src[v];
    }
  }
  if (PAIRS) load_ylm_table<float>(rad, yc);
  for (int t = tid; t < N * (N + 1) / 2; t += kThreads) {
    int n1, n2;
    packed_pair(t, N, &n1, &n2);
    tri[t] = (n1 << 8) | n2;
  }
  for (int j = tid; j < first_block(R.rounds()); j += kThreads) {
    int a = j, b = j;
    if (j >= S) {
      const int jj = j - S, rr = jj / fper;
      R.pair(1 + rr, jj - rr * R.per + R.k0, &a, &b);
    }
    blk_src[j] = int(packed_key_offset(a, b, S, N, L1));
  }
  __syncthreads();  // the block table in place
  int r0 = 0, r1 = R.chunk_end(0, max_blocks, KD4, KB4, PRE);
  fetch(r0, r1);
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int u = tid + kThreads * k;
    if (u < tot_q) *reinterpret_cast<f32x4*>(c4 + 4 * u) = cv[k];
  }
  __syncthreads();  // c and tri in place

  // C_l^T C_l: unit (l, m', quad) of the table
  for (int u = tid; u < GRAM / 4; u += kThreads) {
    int lg = 0;
#pragma unroll
    for (int t = 1; t <= L; ++t) lg += u >= adj_gram_before(t) / 4 ? 1 : 0;
    int g0 = 0;
#pragma unroll
    for (int t = 0; t < L; ++t) g0 += t < lg ? (2 * t + 1) * ((t >> 1) + 1) : 0;
    const int qc = (lg >> 1) + 1, rem = u - g0, m2 = rem / qc, q2 = rem - m2 * qc;
    const float* cc = c4 + 4 * adj_quads_before(lg);
    f32x4 s = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int row = 0; row < S * N; ++row) {
      const float* cr = cc + row * MT;
      s += cr[m2] * *reinterpret_cast<const f32x4*>(cr + 4 * q2);
    }
    *reinterpret_cast<f32x4*>(gram + 4 * u) = s;
  }
  put();
  int r2 = r1 < R.rounds() ? R.chunk_end(r1, max_blocks, KD4, KB4, PRE) : r1;
  if (r1 < r2) fetch(r1, r2);
  __syncthreads();  // the table and the first chunk in place

  f32x4 accG[ROLES][NR], accP[ROLES][NR];
#pragma unroll
  for (int ro = 0; ro < ROLES; ++ro)
#pragma unroll
    for (int r = 0; r < NR; ++r) accG[ro][r] = accP[ro][r] = f32x4{0.f, 0.f, 0.f, 0.f};

  // C_l (C_l^T C_l)
#pragma unroll
  for (int ro = 0; ro < ROLES; ++ro) {
    const int a = wave + kWaves * ro;
    if (a >= S) continue;
    const float* ca = c4 + (a * N + tile * NR) * MT + 4 * adj_quads_before(l);
    const float* mb = gram + go + 4 * mq;
    for (int m2 = 0; m2 < 2 * l + 1; ++m2) {
      const f32x4 mv = *reinterpret_cast<const f32x4*>(mb + m2 * gp);
#pragma unroll
      for (int r = 0; r < NR; ++r) accP[ro][r] += ca[r * MT + m2] * mv;
    }
  }

  // W_l C_l over the chunks
  while (true) {
    const int rm0 = r0 == 0 ? 1 : r0, sd = r0 == 0 ? S : 0;
    for (int r = r0; r < r1; ++r) {
#pragma unroll
      for (int ro = 0; ro < ROLES; ++ro) {
        const int a = wave + kWaves * ro;
        if (a >= S) continue;
        const int k = R.block_of(r, a);
        if (k < 0) continue;
        const int b = R.partner(r, a);
        const int slot = r == 0 ? k : sd + (r - rm0) * R.per + k;
        const float* gb = gl + slot * KB + l;
        const float* cb = c4 + b * N * MT + cq;
        if (a <= b)
          adj_tile_sweep<N, NR, MT, N * L1, L1>(gb + tile * NR * N * L1, cb, accG[ro]);
        else
          adj_tile_sweep<N, NR, MT, L1, N * L1>(gb + tile * NR * L1, cb, accG[ro]);
      }
    }
    if (r1 >= R.rounds()) break;
    __syncthreads();  // every wavefront is done with this chunk
    put();
    r0 = r1;
    r1 = r2;
    r2 = r1 < R.rounds() ? R.chunk_end(r1, max_blocks, KD4, KB4, PRE) : r1;
    if (r1 < r2) fetch(r1, r2);
    __syncthreads();
  }

  // normalisation: alpha = 1/|p|, beta = (X . G)/|p|^2
  double alpha = 1.0, beta = 0.0;
  if (normalize) {
    const double n2 = norm2[i];
    double dotxg;
    if (xg) {
      dotxg = xg_scale * xg[i];
    } else {
      // p . G = 1/2 sum_l f_l <W_l C_l, C_l>, W_l C_l = sqrt 2 accG
      double d = 0;
      if (lane_ok) {
        const double fl = 1.0 / sqrt(double(2 * l + 1));
#pragma unroll
        for (int ro = 0; ro < ROLES; ++ro) {
          const int a = wave + kWaves * ro;
          if (a >= S) continue;
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const f32x4 cc = *reinterpret_cast<const f32x4*>(c4 + (a * N + tile * NR + r) * MT + cq);
#pragma unroll
            for (int v = 0; v < 4; ++v) d += fl * double(accG[ro][r][v]) * double(cc[v]);
          }
        }
      }
      d *= 0.5 * 1.41421356237309504880;
      for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, 64);
      if (lane == 0) red[wave] = d;
      __syncthreads();
      if (tid == 0) {
        double t = 0;
        for (int w = 0; w < kWaves; ++w) t += red[w];
        sh_dot = t;
      }
      __syncthreads();
      dotxg = n2 > 0 ? sh_dot / sqrt(n2) : 0.0;
    }
    if (n2 > 0) {
      alpha = 1.0 / sqrt(n2);
      beta = dotxg / n2;
    }
  }
  __syncthreads();  // every read of c and of the staged blocks is done
  float* dc = c4;   // [S][LL][N]
  if (lane_ok) {
    const float fl = 1.0f / sqrtf(float(2 * l + 1));
    const float sg = float(alpha) * rt2 * fl, sp = float(-2.0 * beta) * fl * fl;
#pragma unroll
    for (int ro = 0; ro < ROLES; ++ro) {
      const int a = wave + kWaves * ro;
      if (a >= S) continue;
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = 4 * mq + v;
          if (m < 2 * l + 1)
            dc[(a * LL + l * l + m) * N + tile * NR + r] = sg * accG[ro][r][v] + sp * accP[ro][r][v];
        }
    }
  }
  __syncthreads();
  if constexpr (PAIRS) {
    adjoint_pairs<float, DIRECT>(rad, S, PB, pv, yc, dc, region, i, centre_virial);
  } else {
    (void)PB;
    (void)centre_virial;
    static_assert((LL * N) % 4 == 0, "dL/dc rows in whole 16-byte values");
    const int E4 = S * LL * N / 4;
    f32x4* dst = reinterpret_cast<f32x4*>(dcg + i * int64_t(S) * LL * N);
    for (int e = tid; e < E4; e += kThreads) dst[e] = reinterpret_cast<const f32x4*>(dc)[e];
  }
}

// Real spherical harmonics Y_l^{+-m} (l = m..L) of one unit direction on the
// lane of order m (0 <= m <= L), written to Y[(l*l + l +- m) * ys]: the
// recursion of ylm_wave without the gradient (no cross-lane operations).
__device__ __forceinline__ void ylm_lane(int L, const float* __restrict__ yc, float x, float y, float z, int m,
                                         float* Y, int ys) {
  const int L1 = L + 1;
  const float* dfac = yc + 4 * L1 * L1;
  const float* s2l1 = dfac + L1;
  float n2 = x * x + y * y + z * z;
  if (fabs(double(n2) - 1.0) > 1e-6) {
    const float inv = 1.0f / sqrtf(n2);
    x *= inv;
    y *= inv;
    z *= inv;
  }
  const float cos_t = z;
  const float sxy = sqrtf(x * x + y * y);
  float cos_p = 1.0f, sin_p = 0.0f;
  if (double(sxy) >= kDblFtol) {
    cos_p = x / sxy;
    sin_p = y / sxy;
  }
  const float sin_t = sqrtf(1.0f - cos_t * cos_t);
  float c_m = 1.0f, s_m = 0.0f, c_m1 = cos_p, s_m1 = sin_p;
  for (int k = 1; k <= m; ++k) {
    const float cn = 2.0f * cos_p * c_m1 - c_m;
    const float sn = 2.0f * cos_p * s_m1 - s_m;
    c_m = c_m1;
    s_m = s_m1;
    c_m1 = cn;
    s_m1 = sn;
  }
  const float s2pi = 0.39894228040143267794f;
  float diag = s2pi;
  for (int k = 1; k <= m; ++k) diag = diag * sin_t * dfac[k];
  const float isq2 = 0.70710678118654752440f;
  float P1 = 0.0f, P2 = 0.0f;
  for (int l = m; l <= L; ++l) {
    float P;
    if (m == l) {
      P = diag;
    } else if (m == l - 1) {
      P = (l == 1) ? cos_t * 1.73205080756887729353f * s2pi : diag * cos_t * s2l1[l];
    } else {
      const Vec4<float>::type c4 = reinterpret_cast<const Vec4<float>::type*>(yc)[l * L1 + m];
      float t = cos_t * P1;
      t += c4.y * P2;
      t *= c4.x;
      P = t;
    }
    const int base = l * l + l;
    if (m == 0) {
      Y[base * ys] = P * isq2;
    } else {
      Y[(base + m) * ys] = P * c_m;
      Y[(base - m) * ys] = P * s_m;
    }
    P2 = P1;
    P1 = P;
  }
}

// ---------------------------------------------------------------------------
// Pair phase of the mixed adjoint as a kernel of its own, after
// adjoint_tile_kernel wrote dL/dc [S][LL][N] per centre: one workgroup per
// centre, a group of 16 lanes per pair (16 pairs in flight).  With D[lm] =
// sum_n dL/dc[s_j][lm][n] R_nl, the pair gradient of adjoint_pairs splits
// over the radial channels: lane n of the group evaluates R_nl and R'_nl of
// its channel for every l (spline) and accumulates
//   A_n  = sum_l (R'_nl fc + R_nl fc') sum_m dc[lm][n] Y_lm,
//   Sl_n = sum_l l R_nl sum_m dc[lm][n] Y_lm,
//   B_n  = sum_l R_nl sum_m dc[lm][n] gY_lm   (gY: the gradient terms of Y),
// with {Y, gY}[lm] of the pair staged once per group in LDS; the group sums
// over n and its lane 0 forms the force in FP64 as adjoint_pairs does.  No
// barrier between pairs (a group lives in one wavefront) and a small LDS
// footprint keep several workgroups per CU in flight.
constexpr int kPairGroup = 16;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ inline size_t pair_kernel_lds(int S, int N, int L1) {
  const int LL = L1 * L1;
  return ylm_table_bytes<float>(L1) + align_up(sizeof(YGradTerm<float>) * 3 * LL, 64) +
         align_up(sizeof(float) * S * LL * N, 64) + sizeof(float) * 4 * LL * (kThreads / kPairGroup);
}

// FP32 spline cell of a pair at distance d whose neighbour has species sj
// (the cell search of pair_fetch, per-species tables included)
__device__ __forceinline__ SplineCell<float> pair_cell(const RadialDev& rad, double d, int sj) {
  int klo;
  double xhi, h;
  if (rad.sp_tab) {
    const int t = rad.sp_tab[sj], last = rad.tab_n[t] - 2;
    int k = int((d - rad.x1) * rad.tab_pu[t]) - 1;
    k = k < 0 ? 0 : (k > last ? last : k);
    xhi = rad.tab_grid[rad.tab_grid0[t] + k + 1];
    h = rad.tab_h[t];
    klo = rad.tab_cell0[t] + k;
  } else {
    int k = int((d - rad.x1) * rad.per_unit) - 1;
    k = k < 0 ? 0 : (k > rad.search_size ? rad.search_size : k);
    xhi = rad.grid[k + 1];
    h = rad.h;
    klo = k;
  }
  const double ca = (xhi - d) / h;
  SplineCell<float> cell;
  cell.klo = klo;
  cell.a = float(ca);
  cell.b = float(1 - ca);
  cell.inv_h = rad.sp_tab ? float(1.0 / h) : float(1.0 / rad.h);
  return cell;
}

// l of the packed index lm = l*l + l + m, for compile-time lm
__host__ __device__ constexpr int l_of_lm_c(int lm) {
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) ++l;
  return l;
}

// ---------------------------------------------------------------------------
// Expansion of the mixed path on pair groups (FP32 spline radial functions,
// compile-time L).  One workgroup per centre; wavefront w takes the
// neighbours of species s = w, w + 4, ... in list order, four at a time, one
// per group of 16 lanes.  Lane n of a group evaluates f_c R_nl of its radial
// channel for every l; the group's lanes m <= L form Y_lm of the pair
// (ylm_lane) in LDS; every lane then adds f_c R_nl Y_lm to its registers
// c[s][n][lm] for all lm.  The four groups' sums are added with shuffles in
// a fixed order (so the result does not depend on timing), the species rows
// pass through LDS and are written once, coalesced, with the centre term.
// The same sums as expansion_kernel (calculator_spherical_expansion.hh
// :2139-2300), accumulated in another order.
template <int L>
__global__ __launch_bounds__(kThreads) void expansion_group_kernel(RadialDev rad, int S, PairView pv,
                                                                   float* __restrict__ cout,
                                                                   unsigned long long* __restrict__ keymask) {
  constexpr int L1 = L + 1, LL = L1 * L1, LLP = (LL + 3) & ~3;
  constexpr int G = 64 / kPairGroup;  // groups per wavefront
  const int N = rad.N;
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int gl = lane & (kPairGroup - 1), grp = lane / kPairGroup;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  float* yc = reinterpret_cast<float*>(smem_raw);
  float* yg = reinterpret_cast<float*>(smem_raw + ylm_table_bytes<float>(L1));  // [waves][G][LLP]
  float* cl = yg + kWaves * G * LLP;                                           // [S][N][LL]
  __shared__ unsigned long long kmask;
  const int si = pv.species_of[i];
  load_ylm_table<float>(rad, yc);
  if (tid == 0) kmask = 1ull << si;
  __syncthreads();
  const int64_t p0 = pv.centre_off[i] + 1;  // the self pair contributes the centre term only
  const int np = int(pv.centre_off[i + 1] - p0);
  float* Yg = yg + (wave * G + grp) * LLP;
  const int n = gl < N ? gl : N - 1;
  for (int s = wave; s < S; s += kWaves) {
    float acc[LL];
#pragma unroll
    for (int lm = 0; lm < LL; ++lm) acc[lm] = 0.f;
    bool any = false;
    for (int b = 0; b < np; b += 64) {
      const int t = b + lane;
      const bool mine = t < np && pv.species_of[pv.pair_j[p0 + t]] == s;
      unsigned long long m = __ballot(mine);
      any = any || m != 0;
      while (m) {
        // the next G pairs of this species, one per group
        int my = -1;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (m) {
            const int bit = __ffsll((long long)m) - 1;
            if (g == grp) my = bit;
            m &= m - 1;
          }
        }
        if (my >= 0) {
          const int64_t p = p0 + b + my;
          const double d = pv.pair_dist[p];
          const double fcd = pv.pair_fc[p];
          const SplineCell<float> cell = pair_cell(rad, d, s);
          float R[L1];
          const float fc = float(fcd);
#pragma unroll
          for (int l = 0; l <= L; ++l) {
            float v;
            spline_eval<float>(rad, cell, n * L1 + l, &v, nullptr);
            R[l] = v * fc;
          }
          if (gl <= L)
            ylm_lane(L, yc, float(pv.pair_vec[3 * p] / d), float(pv.pair_vec[3 * p + 1] / d),
                     float(pv.pair_vec[3 * p + 2] / d), gl, Yg, 1);
          wave_lds_sync();
#pragma unroll
          for (int q = 0; q < LLP / 4; ++q) {
            const f32x4 y4 = *reinterpret_cast<const f32x4*>(Yg + 4 * q);
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (4 * q + u < LL) acc[4 * q + u] += R[l_of_lm_c(4 * q + u)] * y4[u];
          }
          wave_lds_sync();  // Y is read before the group's next pair overwrites it
        }
      }
    }
    // the groups' partial sums, (g0 + g1) + (g2 + g3)
#pragma unroll
    for (int lm = 0; lm < LL; ++lm) {
      acc[lm] += __shfl_xor(acc[lm], kPairGroup, 64);
      acc[lm] += __shfl_xor(acc[lm], 2 * kPairGroup, 64);
    }
    if (grp == 0 && gl < N) {
      float* row = cl + (s * N + gl) * LL;
#pragma unroll
      for (int lm = 0; lm < LL; ++lm) row[lm] = acc[lm];
    }
    if (lane == 0 && any) atomicOr(&kmask, 1ull << s);
  }
  __syncthreads();
  const int E = S * N * LL;
  const double inv_s4pi = 0.28209479177387814347;  // 1/sqrt(4 pi)
  const double* centre = rad.center + (rad.sp_tab ? rad.sp_tab[si] * N : 0);
  float* ci = cout + i * int64_t(E);
  for (int e = tid; e < E; e += kThreads) {
    float v = cl[e];
    const int row = e / LL;
    if (row / N == si && e - row * LL == 0) v += float(centre[row - si * N] * inv_s4pi);
    ci[e] = v;
  }
  if (tid == 0) keymask[i] = kmask;
}

__host__ __device__ inline size_t expansion_group_lds(int S, int N, int L1) {
  const int LL = L1 * L1, LLP = (LL + 3) & ~3;
  return ylm_table_bytes<float>(L1) + sizeof(float) * (size_t(kWaves) * (64 / kPairGroup) * LLP +
                                                       size_t(S) * N * LL);
}

template <int N, int L>
__global__ __launch_bounds__(kThreads) void adjoint_pair_kernel(RadialDev rad, int S, PairView pv,
                                                                const float* __restrict__ dcg,
                                                                double* __restrict__ centre_virial) {
  constexpr int L1 = L + 1, LL = L1 * L1, NG = kThreads / kPairGroup;
  static_assert(N <= kPairGroup && L1 <= kPairGroup, "one lane per radial channel and per order m");
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int gl = lane & (kPairGroup - 1), grp = tid / kPairGroup;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  float* yc = reinterpret_cast<float*>(smem_raw);
  YGradTerm<float>* gt = reinterpret_cast<YGradTerm<float>*>(smem_raw + ylm_table_bytes<float>(L1));
  float* dc = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(gt) +
                                       align_up(sizeof(YGradTerm<float>) * 3 * LL, 64));
  f32x4* yg = reinterpret_cast<f32x4*>(reinterpret_cast<unsigned char*>(dc) +
                                       align_up(sizeof(float) * S * LL * N, 64));
  __shared__ double red[kWaves][9];
  load_ylm_table<float>(rad, yc);
  {
    const YGradTerm<float>* src = ygrad_src<float>(rad);
    for (int e = tid; e < 3 * LL; e += kThreads) gt[e] = src[e];
    const int E4 = S * LL * N / 4;
    const f32x4* d4 = reinterpret_cast<const f32x4*>(dcg + i * int64_t(S) * LL * N);
    for (int e = tid; e < E4; e += kThreads) reinterpret_cast<f32x4*>(dc)[e] = d4[e];
  }
  __syncthreads();
  const int64_t p0 = pv.centre_off[i] + 1;
  const int np = int(pv.centre_off[i + 1] - p0);
  f32x4* Yg = yg + grp * LL;
  double fi[3] = {0, 0, 0}, v6[6] = {0, 0, 0, 0, 0, 0};
  for (int q0 = 0; q0 < np; q0 += NG) {
    const int q = q0 + grp;
    if (q < np) {  // uniform within the group
      const int64_t p = p0 + q;
      const double d = pv.pair_dist[p];
      const double vx = pv.pair_vec[3 * p], vy = pv.pair_vec[3 * p + 1], vz = pv.pair_vec[3 * p + 2];
      const double fcd = pv.pair_fc[p], dfcd = pv.pair_dfc[p];
      const int j = pv.pair_j[p];
      const int sj = pv.species_of[j];
      const SplineCell<float> cell = pair_cell(rad, d, sj);
      float R[L1], dR[L1];
      const int n = gl < N ? gl : N - 1;
#pragma unroll
      for (int l = 0; l <= L; ++l) spline_eval<float>(rad, cell, n * L1 + l, &R[l], &dR[l]);
      const float dx = float(vx / d), dy = float(vy / d), dz = float(vz / d);
      if (gl <= L) ylm_lane(L, yc, dx, dy, dz, gl, reinterpret_cast<float*>(Yg), 4);
      wave_lds_sync();
      for (int lm = gl; lm < LL; lm += kPairGroup) {
        const YGradTerm<float>* g = gt + 3 * lm;
        const float gx = g[0].c0 * Yg[g[0].i0].x + g[0].c1 * Yg[g[0].i1].x;
        const float gy = g[1].c0 * Yg[g[1].i0].x + g[1].c1 * Yg[g[1].i1].x;
        const float gz = g[2].c0 * Yg[g[2].i0].x + g[2].c1 * Yg[g[2].i1].x;
        float* o = reinterpret_cast<float*>(Yg + lm);
        o[1] = gx;
        o[2] = gy;
        o[3] = gz;
      }
      wave_lds_sync();
      float A = 0.f, Sl = 0.f, Bx = 0.f, By = 0.f, Bz = 0.f;
      if (gl < N) {
        const float ffc = float(fcd), fdfc = float(dfcd);
        const float* dcs = dc + sj * LL * N + gl;
#pragma unroll
        for (int l = 0; l <= L; ++l) {
          float s1 = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
          for (int m = 0; m < 2 * l + 1; ++m) {
            const int lm = l * l + m;
            const float t = dcs[lm * N];
            const f32x4 y4 = Yg[lm];
            s1 += t * y4.x;
            sx += t * y4.y;
            sy += t * y4.z;
            sz += t * y4.w;
          }
          A += s1 * (dR[l] * ffc + R[l] * fdfc);
          Sl += s1 * (float(l) * R[l]);
          Bx += sx * R[l];
          By += sy * R[l];
          Bz += sz * R[l];
        }
      }
#pragma unroll
      for (int off = kPairGroup / 2; off > 0; off >>= 1) {
        A += __shfl_xor(A, off, kPairGroup);
        Sl += __shfl_xor(Sl, off, kPairGroup);
        Bx += __shfl_xor(Bx, off, kPairGroup);
        By += __shfl_xor(By, off, kPairGroup);
        Bz += __shfl_xor(Bz, off, kPairGroup);
      }
      if (gl == 0) {
        const double sc = fcd / d;
        const double gv[3] = {double(A) * vx / d + sc * double(Bx - dx * Sl),
                              double(A) * vy / d + sc * double(By - dy * Sl),
                              double(A) * vz / d + sc * double(Bz - dz * Sl)};
        double* pf = pv.pair_force + 3 * p;
        pf[0] = gv[0];
        pf[1] = gv[1];
        pf[2] = gv[2];
        if (j != int(i)) {
          fi[0] -= gv[0];
          fi[1] -= gv[1];
          fi[2] -= gv[2];
        }
        // r_ji = r_i - r_j(image) = -pair vector
        v6[0] -= vx * gv[0];
        v6[1] -= vy * gv[1];
        v6[2] -= vz * gv[2];
        v6[3] -= vy * gv[2];
        v6[4] -= vz * gv[0];
        v6[5] -= vx * gv[1];
      }
      wave_lds_sync();  // the group's {Y, gY} are read before the next pair overwrites them
    }
  }
  double r[9] = {fi[0], fi[1], fi[2], v6[0], v6[1], v6[2], v6[3], v6[4], v6[5]};
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    for (int off = 32; off > 0; off >>= 1) r[k] += __shfl_down(r[k], off, 64);
    if (lane == 0) red[wave][k] = r[k];
  }
  __syncthreads();
  if (tid < 9) {
    double t = 0;
    for (int w = 0; w < kWaves; ++w) t += red[w][tid];
    if (tid < 3) {
      pv.centre_force[3 * i + tid] = t;
    } else {
      centre_virial[6 * i + tid - 3] = t;
    }
  }
}

// dL/dX of one centre per workgroup into the packed layout (packing.hh)
__global__ void pack_gradient_kernel(int S, int N, int L1, const double* __restrict__ G,
                                     float* __restrict__ Gp) {
  const int64_t i = blockIdx.x;
  const int K = S * (S + 1) / 2, KB = N * N * L1, KD = packed_diag_block(N, L1);
  const double* g = G + i * int64_t(K) * KB;
  float* gp = Gp + i * packed_row_stride(S, N, L1);
  for (int k = 0; k < K; ++k) {
    int a, b;
    key_pair(k, S, &a, &b);
    const double* src = g + int64_t(k) * KB;
    float* dst = gp + packed_key_offset(a, b, S, N, L1);
    if (a < b) {
      for (int o = threadIdx.x; o < KB; o += blockDim.x) dst[o] = float(src[o]);
    } else {
      for (int o = threadIdx.x; o < KD; o += blockDim.x) {
        const int t = o / L1, l = o - t * L1;
        int n1, n2;
        packed_pair(t, N, &n1, &n2);
        dst[o] = n1 < n2 ? float((src[(n1 * N + n2) * L1 + l] + src[(n2 * N + n1) * L1 + l]) *
                                 0.70710678118654752440)
                         : float(src[(n1 * N + n1) * L1 + l]);
      }
    }
  }
}

__global__ void fill_f64_kernel(double* __restrict__ p, int64_t n, double v) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x)
    p[e] = v;
}

// n_pairs is read on the device (the last entry of the pair offsets), so one
// launch shape (the pair capacity) serves every neighbour-list rebuild
__global__ void pair_cutoff_kernel(RadialDev rad, const int64_t* __restrict__ n_pairs, int64_t cap,
                                   const double* __restrict__ dist, double* __restrict__ fc,
                                   double* __restrict__ dfc) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= cap || p >= *n_pairs) return;
  double f, df;
  cutoff_eval(rad, dist[p], &f, &df);
  fc[p] = f;
  dfc[p] = df;
}

// per-structure (segment) sums of per-centre rows of width w <= 8, times
// scale[s] when given (+=)
constexpr int kSegThreads = 1024;
struct SegJob {
  const double* in;
  double* out;
  const double* scale;
  int w;
};

// One workgroup per (job, segment): blocks [0, n_seg) run job 0, blocks
// [n_seg, 2 n_seg) job 1.  Each thread takes rows a, a + T, a + 2T, ... in
// that order, four rows' loads issued before their adds; a fixed tree sums
// the threads (the same bits on every run).
__global__ __launch_bounds__(kSegThreads) void segment_sum_kernel(SegJob j0, SegJob j1,
                                                                  const int64_t* __restrict__ seg_off, int n_seg) {
  const bool second = int(blockIdx.x) >= n_seg;
  const int s = second ? int(blockIdx.x) - n_seg : int(blockIdx.x);
  const double* __restrict__ in = second ? j1.in : j0.in;
  double* __restrict__ out = second ? j1.out : j0.out;
  const double* __restrict__ scale = second ? j1.scale : j0.scale;
  const int w = second ? j1.w : j0.w;
  __shared__ double red[kSegThreads / 64][8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t a1 = seg_off[s + 1];
  // unrolled over the largest width (uniform guards): v stays in registers
  for (int64_t a = seg_off[s] + threadIdx.x; a < a1; a += 4 * kSegThreads) {
    double x[4][8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t r = a + int64_t(u) * kSegThreads;
#pragma unroll
      for (int k = 0; k < 8; ++k) x[u][k] = (k < w && r < a1) ? in[int64_t(w) * r + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] += x[u][k];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < w)
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < w) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < w) {
    double t = 0;
    for (int x = 0; x < kSegThreads / 64; ++x) t += red[x][threadIdx.x];
    out[int64_t(w) * s + threadIdx.x] += scale ? t * scale[s] : t;
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
  radial_n(rad, n, r[pt], v, dout ? d : nullptr);
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

// neighbour radial contribution as used by the expansion (spline)
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
__global__ __launch_bounds__(kThreads) void ylm_eval_kernel(RadialDev rad, const double* __restrict__ dirs,
                                                            int64_t n, double* __restrict__ Y,
                                                            double* __restrict__ dY) {
  extern __shared__ __align__(64) unsigned char smem_raw[];
  const int L = rad.L, L1 = L + 1, LL = L1 * L1;
  const int per_wave = 64 / L1, per_block = kWaves * per_wave;
  T* yc = reinterpret_cast<T*>(smem_raw);
  T* dsh = reinterpret_cast<T*>(smem_raw + ylm_table_bytes<T>(L1));
  T* Ysh = dsh + 3 * per_block;
  T* dYsh = Ysh + per_block * LL;
  load_ylm_table<T>(rad, yc);
  const int64_t q0 = int64_t(blockIdx.x) * per_block;
  const int nb = int(min<int64_t>(per_block, n - q0));
  for (int t = threadIdx.x; t < 3 * nb; t += kThreads) dsh[t] = T(dirs[3 * q0 + t]);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  ylm_wave<T>(L, yc, dsh, wave * per_wave, min(nb, (wave + 1) * per_wave), Ysh, dY ? dYsh : nullptr, LL,
              3 * LL);
  __syncthreads();
  for (int t = threadIdx.x; t < nb * LL; t += kThreads) Y[q0 * LL + t] = double(Ysh[t]);
  if (dY)
    for (int t = threadIdx.x; t < nb * 3 * LL; t += kThreads) dY[q0 * 3 * LL + t] = double(dYsh[t]);
}

// ---------------------------------------------------------------------------
// Half neighbour list (adaptor_half_neighbour_list.hh:319-380: a pair i -> j
// is kept when tag_j > tag_i, periodic images carrying the largest tags)
// and the expansion over it (calculator_spherical_expansion.hh:2282-2299:
// c^{j, s_i} += (-1)^l c^{ij}).  Built from the full list on the device:
// the kept pairs are the self pair, the original neighbours j > i and every
// periodic image; a dropped pair (original j < i) takes the row of its
// mirror (j -> i, kept in j's list) with the parity (-1)^l, since
// Y_lm(-r) = (-1)^l Y_lm(r) and R depends on |r| only.

// pair p of centre i is the original neighbour j (no periodic shift): its
// vector is exactly r_j - r_i, as the neighbour list forms it for originals
__device__ __forceinline__ bool pair_original(const PairView& pv, const double* __restrict__ pos,
                                              int64_t p, int64_t i, int j) {
  return pv.pair_vec[3 * p] == pos[3 * j] - pos[3 * i] &&
         pv.pair_vec[3 * p + 1] == pos[3 * j + 1] - pos[3 * i + 1] &&
         pv.pair_vec[3 * p + 2] == pos[3 * j + 2] - pos[3 * i + 2];
}

__device__ __forceinline__ bool half_keep(const PairView& pv, const double* __restrict__ pos,
                                          int64_t p, int64_t i) {
  if (p == pv.centre_off[i]) return true;  // the self pair
  const int j = pv.pair_j[p];
  return j > int(i) || !pair_original(pv, pos, p, i, j);
}

__global__ void half_count_kernel(PairView pv, const double* __restrict__ pos, int* __restrict__ count) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= pv.n_centres) return;
  int k = 0;
  for (int64_t p = pv.centre_off[i]; p < pv.centre_off[i + 1]; ++p) k += half_keep(pv, pos, p, i);
  count[i] = k;
}

// the kept pairs gathered in list order (a CSR by centre, self pair
// first), and for every full pair the row it reads: src >= 0 its own kept
// row, src < 0 the row -1 - src of its mirror
__global__ void half_fill_kernel(PairView pv, const double* __restrict__ pos,
                                 const int64_t* __restrict__ hoff, int64_t* __restrict__ src,
                                 int* __restrict__ hj, double* __restrict__ hvec,
                                 double* __restrict__ hdist, double* __restrict__ hfc,
                                 double* __restrict__ hdfc) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= pv.n_centres) return;
  int64_t h = hoff[i];
  for (int64_t p = pv.centre_off[i]; p < pv.centre_off[i + 1]; ++p) {
    if (!half_keep(pv, pos, p, i)) continue;
    src[p] = h;
    hj[h] = pv.pair_j[p];
    hvec[3 * h] = pv.pair_vec[3 * p];
    hvec[3 * h + 1] = pv.pair_vec[3 * p + 1];
    hvec[3 * h + 2] = pv.pair_vec[3 * p + 2];
    hdist[h] = pv.pair_dist[p];
    hfc[h] = pv.pair_fc[p];
    hdfc[h] = pv.pair_dfc[p];
    ++h;
  }
}

__global__ void half_mirror_kernel(PairView pv, const double* __restrict__ pos, int64_t* __restrict__ src) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= pv.n_centres) return;
  for (int64_t p = pv.centre_off[i] + 1; p < pv.centre_off[i + 1]; ++p) {
    if (half_keep(pv, pos, p, i)) continue;
    const int j = pv.pair_j[p];
    int64_t m = INT64_MIN;
    for (int64_t q = pv.centre_off[j] + 1; q < pv.centre_off[j + 1]; ++q)
      if (pv.pair_j[q] == int(i) && pair_original(pv, pos, q, j, int(i))) {
        m = -1 - src[q];
        break;
      }
    src[p] = m;
  }
}

// rows[h][n][lm] = fc R_nl Y_lm of the kept pairs (not the self pairs), one
// workgroup per centre over its kept pairs (hv: the half list as a
// PairView)
template <class T, bool DIRECT>
__global__ __launch_bounds__(kThreads) void half_rows_kernel(RadialDev rad, int S, int PB, PairView hv,
                                                             T* __restrict__ rows) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, F = N * L1, B = N * LL;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  PairBatch<T> pb = carve_batch<T, DIRECT, false>(smem_raw + ylm_table_bytes<T>(L1), PB, LL, F);
  load_ylm_table<T>(rad, yc);
  const int64_t p0 = hv.centre_off[i] + 1;
  const int np = int(hv.centre_off[i + 1] - p0);
  const FastDiv fB = fast_div(B), fLL = fast_div(LL);
  PairPre pre;
  pair_fetch<DIRECT, false>(rad, hv, p0, max(0, min(PB, np)), pre);
  for (int q0 = 0; q0 < np; q0 += PB) {
    const int nb = min(PB, np - q0);
    stage_pairs<T, DIRECT, false>(rad, yc, pre, nb, pb, nullptr, nullptr, nullptr);
    if (q0 + PB < np) pair_fetch<DIRECT, false>(rad, hv, p0 + q0 + PB, min(PB, np - q0 - PB), pre);
    for (int it = threadIdx.x; it < nb * B; it += kThreads) {
      const int q = it / fB, e = it - q * B, n = e / fLL, lm = e - n * LL;
      rows[(p0 + q0 + q) * B + e] = pb.R[q * F + n * L1 + l_of_lm(lm)] * pb.Y[q * LL + lm];
    }
    __syncthreads();  // the next batch overwrites this one
  }
}

// c[i][s][n][lm] = centre term + sum over the full list of centre i, in
// list order, of the row of each pair (its own, or its mirror's times
// (-1)^l) into the block of the neighbour species; and the key mask
template <class T>
__global__ __launch_bounds__(kThreads) void half_gather_kernel(RadialDev rad, int S, PairView pv,
                                                               const int64_t* __restrict__ src,
                                                               const T* __restrict__ rows,
                                                               T* __restrict__ c,
                                                               unsigned long long* __restrict__ keymask) {
  const int64_t i = blockIdx.x;
  const int N = rad.N, LL = rad.LL, B = N * LL;
  constexpr int kChunk = kThreads;
  __shared__ int sp[kChunk];
  __shared__ int64_t row[kChunk];
  __shared__ unsigned long long kmask;
  const int si = pv.species_of[i];
  const double inv_s4pi = 0.28209479177387814347;  // 1/sqrt(4 pi)
  const int64_t p0 = pv.centre_off[i] + 1, p1 = pv.centre_off[i + 1];
  if (threadIdx.x == 0) kmask = 1ull << si;
  constexpr int kE = 4;  // entries per thread held across the pair chunks
  for (int e0 = 0; e0 < S * B; e0 += kE * kThreads) {
    T acc[kE];
    int es[kE], er[kE], el[kE];
#pragma unroll
    for (int k = 0; k < kE; ++k) {
      const int e = e0 + threadIdx.x + k * kThreads;
      es[k] = e < S * B ? e / B : -1;
      er[k] = e - es[k] * B;
      const int lm = er[k] % LL;
      el[k] = l_of_lm(lm);
      acc[k] = T(0);
      if (es[k] == si && lm == 0)
        acc[k] = T(rad.center[(rad.sp_tab ? rad.sp_tab[si] * rad.N : 0) + er[k] / LL] * inv_s4pi);
    }
    for (int64_t c0 = p0; c0 < p1; c0 += kChunk) {
      const int nq = int(min<int64_t>(kChunk, p1 - c0));
      __syncthreads();  // the previous chunk is read
      if (int(threadIdx.x) < nq) {
        const int64_t p = c0 + threadIdx.x;
        sp[threadIdx.x] = pv.species_of[pv.pair_j[p]];
        row[threadIdx.x] = src[p];
        if (e0 == 0) atomicOr(&kmask, 1ull << sp[threadIdx.x]);
      }
      __syncthreads();
      for (int q = 0; q < nq; ++q) {
        const int s = sp[q];
        const int64_t r = row[q] >= 0 ? row[q] : -1 - row[q];
        const bool mirror = row[q] < 0;
#pragma unroll
        for (int k = 0; k < kE; ++k)
          if (es[k] == s) {
            const T v = rows[r * B + er[k]];
            acc[k] += (mirror && (el[k] & 1)) ? -v : v;
          }
      }
    }
#pragma unroll
    for (int k = 0; k < kE; ++k)
      if (es[k] >= 0) c[i * int64_t(S) * B + es[k] * B + er[k]] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) keymask[i] = kmask;
}

// ---------------------------------------------------------------------------
// Stored representation gradients (FP64), as the reference keeps them for
// get_features_gradient.  Step 1, per-pair expansion derivatives
// (calculator_spherical_expansion.hh:2306-2401): for the pair p = (i, j) that
// is not the self pair,
//   dcp[p][x][n][lm] = d c^{i, s_j}[n][lm] / d r_j,x
//                    = (R' fc + R fc') Y r_hat_x + R fc dY_x / r
// (dY_x from the two-term identity of the adjoint), and for the centre
//   dcs[i][x][s][n][lm] = - sum of dcp over its pairs of species s that are
// not periodic images of i itself.  One workgroup per centre; each thread
// owns fixed (x, n, lm) entries, so it alone updates their centre sums.
template <bool DIRECT>
__global__ __launch_bounds__(kThreads) void pair_dc_kernel(RadialDev rad, int S, int PB, PairView pv,
                                                           double* __restrict__ dcp,
                                                           double* __restrict__ dcs) {
  typedef double T;
  const int64_t i = blockIdx.x;
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, F = N * L1, B = N * LL;
  const int tid = threadIdx.x;
  extern __shared__ __align__(64) unsigned char smem_raw[];
  T* yc = reinterpret_cast<T*>(smem_raw);
  PairBatch<T> pb = carve_batch<T, DIRECT, true>(smem_raw + ylm_table_bytes<T>(L1), PB, LL, F);
  load_ylm_table<T>(rad, yc);
  {
    const YGradTerm<T>* src = ygrad_src<T>(rad);
    for (int e = tid; e < 3 * LL; e += kThreads) pb.gt[e] = src[e];
  }
  double* ds = dcs + i * int64_t(3) * S * B;
  for (int e = tid; e < 3 * S * B; e += kThreads) ds[e] = 0.0;
  const int64_t p0 = pv.centre_off[i] + 1;  // the self pair has no derivative of its own
  const int np = int(pv.centre_off[i + 1] - p0);
  const FastDiv fB = fast_div(B), fLL = fast_div(LL);
  PairPre pre;
  pair_fetch<DIRECT, true>(rad, pv, p0, max(0, min(PB, np)), pre);
  for (int q0 = 0; q0 < np; q0 += PB) {
    const int nb = min(PB, np - q0);
    stage_pairs<T, DIRECT, true>(rad, yc, pre, nb, pb, nullptr, nullptr, nullptr);
    if (q0 + PB < np) pair_fetch<DIRECT, true>(rad, pv, p0 + q0 + PB, min(PB, np - q0 - PB), pre);
    for (int e = tid; e < 3 * B; e += kThreads) {
      const int x = e / fB, rr = e - x * B, n = rr / fLL, lm = rr - n * LL;
      const int l = l_of_lm(lm);
      const YGradTerm<T> g = pb.gt[3 * lm + x];
      for (int q = 0; q < nb; ++q) {
        const T* Yq = pb.Y + q * LL;
        const double rx = pb.dir[3 * q + x];
        const double fc = pb.fc[q], dfc = pb.dfc[q];
        const double R = pb.R[q * F + l * N + n], dR = pb.dR[q * F + l * N + n];
        const double dy = g.c0 * Yq[g.i0] + g.c1 * Yq[g.i1] - double(l) * rx * Yq[lm];
        const double v = (dR * fc + R * dfc) * Yq[lm] * rx + R * fc * dy / pb.dist[q];
        dcp[(p0 + q0 + q) * 3 * B + e] = v;
        if (pb.pj[q] != int(i)) ds[(x * S + pb.sp[q]) * B + rr] -= v;
      }
    }
    __syncthreads();  // the next batch overwrites this one
  }
}

// Step 2, one workgroup per pair row r (centre i, self pair first): the
// three rows dX/dr_x of the feature gradient matrix
// (calculator_spherical_invariants.hh:796-921, normalisation :588-652):
//   dp^{ab}[n1 n2 l] = f_l w_ab sum_m (dA[n1][lm] c_b[n2][lm] + c_a[n1][lm] dB[n2][lm])
// with dA, dB the derivatives of c^{i,a}, c^{i,b}: the centre's own blocks
// for the self row; for a neighbour row the pair derivative in the
// neighbour species' block and zero in the others.  Then, when normalising,
// dX = dp/|p| - p (p . dp)/|p|^3.  Output columns k: the full layout, or the
// selected coefficients col[k] (full-layout index, -1 for a species pair
// outside the batch); pv_rows[i][k] holds the un-normalised p of column k.
__global__ __launch_bounds__(kThreads) void feature_gradient_rows_kernel(
    int N, int L1, int S, PairView pv, const double* __restrict__ c,
    const double* __restrict__ pv_rows, const double* __restrict__ norm2, int normalize,
    const double* __restrict__ dcp, const double* __restrict__ dcs, int64_t Dout,
    const int64_t* __restrict__ col, double* __restrict__ out) {
  const int64_t r = blockIdx.x;
  int64_t lo = 0, hi = pv.n_centres;  // the centre i with off[i] <= r < off[i + 1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (pv.centre_off[mid] <= r)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t i = lo;
  const bool self = r == pv.centre_off[i];
  const int LL = L1 * L1, B = N * LL, KB = N * N * L1;
  const int s = self ? -1 : pv.species_of[pv.pair_j[r]];
  const double* ci = c + i * int64_t(S) * B;
  const double* pi = pv_rows + i * Dout;
  double dot[3] = {0, 0, 0};
  for (int64_t k = threadIdx.x; k < Dout; k += kThreads) {
    const int64_t o = col ? col[k] : k;
    double g[3] = {0, 0, 0};
    if (o >= 0) {
      const int key = int(o / KB), rem = int(o - int64_t(key) * KB);
      const int n1 = rem / (N * L1), rem2 = rem - n1 * N * L1, n2 = rem2 / L1, l = rem2 - n2 * L1;
      int a, b;
      key_pair(key, S, &a, &b);
      const double f = (a < b ? 1.41421356237309504880 : 1.0) / sqrt(double(2 * l + 1));
      const double* ca = ci + (a * N + n1) * LL;
      const double* cb = ci + (b * N + n2) * LL;
      for (int x = 0; x < 3; ++x) {
        const double* dA = self ? dcs + ((i * 3 + x) * S + a) * B + n1 * LL
                                : (a == s ? dcp + (r * 3 + x) * B + n1 * LL : nullptr);
        const double* dB = self ? dcs + ((i * 3 + x) * S + b) * B + n2 * LL
                                : (b == s ? dcp + (r * 3 + x) * B + n2 * LL : nullptr);
        double acc = 0;
        for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm) {
          if (dA) acc += dA[lm] * cb[lm];
          if (dB) acc += ca[lm] * dB[lm];
        }
        g[x] = f * acc;
        dot[x] += pi[k] * g[x];
      }
    }
    for (int x = 0; x < 3; ++x) out[(r * 3 + x) * Dout + k] = g[x];
  }
  if (!normalize) return;
  __shared__ double red[kWaves][3];
  for (int x = 0; x < 3; ++x) {
    double v = dot[x];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][x] = v;
  }
  __syncthreads();
  double tot[3];
  for (int x = 0; x < 3; ++x) {
    tot[x] = 0;
    for (int w = 0; w < kWaves; ++w) tot[x] += red[w][x];
  }
  const double n2 = norm2[i];
  const double inv = n2 > 0 ? 1.0 / sqrt(n2) : 0.0;
  for (int64_t k = threadIdx.x; k < Dout; k += kThreads)
    for (int x = 0; x < 3; ++x) {
      double& v = out[(r * 3 + x) * Dout + k];
      v = v * inv - pi[k] * tot[x] * inv * inv * inv;
    }
}

// Sparsified invariants (coefficient_subselection): xs[i][k] = p[i][col[k]]
// (0 for col[k] < 0) and norm2[i] = |xs_i|^2; then row i of p becomes S^T xs_i
// for the selection matrix S (the selected coefficients, summed over
// repeated selections, zero elsewhere): with it and |xs_i| the adjoint's
// normalisation term is the one of the selected features.
__global__ __launch_bounds__(kThreads) void sparsify_kernel(int64_t D, int64_t Dsel,
                                                            const int64_t* __restrict__ col,
                                                            const int64_t* __restrict__ inv_off,
                                                            const int64_t* __restrict__ inv_k,
                                                            double* __restrict__ p,
                                                            double* __restrict__ xs,
                                                            double* __restrict__ norm2) {
  const int64_t i = blockIdx.x;
  double* pi = p + i * D;
  double* xi = xs + i * Dsel;
  double ss = 0;
  for (int64_t k = threadIdx.x; k < Dsel; k += kThreads) {
    const double v = col[k] >= 0 ? pi[col[k]] : 0.0;
    xi[k] = v;
    ss += v * v;
  }
  __shared__ double red[kWaves];
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();  // also: every selected value has been read and xi written
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kWaves; ++w) t += red[w];
    norm2[i] = t;
  }
  // column o takes the selections naming it, in ascending selection order
  for (int64_t o = threadIdx.x; o < D; o += kThreads) {
    double t = 0.0;
    for (int64_t q = inv_off[o]; q < inv_off[o + 1]; ++q) t += xi[inv_k[q]];
    pi[o] = t;
  }
}

// dst[i] = S^T src[i] (rows of Dsel selected columns into rows of D
// columns, repeated selections summed in ascending selection order through
// the inverse map inv_off/inv_k), one workgroup per row
__global__ __launch_bounds__(kThreads) void scatter_columns_kernel(int64_t Dsel,
                                                                   const int64_t* __restrict__ inv_off,
                                                                   const int64_t* __restrict__ inv_k,
                                                                   const double* __restrict__ src,
                                                                   int64_t D, double* __restrict__ dst) {
  const int64_t i = blockIdx.x;
  double* di = dst + i * D;
  const double* si = src + i * Dsel;
  for (int64_t o = threadIdx.x; o < D; o += kThreads) {
    double t = 0.0;
    for (int64_t q = inv_off[o]; q < inv_off[o + 1]; ++q) t += si[inv_k[q]];
    di[o] = t;
  }
}

// LDS per workgroup: the forward and the power spectrum stay small enough
// for several workgroups per CU; the adjoint may use more.
constexpr size_t kFwdBudget = 32 * 1024;
constexpr size_t kPsBudget = 48 * 1024;
constexpr size_t kAdjBudget = 64 * 1024;
constexpr size_t kResidentBudget = 52 * 1024;  // three workgroups per CU

// LDS limit per workgroup of the current device (queried once per device;
// launches may be captured into graphs, where device queries are avoided)
size_t device_lds_limit() {
  static int cached[64] = {0};
  int dev = 0;
  RGPU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw RuntimeError("device index out of range");
  if (!cached[dev]) {
    int v = 0;
    RGPU_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached[dev] = v;
  }
  return size_t(cached[dev]);
}

template <class K>
void allow_lds(K kernel, size_t smem) {
  if (smem > device_lds_limit())
    throw InvalidArgument("hypers need " + std::to_string(smem) +
                          " bytes of LDS per workgroup, more than the device provides");
  if (smem > 64 * 1024)
    RGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(smem)));
}

template <class T, int KPT, bool DIRECT, int NC = -1, int LC = -1>
void launch_expansion_t(const RadialDev& rad, int S, const PairView& pv, T* c,
                        unsigned long long* keymask, hipStream_t stream, const SoapTiles& tiles) {
  const int F = rad.N * rad.L1;
  const size_t yt = ylm_table_bytes<T>(rad.L1);
  int PB = 64;
  if (tiles.pb_fwd > 0)
    PB = std::min(tiles.pb_fwd, 64);
  else
    while (PB > 1 && yt + pair_batch_bytes(PB, rad.LL, F, sizeof(T), DIRECT, false) > kFwdBudget)
      PB /= 2;
  const size_t smem = yt + pair_batch_bytes(PB, rad.LL, F, sizeof(T), DIRECT, false);
  auto kern = expansion_kernel<T, KPT, DIRECT, NC, LC>;
  allow_lds(kern, smem);
  hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB, pv,
                     c, keymask);
  RGPU_LAUNCH_CHECK();
}

template <class T, int KPT>
void launch_expansion_k(const RadialDev& rad, int S, const PairView& pv, T* c,
                        unsigned long long* keymask, hipStream_t stream, const SoapTiles& tiles) {
  if (rad.use_spline)
    launch_expansion_t<T, KPT, false>(rad, S, pv, c, keymask, stream, tiles);
  else
    launch_expansion_t<T, KPT, true>(rad, S, pv, c, keymask, stream, tiles);
}

// dynamic LDS of the register-tile adjoint: two workgroups per CU with the
// pair phase in the same workgroup, three without it
constexpr size_t kTileAdjBudget = 77 * 1024;
constexpr size_t kTileSplitBudget = 52 * 1024;

template <int N, int L, int NR, int ROLES, int PRE, bool DIRECT, bool PAIRS>
bool go_adjoint_tile(const RadialDev& rad, int S, const PairView& pv, const float* c, const double* norm2,
                     bool normalize, const double* xg, double xg_scale, const float* G,
                     double* centre_virial, float* work, hipStream_t stream, const SoapTiles& tiles) {
  constexpr int L1 = L + 1, LL = L1 * L1, F = N * L1;
  constexpr int KB = N * N * L1, KD = N * (N + 1) / 2 * L1;
  constexpr int MT = 4 * adj_quads_before(L1), GRAM = adj_gram_before(L1);
  const AdjRounds R(S);
  // one round must fit the prefetch registers
  if (S * (KD / 4) > PRE * kThreads || R.per * (KB / 4) > PRE * kThreads) return false;
  const size_t budget = PAIRS ? kTileAdjBudget : kTileSplitBudget;
  const size_t fixed = (PAIRS ? ylm_table_bytes<float>(L1) : 0) + align_up(sizeof(float) * S * N * MT, 64) +
                       align_up(sizeof(float) * GRAM, 64);
  int maxb = tiles.keys_adj;
  if (maxb <= 0) maxb = fixed < budget ? int((budget - fixed) / (sizeof(float) * KB)) : 0;
  maxb = std::max(maxb, std::max(S, R.per));
  int big = 0;  // blocks of the largest chunk
  for (int r0 = 0; r0 < R.rounds();) {
    const int r1 = R.chunk_end(r0, maxb, KD / 4, KB / 4, PRE);
    int nb = 0;
    for (int r = r0; r < r1; ++r) nb += R.blocks(r);
    big = std::max(big, nb);
    r0 = r1;
  }
  const size_t region = align_up(sizeof(float) * size_t(big) * KB, 64);
  int PB = 0;
  size_t smem = fixed + region;
  if (PAIRS) {
    PB = tiles.pb_adj > 0 ? std::min(tiles.pb_adj, 64) : 64;
    if (tiles.pb_adj <= 0)
      while (PB > 1 && pair_batch_bytes(PB, LL, F, sizeof(float), DIRECT, true) > region) --PB;
    smem = fixed + std::max(region, pair_batch_bytes(PB, LL, F, sizeof(float), DIRECT, true));
  }
  auto kern = adjoint_tile_kernel<N, L, NR, ROLES, PRE, DIRECT, PAIRS>;
  allow_lds(kern, smem);
  hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB, maxb, pv,
                     c, norm2, int(normalize), xg, xg_scale, G, centre_virial, work);
  RGPU_LAUNCH_CHECK();
  if (!PAIRS) {
    const size_t ps = pair_kernel_lds(S, N, L1);
    auto pk = adjoint_pair_kernel<N, L>;
    allow_lds(pk, ps);
    hipLaunchKernelGGL(pk, dim3(unsigned(pv.n_centres)), dim3(kThreads), ps, stream, rad, S, pv,
                       static_cast<const float*>(work), centre_virial);
    RGPU_LAUNCH_CHECK();
  }
  return true;
}

// The register-tile adjoint for the (max_radial, max_angular) it is built
// for (up to 8 species); false where it does not apply.  With a workspace
// of n_centres S (L+1)^2 N floats and spline radial functions, the pair
// phase runs as adjoint_pair_kernel (split); otherwise in the same
// workgroup.
template <int N, int L, int NR, bool DIRECT>
bool launch_adjoint_tile_nl(const RadialDev& rad, int S, const PairView& pv, const float* c, const double* norm2,
                            bool normalize, const double* xg, double xg_scale, const float* G,
                            double* centre_virial, float* work, hipStream_t stream, const SoapTiles& tiles) {
  if (work && !DIRECT) {
    if (S <= kWaves)
      return go_adjoint_tile<N, L, NR, 1, 6, false, false>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                           centre_virial, work, stream, tiles);
    return go_adjoint_tile<N, L, NR, 2, 8, false, false>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                         centre_virial, work, stream, tiles);
  }
  if (S <= kWaves)
    return go_adjoint_tile<N, L, NR, 1, 10, DIRECT, true>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                          centre_virial, nullptr, stream, tiles);
  return go_adjoint_tile<N, L, NR, 2, 10, DIRECT, true>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                        centre_virial, nullptr, stream, tiles);
}

template <bool DIRECT>
bool launch_adjoint_tile(const RadialDev& rad, int S, const PairView& pv, const float* c, const double* norm2,
                         bool normalize, const double* xg, double xg_scale, const float* G,
                         double* centre_virial, float* work, hipStream_t stream, const SoapTiles& tiles) {
  if (S > 2 * kWaves) return false;
  if (rad.N == 12 && rad.L == 9)
    return launch_adjoint_tile_nl<12, 9, 6, DIRECT>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                    centre_virial, work, stream, tiles);
  if (rad.N == 8 && rad.L == 6)
    return launch_adjoint_tile_nl<8, 6, 2, DIRECT>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G,
                                                   centre_virial, work, stream, tiles);
  return false;
}

template <class T, bool DIRECT>
void launch_adjoint_t(const RadialDev& rad, int S, const PairView& pv, const T* c, const T* p,
                      const unsigned short* p_hi, const unsigned short* p_lo,
                      const double* norm2, bool normalize, const double* xg, double xg_scale,
                      const T* G, double* centre_virial, hipStream_t stream,
                      const SoapTiles& tiles, const T* Grow = nullptr,
                      const double* Gscale = nullptr, float* work = nullptr) {
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, F = N * L1, E = S * N * LL;
  const size_t yt = ylm_table_bytes<T>(L1);
  const size_t blk = align_up(sizeof(T) * size_t(E), 64);
  if constexpr (sizeof(T) == 4) {
    const int K = S * (S + 1) / 2;
    if ((tiles.adj_kind == 2 || tiles.adj_kind == 3) && Grow == nullptr &&
        launch_adjoint_tile<DIRECT>(rad, S, pv, c, norm2, normalize, xg, xg_scale, G, centre_virial,
                                    tiles.adj_kind == 3 ? work : nullptr, stream, tiles))
      return;
    // matrix-core adjoint: packed dL/dX staged KCH key blocks at a time,
    // the invariants not read
    int T_col = 0;
    for (int l = 0; l <= rad.L; ++l) T_col += (2 * l + 16) >> 4;
    const int maxt = (S * T_col + kWaves - 1) / kWaves;
    if (tiles.adj_kind == 1 && N <= 16 && maxt <= 16 && Grow == nullptr) {
      const size_t cbytes = align_up(sizeof(float) * size_t(S) * N * ((LL + 15) / 32 * 32 + 16), 64);
      size_t mbytes = 0;
      for (int l = 0; l <= rad.L; ++l) {
        const size_t w = 16 * ((2 * l + 16) >> 4);
        mbytes += sizeof(float) * w * w;
      }
      // packed G of a chunk of keys (the largest chunk: diagonal keys are
      // the smaller blocks)
      const size_t kbytes = sizeof(float) * size_t(L1) * N * (4 * ((N + 1) / 4) + 2);
      int KCH = tiles.keys_adj > 0 ? std::min(tiles.keys_adj, K) : K;
      // (A fragments of rows N..15 read past the last staged block)
      const size_t slack = sizeof(float) * 16 * (4 * ((N + 1) / 4) + 2);
      auto region_of = [&](int kch, int pb) {
        return std::max({align_up(kch * kbytes + slack, 64), align_up(mbytes, 64),
                         pair_batch_bytes(pb, LL, F, sizeof(T), DIRECT, true)});
      };
      int PB = tiles.pb_adj > 0 ? std::min(tiles.pb_adj, 64) : 16;
      if (tiles.keys_adj <= 0)
        while (KCH > 1 && yt + cbytes + region_of(KCH, 1) > kResidentBudget) --KCH;
      if (tiles.pb_adj <= 0)
        while (PB > 1 && pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, true) >
                             std::max(align_up(KCH * kbytes, 64), align_up(mbytes, 64)))
          --PB;
      const size_t smem = yt + cbytes + region_of(KCH, PB);
      auto go = [&](auto kern) {
        allow_lds(kern, smem);
        hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB,
                           KCH, pv, c, norm2, int(normalize), xg, xg_scale, G, centre_virial);
      };
      if (maxt <= 4)
        go(adjoint_mfma_kernel<DIRECT, 4>);
      else if (maxt <= 8)
        go(adjoint_mfma_kernel<DIRECT, 8>);
      else if (maxt <= 12)
        go(adjoint_mfma_kernel<DIRECT, 12>);
      else
        go(adjoint_mfma_kernel<DIRECT, 16>);
      RGPU_LAUNCH_CHECK();
      return;
    }
    // mixed path: packed dL/dX and BF16 planes of p, staged in LDS KCH key
    // blocks at a time beside c
    if (!p_hi || !p_lo) throw InvalidArgument("the FP32 adjoint takes the split invariants");
    const size_t kbytes = sizeof(T) * size_t(N) * N * L1;
    const int upt = (S * ((N + 3) / 4) * LL + kThreads - 1) / kThreads;
    if (upt > 8) throw InvalidArgument("max_radial x (max_angular+1)^2 too large for the FP32 adjoint");
    int KCH = tiles.keys_adj > 0 ? std::min(tiles.keys_adj, K) : K;
    if (tiles.keys_adj <= 0)
      while (KCH > 1 && yt + blk + align_up(KCH * kbytes, 64) > kResidentBudget) --KCH;
    const size_t dbytes = align_up(KCH * kbytes, 64);
    int PB = tiles.pb_adj > 0 ? std::min(tiles.pb_adj, 64) : 16;
    if (tiles.pb_adj <= 0)
      while (PB > 1 && pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, true) > dbytes) --PB;
    const size_t smem = yt + blk + std::max(dbytes, pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, true));
    auto go = [&](auto kern) {
      allow_lds(kern, smem);
      hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB, KCH,
                         pv, c, p_hi, p_lo, norm2, int(normalize), xg, xg_scale, G, centre_virial);
    };
    if (upt <= 2)
      go(adjoint_staged_kernel<DIRECT, 2>);
    else if (upt <= 3)
      go(adjoint_staged_kernel<DIRECT, 3>);
    else if (upt <= 4)
      go(adjoint_staged_kernel<DIRECT, 4>);
    else if (upt <= 5)
      go(adjoint_staged_kernel<DIRECT, 5>);
    else
      go(adjoint_staged_kernel<DIRECT, 8>);
    RGPU_LAUNCH_CHECK();
  } else {
    (void)p_hi;
    (void)p_lo;
    const int N4 = (N + 3) & ~3;
    const size_t gbytes = sizeof(T) * 2 * size_t(L1) * (N * N4 + 2);  // two l-plane sets, LS apart
    const int c_in_lds = yt + blk + blk + gbytes <= kAdjBudget ? 1 : 0;
    const size_t region = (c_in_lds ? blk : 0) + gbytes;
    const size_t fixed = yt + blk;
    const size_t batch_room = std::max(region, kAdjBudget > fixed ? kAdjBudget - fixed : size_t(0));
    int PB = 8;
    if (tiles.pb_adj > 0)
      PB = std::min(tiles.pb_adj, 64);
    else
      while (PB > 1 && pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, true) > batch_room) --PB;
    const size_t smem = fixed + std::max(region, pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, true));
    auto kern = adjoint_kernel<T, DIRECT>;
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB,
                       c_in_lds, pv, c, p, norm2, int(normalize), xg, xg_scale, G, Grow, Gscale,
                       centre_virial);
    RGPU_LAUNCH_CHECK();
  }
}

}  // namespace

// The pair-group expansion for the max_angular it is built for (FP32 spline
// radial functions, max_radial <= 16, at least 3 species: with fewer, most
// of its wavefronts would idle); false where it does not apply.
template <class T>
bool launch_expansion_group(const RadialDev& rad, int S, const PairView& pv, T* c,
                            unsigned long long* keymask, hipStream_t stream, const SoapTiles& tiles) {
  if constexpr (sizeof(T) != 4) {
    return false;
  } else {
    if (!tiles.fwd_group || !rad.use_spline || rad.N > kPairGroup || S < 3) return false;
    auto go = [&](auto kern) {
      const size_t smem = expansion_group_lds(S, rad.N, rad.L1);
      allow_lds(kern, smem);
      hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, pv, c,
                         keymask);
      RGPU_LAUNCH_CHECK();
      return true;
    };
    if (rad.L == 9) return go(expansion_group_kernel<9>);
    if (rad.L == 6) return go(expansion_group_kernel<6>);
    return false;
  }
}

template <class T>
void launch_expansion(const RadialDev& rad, int S, const PairView& pv, T* c,
                      unsigned long long* keymask, hipStream_t stream, const SoapTiles& tiles) {
  if (pv.n_centres == 0) return;
  if (S > kMaxSpecies) throw InvalidArgument("at most 64 species");
  if (launch_expansion_group<T>(rad, S, pv, c, keymask, stream, tiles)) return;
  const int B = rad.N * rad.LL;
  const int kpt = (B + kThreads - 1) / kThreads;
  // the benchmark shape (n_max 12, l_max 9, spline radial functions) with
  // its sizes fixed at compile time
  constexpr bool kFixedShape = true;
  if (kFixedShape && rad.use_spline && rad.N == 12 && rad.L == 9) {
    launch_expansion_t<T, 5, false, 12, 9>(rad, S, pv, c, keymask, stream, tiles);
    return;
  }
  if (kpt <= 2)
    launch_expansion_k<T, 2>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 4)
    launch_expansion_k<T, 4>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 5)
    launch_expansion_k<T, 5>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 8)
    launch_expansion_k<T, 8>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 12)
    launch_expansion_k<T, 12>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 16)
    launch_expansion_k<T, 16>(rad, S, pv, c, keymask, stream, tiles);
  else if (kpt <= 28)
    launch_expansion_k<T, 28>(rad, S, pv, c, keymask, stream, tiles);
  else
    throw InvalidArgument("max_radial x (max_angular+1)^2 too large (> 7168)");
}

template <class T>
void launch_power_spectrum_impl(const RadialDev& rad, int S, int64_t n, const T* c, T* p,
                                unsigned short* p_hi, unsigned short* p_lo, double* norm2,
                                hipStream_t stream, const SoapTiles& tiles) {
  if (n == 0) return;
  const int K = S * (S + 1) / 2;
  if (S > kMaxSpecies) throw InvalidArgument("at most 64 species");
  const size_t cbytes = sizeof(T) * align_up((size_t(S) * rad.N + 3) * (rad.LL | 1), 4);
  const size_t kb = sizeof(T) * size_t(rad.N) * rad.N * rad.L1;
  int KC = cbytes + kb < kPsBudget ? int((kPsBudget - cbytes) / kb) : 1;
  if (tiles.kc > 0) KC = tiles.kc;
  KC = std::max(1, std::min(KC, K));
  const size_t smem = cbytes + size_t(KC) * kb + sizeof(int) * size_t(K + rad.N * (rad.N + 1) / 2);
  // the benchmark shape (n_max 12, l_max 9) with its sizes fixed at compile time
  constexpr bool kFixedShape = true;
  if (p_hi && kFixedShape && rad.N == 12 && rad.L == 9) {
    auto kern = power_spectrum_kernel<T, true, 12, 9>;
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad.N, rad.L, S, KC,
                       packed_row_stride(S, rad.N, rad.L1), c, p, p_hi, p_lo, norm2);
  } else if (!p_hi && kFixedShape && rad.N == 12 && rad.L == 9) {
    auto kern = power_spectrum_kernel<T, false, 12, 9>;
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad.N, rad.L, S, KC,
                       packed_row_stride(S, rad.N, rad.L1), c, p, p_hi, p_lo, norm2);
  } else if (p_hi) {
    auto kern = power_spectrum_kernel<T, true>;
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad.N, rad.L, S, KC,
                       packed_row_stride(S, rad.N, rad.L1), c, p, p_hi, p_lo, norm2);
  } else {
    auto kern = power_spectrum_kernel<T, false>;
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad.N, rad.L, S, KC,
                       packed_row_stride(S, rad.N, rad.L1), c, p, p_hi, p_lo, norm2);
  }
  RGPU_LAUNCH_CHECK();
}

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  RGPU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw RuntimeError("device index out of range");
  if (!cached[dev]) {
    int v = 0;
    RGPU_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
    cached[dev] = v;
  }
  return cached[dev];
}

template <class T, int KPT, bool DIRECT, bool SPLIT, bool PS>
void launch_fused_t(const RadialDev& rad, int S, const PairView& pv, T* c, unsigned long long* keymask,
                    T* p, unsigned short* p_hi, unsigned short* p_lo, double* norm2, hipStream_t stream,
                    const SoapTiles& tiles) {
  const int N = rad.N, L1 = rad.L1, LL = rad.LL, F = N * L1, K = S * (S + 1) / 2;
  const size_t yt = ylm_table_bytes<T>(L1);
  const size_t cbytes = sizeof(T) * (size_t(S) * N + 3) * (LL | 1);
  const size_t fixed = align_up(yt + cbytes + sizeof(int) * size_t(K + N * (N + 1) / 2), 64);
  const size_t kb = sizeof(T) * size_t(N) * N * L1;
  int PB = tiles.pb_fwd > 0 ? std::min(tiles.pb_fwd, 64) : 16;
  int KC = tiles.kc > 0 ? std::min(tiles.kc, K) : 2;
  const size_t region = std::max(pair_batch_bytes(PB, LL, F, sizeof(T), DIRECT, false), PS ? KC * kb : 0);
  const size_t smem = fixed + region;
  auto kern = expansion_ps_kernel<T, KPT, DIRECT, SPLIT, PS>;
  allow_lds(kern, smem);
  // persistent: as many workgroups as fit at once (cached per launch shape)
  static std::vector<std::pair<std::pair<const void*, size_t>, int>> occ_cache;
  int per_cu = 0;
  for (auto& e : occ_cache)
    if (e.first.first == reinterpret_cast<const void*>(kern) && e.first.second == smem) per_cu = e.second;
  if (!per_cu) {
    RGPU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, smem));
    per_cu = std::max(per_cu, 1);
    occ_cache.push_back({{reinterpret_cast<const void*>(kern), smem}, per_cu});
  }
  const int64_t grid = std::min<int64_t>(pv.n_centres, int64_t(per_cu) * device_cu_count());
  hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(kThreads), smem, stream, rad, S, PB, KC,
                     packed_row_stride(S, N, L1), pv, c, keymask, p, p_hi, p_lo, norm2);
  RGPU_LAUNCH_CHECK();
}

template <class T, int KPT>
void launch_fused_k(const RadialDev& rad, int S, const PairView& pv, T* c, unsigned long long* keymask,
                    T* p, unsigned short* p_hi, unsigned short* p_lo, double* norm2, hipStream_t stream,
                    const SoapTiles& tiles, bool ps) {
  constexpr bool split = std::is_same<T, float>::value;
  if (ps) {
    if (rad.use_spline)
      launch_fused_t<T, KPT, false, split, true>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles);
    else
      launch_fused_t<T, KPT, true, split, true>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles);
  } else {
    if (rad.use_spline)
      launch_fused_t<T, KPT, false, split, false>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles);
    else
      launch_fused_t<T, KPT, true, split, false>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles);
  }
}

template <class T>
void launch_expansion_ps(const RadialDev& rad, int S, const PairView& pv, T* c,
                         unsigned long long* keymask, T* p, unsigned short* p_hi, unsigned short* p_lo,
                         double* norm2, hipStream_t stream, const SoapTiles& tiles) {
  if (pv.n_centres == 0) return;
  if (S > kMaxSpecies) throw InvalidArgument("at most 64 species");
  const int kpt = (rad.N * rad.LL + kThreads - 1) / kThreads;
  const bool ps = p || p_hi;
  if (kpt <= 2)
    launch_fused_k<T, 2>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles, ps);
  else if (kpt <= 5)
    launch_fused_k<T, 5>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles, ps);
  else if (kpt <= 8)
    launch_fused_k<T, 8>(rad, S, pv, c, keymask, p, p_hi, p_lo, norm2, stream, tiles, ps);
  else
    throw InvalidArgument("max_radial x (max_angular+1)^2 too large for the fused kernel");
}

template void launch_expansion_ps<double>(const RadialDev&, int, const PairView&, double*,
                                          unsigned long long*, double*, unsigned short*,
                                          unsigned short*, double*, hipStream_t, const SoapTiles&);
template void launch_expansion_ps<float>(const RadialDev&, int, const PairView&, float*,
                                         unsigned long long*, float*, unsigned short*, unsigned short*,
                                         double*, hipStream_t, const SoapTiles&);

template <class T>
void launch_power_spectrum(const RadialDev& rad, int S, int64_t n, const T* c, T* p, double* norm2,
                           hipStream_t stream, const SoapTiles& tiles) {
  launch_power_spectrum_impl<T>(rad, S, n, c, p, nullptr, nullptr, norm2, stream, tiles);
}

void launch_power_spectrum_split(const RadialDev& rad, int S, int64_t n, const float* c,
                                 unsigned short* p_hi, unsigned short* p_lo, double* norm2,
                                 hipStream_t stream, const SoapTiles& tiles) {
  launch_power_spectrum_impl<float>(rad, S, n, c, nullptr, p_hi, p_lo, norm2, stream, tiles);
}

template <class T>
void launch_normalise(int64_t n, int64_t D, const T* p, const double* norm2, bool normalize, double* out,
                      hipStream_t stream, const int64_t* rows) {
  if (n == 0 || D == 0) return;
  int64_t total = n * D;
  unsigned blocks = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(normalise_kernel<T>, dim3(blocks), dim3(256), 0, stream, n, D, p, norm2,
                     int(normalize), rows, out);
  RGPU_LAUNCH_CHECK();
}

void launch_fill(double* p, int64_t n, double v, hipStream_t stream) {
  if (n <= 0) return;
  const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks), dim3(256), 0, stream, p, n, v);
  RGPU_LAUNCH_CHECK();
}

void reserve_half_list_scan(int64_t n, DeviceBuffer<int64_t>& scratch) {
  exclusive_scan_reserve(n, scratch);
}

template <class T>
void launch_expansion_half(const RadialDev& rad, int S, const PairView& pv, const double* pos,
                           const HalfList& hl, T* c, unsigned long long* keymask, hipStream_t stream,
                           const SoapTiles& tiles) {
  const int64_t n = pv.n_centres;
  if (n == 0) return;
  const unsigned g1 = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(half_count_kernel, dim3(g1), dim3(256), 0, stream, pv, pos, hl.count);
  RGPU_LAUNCH_CHECK();
  exclusive_scan_device(hl.count, hl.off, n, stream, *hl.scan_scratch);
  hipLaunchKernelGGL(half_fill_kernel, dim3(g1), dim3(256), 0, stream, pv, pos, hl.off, hl.src, hl.j,
                     hl.vec, hl.dist, hl.fc, hl.dfc);
  RGPU_LAUNCH_CHECK();
  hipLaunchKernelGGL(half_mirror_kernel, dim3(g1), dim3(256), 0, stream, pv, pos, hl.src);
  RGPU_LAUNCH_CHECK();
  PairView hv = pv;
  hv.centre_off = hl.off;
  hv.pair_j = hl.j;
  hv.pair_vec = hl.vec;
  hv.pair_dist = hl.dist;
  hv.pair_fc = hl.fc;
  hv.pair_dfc = hl.dfc;
  const int F = rad.N * rad.L1;
  const bool direct = !rad.use_spline;
  const size_t yt = ylm_table_bytes<T>(rad.L1);
  int PB = tiles.pb_fwd > 0 ? std::min(tiles.pb_fwd, 64) : 64;
  if (tiles.pb_fwd <= 0)
    while (PB > 1 && yt + pair_batch_bytes(PB, rad.LL, F, sizeof(T), direct, false) > kFwdBudget) PB /= 2;
  const size_t smem = yt + pair_batch_bytes(PB, rad.LL, F, sizeof(T), direct, false);
  T* rows = static_cast<T*>(hl.rows);
  auto go = [&](auto kern) {
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(n)), dim3(kThreads), smem, stream, rad, S, PB, hv, rows);
  };
  if (direct)
    go(half_rows_kernel<T, true>);
  else
    go(half_rows_kernel<T, false>);
  RGPU_LAUNCH_CHECK();
  hipLaunchKernelGGL(half_gather_kernel<T>, dim3(unsigned(n)), dim3(kThreads), 0, stream, rad, S, pv,
                     static_cast<const int64_t*>(hl.src), static_cast<const T*>(rows), c, keymask);
  RGPU_LAUNCH_CHECK();
}

template void launch_expansion_half<double>(const RadialDev&, int, const PairView&, const double*,
                                            const HalfList&, double*, unsigned long long*, hipStream_t,
                                            const SoapTiles&);
template void launch_expansion_half<float>(const RadialDev&, int, const PairView&, const double*,
                                           const HalfList&, float*, unsigned long long*, hipStream_t,
                                           const SoapTiles&);

void launch_pair_gradients(const RadialDev& rad, int S, const PairView& pv, double* dcp, double* dcs,
                           hipStream_t stream) {
  if (pv.n_centres == 0) return;
  const int F = rad.N * rad.L1;
  const size_t yt = ylm_table_bytes<double>(rad.L1);
  const bool direct = !rad.use_spline;
  int PB = 16;
  while (PB > 1 && yt + pair_batch_bytes(PB, rad.LL, F, sizeof(double), direct, true) > kAdjBudget)
    PB /= 2;
  const size_t smem = yt + pair_batch_bytes(PB, rad.LL, F, sizeof(double), direct, true);
  auto go = [&](auto kern) {
    allow_lds(kern, smem);
    hipLaunchKernelGGL(kern, dim3(unsigned(pv.n_centres)), dim3(kThreads), smem, stream, rad, S, PB,
                       pv, dcp, dcs);
  };
  if (direct)
    go(pair_dc_kernel<true>);
  else
    go(pair_dc_kernel<false>);
  RGPU_LAUNCH_CHECK();
}

void launch_feature_gradient_rows(int N, int L1, int S, const PairView& pv, int64_t n_pairs,
                                  const double* c, const double* p_rows, const double* norm2,
                                  bool normalize, const double* dcp, const double* dcs, int64_t Dout,
                                  const int64_t* col, double* out, hipStream_t stream) {
  if (n_pairs == 0 || Dout == 0) return;
  hipLaunchKernelGGL(feature_gradient_rows_kernel, dim3(unsigned(n_pairs)), dim3(kThreads), 0, stream,
                     N, L1, S, pv, c, p_rows, norm2, int(normalize), dcp, dcs, Dout, col, out);
  RGPU_LAUNCH_CHECK();
}

void launch_sparsify(int64_t n, int64_t D, int64_t Dsel, const SelectionMap& sel, double* p, double* xs,
                     double* norm2, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(sparsify_kernel, dim3(unsigned(n)), dim3(kThreads), 0, stream, D, Dsel, sel.col,
                     sel.inv_off, sel.inv_k, p, xs, norm2);
  RGPU_LAUNCH_CHECK();
}

void launch_scatter_columns(int64_t n, int64_t Dsel, const SelectionMap& sel, const double* src, int64_t D,
                            double* dst, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(scatter_columns_kernel, dim3(unsigned(n)), dim3(kThreads), 0, stream, Dsel, sel.inv_off,
                     sel.inv_k, src, D, dst);
  RGPU_LAUNCH_CHECK();
}

template <class T>
void launch_backward(const RadialDev& rad, int S, const PairView& pv, const T* c, const T* p,
                     const double* norm2, bool normalize, const double* xg, double xg_scale,
                     const T* G, double* centre_virial, hipStream_t stream,
                     const SoapTiles& tiles) {
  if (pv.n_centres == 0) return;
  if (rad.use_spline)
    launch_adjoint_t<T, false>(rad, S, pv, c, p, nullptr, nullptr, norm2, normalize, xg, xg_scale, G,
                               centre_virial, stream, tiles);
  else
    launch_adjoint_t<T, true>(rad, S, pv, c, p, nullptr, nullptr, norm2, normalize, xg, xg_scale, G,
                              centre_virial, stream, tiles);
}

void launch_backward_rank1(const RadialDev& rad, int S, const PairView& pv, const double* c,
                           const double* p, const double* norm2, bool normalize, const double* Grow,
                           const double* Gscale, double* centre_virial,
                           hipStream_t stream, const SoapTiles& tiles) {
  if (pv.n_centres == 0) return;
  if (rad.use_spline)
    launch_adjoint_t<double, false>(rad, S, pv, c, p, nullptr, nullptr, norm2, normalize, nullptr, 1.0,
                                    nullptr, centre_virial, stream, tiles, Grow, Gscale);
  else
    launch_adjoint_t<double, true>(rad, S, pv, c, p, nullptr, nullptr, norm2, normalize, nullptr, 1.0,
                                   nullptr, centre_virial, stream, tiles, Grow, Gscale);
}

void launch_backward_split(const RadialDev& rad, int S, const PairView& pv, const float* c,
                           const unsigned short* p_hi, const unsigned short* p_lo,
                           const double* norm2, bool normalize, const double* xg, double xg_scale,
                           const float* G, double* centre_virial, hipStream_t stream,
                           const SoapTiles& tiles, float* work) {
  if (pv.n_centres == 0) return;
  if (rad.use_spline)
    launch_adjoint_t<float, false>(rad, S, pv, c, nullptr, p_hi, p_lo, norm2, normalize, xg, xg_scale,
                                   G, centre_virial, stream, tiles, nullptr, nullptr, work);
  else
    launch_adjoint_t<float, true>(rad, S, pv, c, nullptr, p_hi, p_lo, norm2, normalize, xg, xg_scale,
                                  G, centre_virial, stream, tiles, nullptr, nullptr, work);
}

void launch_pack_gradient(int S, int N, int L1, int64_t n, const double* G, float* Gp,
                          hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(pack_gradient_kernel, dim3(unsigned(n)), dim3(kThreads), 0, stream, S, N, L1, G, Gp);
  RGPU_LAUNCH_CHECK();
}

__global__ void knm_gradient_scale_kernel(int64_t n, const double* __restrict__ d, int64_t ld,
                                          const int* __restrict__ species_of, int s, int zeta,
                                          double* __restrict__ scale) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  if (species_of[i] == s) {
    const double x = d[i * ld];
    v = double(zeta);
    for (int k = 1; k < zeta; ++k) v *= x;
  }
  scale[i] = v;
}

__global__ void strided_copy_kernel(const double* __restrict__ src, int64_t n, double* __restrict__ dst,
                                    int64_t ld) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n) dst[r * ld] = src[r];
}

void launch_knm_gradient_scale(int64_t n, const double* d, int64_t ld, const int* species_of, int s,
                               int zeta, double* scale, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(knm_gradient_scale_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, n,
                     d, ld, species_of, s, zeta, scale);
  RGPU_LAUNCH_CHECK();
}

void launch_strided_copy(const double* src, int64_t n, double* dst, int64_t ld, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(strided_copy_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, src, n,
                     dst, ld);
  RGPU_LAUNCH_CHECK();
}

void launch_pair_cutoff(const RadialDev& rad, const int64_t* n_pairs, int64_t cap, const double* dist,
                        double* fc, double* dfc, hipStream_t stream) {
  if (cap == 0) return;
  hipLaunchKernelGGL(pair_cutoff_kernel, dim3(unsigned((cap + 255) / 256)), dim3(256), 0, stream, rad,
                     n_pairs, cap, dist, fc, dfc);
  RGPU_LAUNCH_CHECK();
}

void launch_segment_sum(const double* in, int w, const int64_t* seg_off, int n_seg, double* out,
                        hipStream_t stream, const double* scale) {
  if (n_seg == 0) return;
  if (w < 1 || w > 8) throw InvalidArgument("segment sum: row width must be 1 to 8");
  const SegJob j{in, out, scale, w};
  hipLaunchKernelGGL(segment_sum_kernel, dim3(unsigned(n_seg)), dim3(kSegThreads), 0, stream, j, j, seg_off,
                     n_seg);
  RGPU_LAUNCH_CHECK();
}

void launch_segment_sum2(const double* in0, int w0, double* out0, const double* scale0, const double* in1, int w1,
                         double* out1, const double* scale1, const int64_t* seg_off, int n_seg,
                         hipStream_t stream) {
  if (n_seg == 0) return;
  if (w0 < 1 || w0 > 8 || w1 < 1 || w1 > 8) throw InvalidArgument("segment sum: row width must be 1 to 8");
  hipLaunchKernelGGL(segment_sum_kernel, dim3(unsigned(2 * n_seg)), dim3(kSegThreads), 0, stream,
                     SegJob{in0, out0, scale0, w0}, SegJob{in1, out1, scale1, w1}, seg_off, n_seg);
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
void launch_ylm_eval(const RadialDev& rad, const double* dirs, int64_t n, double* Y, double* dY,
                     hipStream_t stream) {
  if (n == 0) return;
  const int L1 = rad.L + 1, LL = L1 * L1, per_block = kWaves * (64 / L1);
  size_t smem = ylm_table_bytes<T>(L1) + sizeof(T) * (3 * per_block + per_block * LL + per_block * 3 * LL);
  hipLaunchKernelGGL(ylm_eval_kernel<T>, dim3(unsigned((n + per_block - 1) / per_block)), dim3(kThreads),
                     smem, stream, rad, dirs, n, Y, dY);
  RGPU_LAUNCH_CHECK();
}

template void launch_radial_spline_eval<double>(const RadialDev&, const double*, int64_t, double*, double*,
                                                hipStream_t);
template void launch_radial_spline_eval<float>(const RadialDev&, const double*, int64_t, double*, double*,
                                               hipStream_t);
template void launch_ylm_eval<double>(const RadialDev&, const double*, int64_t, double*, double*,
                                      hipStream_t);
template void launch_ylm_eval<float>(const RadialDev&, const double*, int64_t, double*, double*,
                                     hipStream_t);

template void launch_expansion<double>(const RadialDev&, int, const PairView&, double*,
                                       unsigned long long*, hipStream_t, const SoapTiles&);
template void launch_expansion<float>(const RadialDev&, int, const PairView&, float*,
                                      unsigned long long*, hipStream_t, const SoapTiles&);
template void launch_power_spectrum<double>(const RadialDev&, int, int64_t, const double*, double*,
                                            double*, hipStream_t, const SoapTiles&);
template void launch_power_spectrum<float>(const RadialDev&, int, int64_t, const float*, float*,
                                           double*, hipStream_t, const SoapTiles&);
template void launch_normalise<double>(int64_t, int64_t, const double*, const double*, bool, double*,
                                       hipStream_t, const int64_t*);
template void launch_normalise<float>(int64_t, int64_t, const float*, const double*, bool, double*,
                                      hipStream_t, const int64_t*);
namespace {

// ---------------------------------------------------------------------------
// Deterministic force sum.  The adjoint writes each pair's force and each
// centre's own share; the force on atom j is its centre share plus the
// forces of the pairs whose neighbour is j, added in ascending pair index
// over reverse lists (counted, scanned, filled and sorted per atom), so the
// result does not depend on the order in which workgroups ran.
// one wavefront per centre, its lanes over the centre's pairs (the self
// pair skipped)
__global__ __launch_bounds__(256) void rev_count_kernel(int64_t n, const int64_t* __restrict__ centre_off,
                                                        const int* __restrict__ pair_j, int* __restrict__ count) {
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  for (int64_t p = centre_off[i] + 1 + (threadIdx.x & 63); p < centre_off[i + 1]; p += 64)
    atomicAdd(count + pair_j[p], 1);
}

// slots of atom j's list taken from the top down (count[j] returns to 0);
// the order inside a list is arbitrary: force_gather_kernel ranks it
__global__ __launch_bounds__(256) void rev_fill_kernel(int64_t n, const int64_t* __restrict__ centre_off,
                                                       const int* __restrict__ pair_j,
                                                       const int64_t* __restrict__ roff, int* __restrict__ count,
                                                       int64_t* __restrict__ rlist) {
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  for (int64_t p = centre_off[i] + 1 + (threadIdx.x & 63); p < centre_off[i + 1]; p += 64) {
    const int j = pair_j[p];
    rlist[roff[j] + atomicSub(count + j, 1) - 1] = p;
  }
}

// One wavefront per atom j: grad[j] = centre_force[j] + the pair forces of
// the pairs naming j, lane l summing the pairs of rank l, l + 64, ... in
// ascending pair index, then a fixed butterfly over the lanes.  Lists of at
// most kGatherCap pairs are ranked in LDS; longer ones take their pairs in
// ascending order one per step (the same sums, more slowly).
constexpr int kGatherCap = 256;

__global__ __launch_bounds__(256) void force_gather_kernel(int64_t n, const int64_t* __restrict__ roff,
                                                           const int64_t* __restrict__ rlist,
                                                           const double* __restrict__ pf,
                                                           const double* __restrict__ cf, double* __restrict__ grad) {
  __shared__ int64_t lst[4][kGatherCap], ord[4][kGatherCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = int64_t(blockIdx.x) * 4 + w;
  if (j >= n) return;  // whole wavefronts; no workgroup barrier follows
  const int64_t b = roff[j];
  const int m = int(roff[j + 1] - b);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (m <= kGatherCap) {
    for (int k = lane; k < m; k += 64) lst[w][k] = rlist[b + k];
    wave_lds_sync();
    for (int k = lane; k < m; k += 64) {
      const int64_t v = lst[w][k];
      int r = 0;
      for (int q = 0; q < m; ++q) r += lst[w][q] < v ? 1 : 0;
      ord[w][r] = v;
    }
    wave_lds_sync();
    for (int k = lane; k < m; k += 64) {
      const double* f = pf + 3 * ord[w][k];
      s0 += f[0];
      s1 += f[1];
      s2 += f[2];
    }
  } else {
    int64_t prev = -1;
    for (int k = 0; k < m; ++k) {
      long long best = LLONG_MAX;
      for (int q = lane; q < m; q += 64) {
        const long long v = rlist[b + q];
        if (v > prev && v < best) best = v;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const long long o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
      }
      prev = best;
      if (lane == (k & 63)) {
        const double* f = pf + 3 * best;
        s0 += f[0];
        s1 += f[1];
        s2 += f[2];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (lane == 0) {
    grad[3 * j] = cf[3 * j] + s0;
    grad[3 * j + 1] = cf[3 * j + 1] + s1;
    grad[3 * j + 2] = cf[3 * j + 2] + s2;
  }
}

}  // namespace

void launch_reverse_pairs(const PairView& pv, const ReverseLists& rl, DeviceBuffer<int64_t>& scan_scratch,
                          hipStream_t stream, bool clear_count) {
  const int64_t n = pv.n_centres;
  if (n == 0) return;
  const unsigned grid = unsigned((n + 3) / 4);  // a wavefront per centre
  if (clear_count) RGPU_HIP(hipMemsetAsync(rl.count, 0, sizeof(int) * n, stream));
  hipLaunchKernelGGL(rev_count_kernel, dim3(grid), dim3(256), 0, stream, n, pv.centre_off, pv.pair_j, rl.count);
  RGPU_LAUNCH_CHECK();
  exclusive_scan_reserve(n, scan_scratch);
  exclusive_scan_device(rl.count, rl.roff, n, stream, scan_scratch);
  hipLaunchKernelGGL(rev_fill_kernel, dim3(grid), dim3(256), 0, stream, n, pv.centre_off, pv.pair_j, rl.roff,
                     rl.count, rl.rlist);
  RGPU_LAUNCH_CHECK();
}

void launch_force_gather(const PairView& pv, const ReverseLists& rl, double* grad, hipStream_t stream) {
  const int64_t n = pv.n_centres;
  if (n == 0) return;
  hipLaunchKernelGGL(force_gather_kernel, dim3(unsigned((n + 3) / 4)), dim3(256), 0, stream, n, rl.roff,
                     rl.rlist, pv.pair_force, pv.centre_force, grad);
  RGPU_LAUNCH_CHECK();
}

template void launch_backward<double>(const RadialDev&, int, const PairView&, const double*,
                                      const double*, const double*, bool, const double*, double,
                                      const double*, double*, hipStream_t, const SoapTiles&);

}  // namespace rgpu
