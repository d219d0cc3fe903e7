// Context object and C ABI (include/rascal_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/rascal_mi355x.h"
#include "common.hh"
#include "device_buffer.hh"
#include "gap_f64.hh"
#include "gap_kernels.hh"
#include "gap_mfma.hh"
#include "hypers.hh"
#include "neighbours.hh"
#include "packing.hh"
#include "radial.hh"
#include "scan.hh"
#include "soap_kernels.hh"

namespace rgpu {

namespace {
thread_local std::string g_last_error;
const char* const kNoSparse =
    "no sparse points set (sparse points are dropped when a batch with another species list is "
    "set)";
}

struct Context {
  SoapHypers hp;
  std::unique_ptr<RadialHost> rh;
  RadialDev rd{};
  hipStream_t stream = nullptr;
  int precision = RGPU_FP64;

  // radial tables on the device
  DeviceBuffer<double> d_fac_b, d_O, d_center, d_hc, d_ha, d_grid, d_spl64, d_dvr_x, d_dvr_f;
  DeviceBuffer<float> d_spl32, d_ylm32;
  DeviceBuffer<double> d_ylm64;
  DeviceBuffer<HypCoef> d_hyp;

  // batch
  bool have_batch = false;
  int n_struct = 0;
  int64_t n_atoms = 0;
  std::vector<int64_t> atom_off;
  std::vector<double> cells;
  std::vector<int> pbc, types;
  std::vector<int> species;      // sorted
  std::vector<int> species_of;   // [n_atoms] index into species
  DeviceBuffer<int> d_species_of;
  DeviceBuffer<int> d_sp_tab;  // batch species -> spline table (RadialDimReduction)
  DeviceBuffer<int> d_tab_n, d_tab_cell0, d_tab_grid0;
  DeviceBuffer<YGradTerm<double>> d_ygrad64;
  DeviceBuffer<YGradTerm<float>> d_ygrad32;
  DeviceBuffer<double> d_tab_h, d_tab_pu;
  DeviceBuffer<int64_t> d_atom_off;
  PairList pl;
  // a deferred size check found the list too small and no rebuild followed
  // (the allocation-free path cannot rebuild): the next check rebuilds
  bool nl_stale = false;
  // centres: every atom, or the atoms selected by a centre mask, in atom
  // order (StructureManagerCenters puts them first,
  // structure_manager_centers.cc:37-100)
  bool masked = false;
  int64_t n_centres = 0;
  std::vector<int64_t> centre_idx;  // atom index of each centre
  DeviceBuffer<int64_t> d_centre_idx;
  // coefficient_subselection: full-layout column of each selected
  // coefficient (-1 for a species pair outside the batch's species list)
  std::vector<int64_t> sel_col;
  DeviceBuffer<int64_t> d_sel_col, d_sel_inv_off, d_sel_inv_k;  // selection and its inverse
  SelectionMap selection() const {
    SelectionMap m;
    m.col = d_sel_col.get();
    m.inv_off = d_sel_inv_off.get();
    m.inv_k = d_sel_inv_k.get();
    return m;
  }

  // representation
  bool have_rep = false;
  DeviceBuffer<double> c64, p64;
  DeviceBuffer<double> ps64;         // selected invariants [n_atoms][n_sel] (un-normalised)
  DeviceBuffer<double> gfull;        // dL/dX scattered into the full layout
  DeviceBuffer<double> dcp, dcs;     // stored expansion derivatives (gradient export)
  DeviceBuffer<float> adj_work;      // dL/dc between the adjoint's two kernels (adj_kind 3)
  // deterministic force sum: the adjoint's pair and centre forces, and the
  // reverse pair lists of the current neighbour list (rev_valid)
  DeviceBuffer<double> pair_force, centre_force;
  DeviceBuffer<int> rev_count;
  int64_t rev_count_zero = 0;  // rev_count[0, rev_count_zero) holds zeros
  // the reverse lists of a prediction are built on a second stream while
  // its GEMMs run (rev_pending: the context stream has not waited yet)
  hipStream_t side = nullptr;
  hipEvent_t ev_list = nullptr, ev_rev = nullptr;
  bool rev_pending = false;
  // the context stream waits for the side stream's reverse lists
  void join_side() {
    if (!rev_pending) return;
    RGPU_HIP(hipStreamWaitEvent(stream, ev_rev, 0));
    rev_pending = false;
  }
  DeviceBuffer<int64_t> rev_off, rev_list, rev_scan;
  bool rev_valid = false;
  // half neighbour list (rgpu_set_half_list) and its contribution rows
  DeviceBuffer<int> half_count, half_j;
  DeviceBuffer<int64_t> half_off, half_src, half_scan;
  DeviceBuffer<double> half_vec, half_dist, half_fc, half_dfc, half_rows64;
  DeviceBuffer<float> half_rows32;
  DeviceBuffer<float> c32, p32;
  DeviceBuffer<double> norm2, tmp_norm2;
  DeviceBuffer<unsigned long long> keymask;

  // sparse GAP model
  bool have_sparse = false;
  int zeta = 1;
  int64_t M = 0;
  std::vector<int> sp_species;    // species-sorted order
  std::vector<int64_t> sp_order;  // sorted position -> given position
  std::vector<int64_t> sp_off;    // per species index s: [sp_off[s], sp_off[s+1])
  bool have_weights = false;
  DeviceBuffer<double> T64, alpha;
  DeviceBuffer<float> T32;
  MixedSparse mixed_sparse;

  // per-species centre rows
  std::vector<int64_t> row_off;   // [S+1]
  DeviceBuffer<int> d_rows;

  // scratch
  DeviceBuffer<double> e_centre, grad, cvir, tmp64, out64, knm_grad, knm_scale;
  DeviceBuffer<double> e_part;  // energy GEMM partials per column tile and centre
  PinnedBuffer<double> out_stage;  // page-locked staging of a prediction's outputs
  DeviceBuffer<double> d_inv_vol;  // [n_struct] 1 / cell volume
  std::vector<double> inv_vol;     // host copy of the uploaded values
  DeviceBuffer<double> pair_fc, pair_dfc;  // cutoff function per pair
  DeviceBuffer<int> d_sp_rows;             // 0..M-1 (K_MM row gather)
  DeviceBuffer<double> wt64, G64;
  DeviceBuffer<float> G32;
  // mixed path: the invariants and W~ as BF16 operand planes; FP32
  // invariants are formed only for export and K_NM (p32_valid)
  DeviceBuffer<unsigned short> p_hi, p_lo, wth, wtl;
  bool p32_valid = false;
  DeviceBuffer<int64_t> knm_perm;

  // kernel work partitions chosen by the autotuner, per precision
  SoapTiles tiles[2];
  SoapTiles& tile() { return tiles[precision == RGPU_FP64 ? 0 : 1]; }

  // HIP graph of the representation stage (cutoff, expansion, invariants)
  bool use_graphs = false;
  hipStream_t own_stream = nullptr;  // created for capture when no stream was given
  hipGraphExec_t rep_exec = nullptr;
  std::array<uintptr_t, 29> rep_sig{};
  int64_t graph_captures = 0, graph_launches = 0;
  void drop_graph() {
    if (rep_exec) (void)hipGraphExecDestroy(rep_exec);
    rep_exec = nullptr;
    rep_sig.fill(0);
  }

  // stage timers (HIP events on `stream`)
  bool timing = false;
  struct Pending {
    hipEvent_t a, b;
    int stage;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;
  double stage_ms[RGPU_N_STAGES] = {0, 0, 0, 0, 0, 0};
  int64_t stage_n[RGPU_N_STAGES] = {0, 0, 0, 0, 0, 0};

  hipEvent_t event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    RGPU_HIP(hipEventCreate(&e));
    return e;
  }
  void collect() {
    if (pending.empty()) return;
    RGPU_HIP(hipStreamSynchronize(stream));
    for (auto& p : pending) {
      float ms = 0;
      RGPU_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      stage_ms[p.stage] += ms;
      stage_n[p.stage] += 1;
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }
  ~Context() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipEventDestroy(ev_list);
      (void)hipEventDestroy(ev_rev);
      (void)hipStreamDestroy(side);
    }
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  // feature size: the selected coefficients of a sparsified power
  // spectrum, else the full layout
  int64_t D() const {
    if (hp.kind == SoapHypers::kPowerSpectrum && hp.sparsified) return int64_t(hp.sel_a.size());
    return D_full();
  }
  int64_t D_full() const {
    const int S = int(species.size());
    const int N = hp.max_radial, L1 = hp.max_angular + 1;
    if (hp.kind == SoapHypers::kPowerSpectrum) return int64_t(S) * (S + 1) / 2 * N * N * L1;
    return int64_t(S) * N * L1 * L1;
  }
  bool sparsified() const { return hp.kind == SoapHypers::kPowerSpectrum && hp.sparsified; }
  // row stride of the packed invariants of the mixed path (PowerSpectrum):
  // the packed columns rounded up to 32, pad columns zero
  int64_t Dp() const {
    return packed_row_stride(int(species.size()), hp.max_radial, hp.max_angular + 1);
  }
};

namespace {

// Records HIP events around a stage's kernels when timing is enabled.
class StageTimer {
 public:
  StageTimer(Context& c, int stage) : c_(c), stage_(stage) {
    if (c_.timing) {
      a_ = c_.event();
      RGPU_HIP(hipEventRecord(a_, c_.stream));
    }
  }
  ~StageTimer() {
    if (!a_) return;
    hipEvent_t b = c_.event();
    if (hipEventRecord(b, c_.stream) == hipSuccess) c_.pending.push_back({a_, b, stage_});
  }

 private:
  Context& c_;
  int stage_;
  hipEvent_t a_ = nullptr;
};

void build_tables(Context& c) {
  c.rh.reset(new RadialHost(c.hp));
  RadialHost& h = *c.rh;
  const int N = h.N, L1 = h.L + 1;
  auto up = [&](auto& buf, const auto& vec) {
    using V = typename std::decay<decltype(vec)>::type::value_type;
    auto* p = buf.reserve(vec.size());
    RGPU_HIP(hipMemcpy(p, vec.data(), sizeof(V) * vec.size(), hipMemcpyHostToDevice));
    return p;
  };
  RadialDev& r = c.rd;
  r.N = N;
  r.L = h.L;
  r.L1 = L1;
  r.LL = L1 * L1;
  r.recursion = h.L >= 3 ? 1 : 0;  // hyp1f1.cc:383-390
  r.cutoff_type = c.hp.cutoff_type;
  r.rc = c.hp.cutoff;
  r.sw = c.hp.smooth_width;
  r.rs_rate = c.hp.rs_rate;
  r.rs_scale = c.hp.rs_scale;
  r.rs_exponent = c.hp.rs_exponent;
  r.fac_a = h.fac_a;
  r.fac_b = up(c.d_fac_b, h.fac_b);
  r.O = up(c.d_O, h.O);
  r.hyp = up(c.d_hyp, h.hyp);
  r.hc = up(c.d_hc, h.hc);
  r.ha = up(c.d_ha, h.ha);
  r.basis = h.basis;
  r.dvr_x = h.basis == 1 ? up(c.d_dvr_x, h.dvr_x) : nullptr;
  r.dvr_f = h.basis == 1 ? up(c.d_dvr_f, h.dvr_f) : nullptr;
  {
    const std::vector<double> yt = ylm_coefficients(h.L);
    std::vector<float> yt32(yt.begin(), yt.end());
    r.ylm64 = up(c.d_ylm64, yt);
    r.ylm32 = up(c.d_ylm32, yt32);
  }
  {
    // Gradient terms of Y_lm: coefficients of d_c (r^l Y_lm) = dY_c[lm] +
    // l r_c Y_lm over the harmonics of degree l - 1, projected with a
    // quadrature exact for these band-limited products (Gauss-Legendre in
    // cos theta, uniform in phi), from the FP64 Y_lm and gradient kernel
    const int L = h.L, LL = L1 * L1, nt = L + 2, np = 2 * L + 3;
    std::vector<double> ct, wt;
    gauss_legendre(-1.0, 1.0, nt, ct, wt);
    const int nq = nt * np;
    std::vector<double> dirs(size_t(3) * nq), wq(nq);
    for (int it = 0; it < nt; ++it)
      for (int ip = 0; ip < np; ++ip) {
        const int q = it * np + ip;
        const double ph = 2 * 3.14159265358979323846 * (ip + 0.5) / np, st = std::sqrt(1 - ct[it] * ct[it]);
        dirs[3 * q] = st * std::cos(ph);
        dirs[3 * q + 1] = st * std::sin(ph);
        dirs[3 * q + 2] = ct[it];
        wq[q] = wt[it] * 2 * 3.14159265358979323846 / np;
      }
    DeviceBuffer<double> d_dirs, d_Y, d_dY;
    double* dd = d_dirs.reserve(dirs.size());
    RGPU_HIP(hipMemcpy(dd, dirs.data(), sizeof(double) * dirs.size(), hipMemcpyHostToDevice));
    double* dY0 = d_Y.reserve(size_t(nq) * LL);
    double* dY1 = d_dY.reserve(size_t(nq) * 3 * LL);
    launch_ylm_eval<double>(r, dd, nq, dY0, dY1, nullptr);
    std::vector<double> Y(size_t(nq) * LL), dY(size_t(nq) * 3 * LL);
    RGPU_HIP(hipMemcpy(Y.data(), dY0, sizeof(double) * Y.size(), hipMemcpyDeviceToHost));
    RGPU_HIP(hipMemcpy(dY.data(), dY1, sizeof(double) * dY.size(), hipMemcpyDeviceToHost));
    std::vector<YGradTerm<double>> g64(size_t(3) * LL, YGradTerm<double>{0.0, 0.0, 0, 0});
    for (int l = 1; l <= L; ++l)
      for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm)
        for (int cc = 0; cc < 3; ++cc) {
          std::vector<std::pair<double, int>> terms;
          double amax = 0;
          std::vector<double> a(2 * l - 1);
          for (int k = 0; k < 2 * l - 1; ++k) {
            const int lp = (l - 1) * (l - 1) + k;
            double num = 0, den = 0;
            for (int q = 0; q < nq; ++q) {
              const double f = dY[(size_t(q) * 3 + cc) * LL + lm] + l * dirs[3 * q + cc] * Y[size_t(q) * LL + lm];
              num += wq[q] * f * Y[size_t(q) * LL + lp];
              den += wq[q] * Y[size_t(q) * LL + lp] * Y[size_t(q) * LL + lp];
            }
            a[k] = num / den;
            amax = std::max(amax, std::fabs(a[k]));
          }
          for (int k = 0; k < 2 * l - 1; ++k)
            // the nonzero coefficients are O(1) and above (d_z (x +- iy)^l is
            // exactly zero), so round-off is cut on an absolute scale
            if (std::fabs(a[k]) > 1e-9 * std::max(amax, 1.0)) terms.push_back({a[k], (l - 1) * (l - 1) + k});
          if (terms.size() > 2) throw RuntimeError("Y_lm gradient identity: more than two terms");
          YGradTerm<double>& e = g64[size_t(lm) * 3 + cc];
          if (terms.size() > 0) e.c0 = terms[0].first, e.i0 = terms[0].second;
          if (terms.size() > 1) e.c1 = terms[1].first, e.i1 = terms[1].second;
          // the identity reproduces the gradients at every quadrature point
          double err = 0, ref = 0;
          for (int q = 0; q < nq; ++q) {
            const double* yq = &Y[size_t(q) * LL];
            const double v = e.c0 * yq[e.i0] + e.c1 * yq[e.i1] - l * dirs[3 * q + cc] * yq[lm];
            const double d = dY[(size_t(q) * 3 + cc) * LL + lm];
            err = std::max(err, std::fabs(v - d));
            ref = std::max(ref, std::fabs(d));
          }
          if (err > 1e-11 * std::max(ref, 1.0)) throw RuntimeError("Y_lm gradient identity does not hold");
        }
    std::vector<YGradTerm<float>> g32(g64.size());
    for (size_t k = 0; k < g64.size(); ++k)
      g32[k] = YGradTerm<float>{float(g64[k].c0), float(g64[k].c1), g64[k].i0, g64[k].i1};
    r.ygrad64 = up(c.d_ygrad64, g64);
    r.ygrad32 = up(c.d_ygrad32, g32);
  }
  r.use_spline = 0;
  r.n_tabs = h.n_tabs;
  r.sp_tab = nullptr;  // set with the batch's species
  // finalised: R is finalised per pair; one centre term per table
  r.center = up(c.d_center, h.tab_centers);
  if (c.hp.optimization != 1) return;

  // Spline of O^T R(r) (the DVR functions as they are: O = identity; with
  // RadialDimReduction one spline per species of its projection P_k[l]
  // O^T R[:, l], calculator_spherical_expansion.hh:1431-1454) on [1e-5, rc]
  // (:1107-1135): grids of 2<<k points, k = 5, 6, ..., refined until the
  // mean absolute error at the cell midpoints is below the accuracy
  // (math/interpolator.hh:705-712, 1229-1248).  Function values come from
  // the FP64 direct kernel; the tridiagonal solve runs on the host.
  const int F = N * L1;
  const double x1 = 1e-5, x2 = c.hp.cutoff;
  DeviceBuffer<double> d_r, d_raw, d_fin;
  auto eval_direct = [&](const std::vector<double>& xs, std::vector<double>& out) {
    int64_t n = int64_t(xs.size());
    double* dr = d_r.reserve(n);
    RGPU_HIP(hipMemcpy(dr, xs.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    double* raw = d_raw.reserve(size_t(n) * F);
    double* fin = d_fin.reserve(size_t(n) * F);
    launch_radial_direct(r, dr, n, raw, nullptr, nullptr);
    launch_radial_finalise(r, raw, n, fin, nullptr);
    out.resize(size_t(n) * F);
    RGPU_HIP(hipMemcpy(out.data(), fin, sizeof(double) * n * F, hipMemcpyDeviceToHost));
  };
  // values of table t at the points xs: out[pt][c * L1 + l]
  auto eval_table = [&](int t, const std::vector<double>& xs, std::vector<double>& out) {
    eval_direct(xs, out);
    if (h.tab_species.empty()) return;
    std::vector<double> fin = out;
    for (size_t pt = 0; pt < xs.size(); ++pt)
      for (int l = 0; l < L1; ++l)
        for (int cc = 0; cc < N; ++cc) {
          const double* P = &h.proj[((size_t(t) * L1 + l) * N + cc) * N];
          double v = 0;
          for (int n = 0; n < N; ++n) v += P[n] * fin[pt * F + n * L1 + l];
          out[pt * F + cc * L1 + l] = v;
        }
  };
  struct Fit {
    std::vector<double> grid, y, d2;
    double h = 0, per_unit = 0;
  };
  // spline of table t on 2 << dof points (fixed_npts > 0: that grid, no
  // refinement)
  auto fit_table = [&](int t, int fixed_npts) {
    Fit f;
    std::vector<double> test, ref;
    int dof = 5;
    while (true) {
      const int npts = fixed_npts > 0 ? fixed_npts : 2 << dof;
      f.grid = linspaced(npts, x1, x2);
      eval_table(t, f.grid, f.y);
      f.h = f.grid[1] - f.grid[0];
      spline_second_derivatives(npts, F, f.h, f.y, f.d2);
      f.per_unit = double(npts) / (f.grid[npts - 1] - f.grid[0]);
      if (fixed_npts > 0) return f;
      const int tdof = std::max(dof, 5);
      const int tn = 2 << tdof;
      const double off = (x2 - x1) / (2 * (tn - 1));
      test = linspaced(tn - 1, x1 + off, x2 - off);
      eval_table(t, test, ref);
      double err = 0;
      for (size_t q = 0; q < test.size(); ++q) {
        const double x = test[q];
        int k = std::max(0, std::min(npts - 2, int((x - x1) * f.per_unit) - 1));
        const double a = (f.grid[k + 1] - x) / f.h, b = 1 - a;
        for (int ff = 0; ff < F; ++ff) {
          const double v = a * (f.y[size_t(k) * F + ff] + (a * a - 1) * f.d2[size_t(k) * F + ff]) +
                           b * (f.y[size_t(k + 1) * F + ff] + (b * b - 1) * f.d2[size_t(k + 1) * F + ff]);
          err += std::fabs(v - ref[q * F + ff]);
        }
      }
      err /= double(test.size()) * F;
      if (err < c.hp.spline_accuracy || npts >= 100000) return f;
      ++dof;
    }
  };
  // one spline per table, each refined on its own (as the reference keeps
  // one interpolator per species), the tables stored one after another
  std::vector<Fit> fits;
  for (int t = 0; t < h.n_tabs; ++t) fits.push_back(fit_table(t, 0));
  std::vector<int> tab_n, tab_cell0, tab_grid0;
  std::vector<double> tab_h, tab_pu, grids;
  int cells = 0;
  for (const Fit& f : fits) {
    tab_n.push_back(int(f.grid.size()));
    tab_cell0.push_back(cells);
    tab_grid0.push_back(int(grids.size()));
    tab_h.push_back(f.h);
    tab_pu.push_back(f.per_unit);
    grids.insert(grids.end(), f.grid.begin(), f.grid.end());
    cells += int(f.grid.size()) - 1;
  }
  // packed per cell: {y_k, d_k, y_k+1, d_k+1}[f]
  std::vector<double> tab(size_t(cells) * F * 4);
  for (int t = 0; t < h.n_tabs; ++t)
    for (int k = 0; k + 1 < tab_n[t]; ++k)
      for (int ff = 0; ff < F; ++ff) {
        double* e = &tab[((size_t(tab_cell0[t]) + k) * F + ff) * 4];
        e[0] = fits[t].y[size_t(k) * F + ff];
        e[1] = fits[t].d2[size_t(k) * F + ff];
        e[2] = fits[t].y[size_t(k + 1) * F + ff];
        e[3] = fits[t].d2[size_t(k + 1) * F + ff];
      }
  std::vector<float> tab32(tab.begin(), tab.end());
  const Fit& f0 = fits[0];
  const int npts = tab_n[0];
  h.grid_n = npts;
  h.x1 = x1;
  h.x2 = x2;
  h.h = f0.h;
  h.per_unit = f0.per_unit;
  h.spl64 = tab;
  r.use_spline = 1;
  // the scalar grid is table 0's (the only one without RadialDimReduction)
  r.grid_n = npts;
  r.search_size = npts - 2;
  r.x1 = x1;
  r.h = f0.h;
  r.per_unit = f0.per_unit;
  r.grid = up(c.d_grid, grids);
  r.spl64 = up(c.d_spl64, tab);
  r.spl32 = up(c.d_spl32, tab32);
  r.tab_n = up(c.d_tab_n, tab_n);
  r.tab_cell0 = up(c.d_tab_cell0, tab_cell0);
  r.tab_grid0 = up(c.d_tab_grid0, tab_grid0);
  r.tab_h = up(c.d_tab_h, tab_h);
  r.tab_pu = up(c.d_tab_pu, tab_pu);
  r.tab_grid = r.grid;
}

// the batch is set; a neighbour list built with its size check deferred
// may still be unchecked (only the prediction path works on such a list)
void require_batch_unsettled(const Context& c) {
  if (!c.have_batch) throw InvalidArgument("no structures set (call rgpu_set_structures first)");
}

// Completes a deferred neighbour-list build; when a capacity was short, the
// list is built again (synchronously) and the representation, computed from
// a list without pairs, is invalidated.  Returns false in that case.
bool settle_nl(Context& c) {
  if (!c.nl_stale && finish_pair_list(c.pl, c.stream)) return true;
  c.nl_stale = false;
  c.join_side();  // the side stream may still read the list
  rebuild_pair_list(c.pl, c.hp.cutoff, c.stream);
  c.have_rep = false;
  c.rev_valid = false;
  return false;
}

void require_batch(Context& c) {
  require_batch_unsettled(c);
  settle_nl(c);
}

double cell_volume(const double* C) {
  return std::fabs(C[0] * (C[4] * C[8] - C[5] * C[7]) - C[1] * (C[3] * C[8] - C[5] * C[6]) +
                   C[2] * (C[3] * C[7] - C[4] * C[6]));
}

void build_batch_nl(Context& c, const double* positions, bool defer = false) {
  c.join_side();  // the side stream may still read the list
  StageTimer t(c, RGPU_STAGE_NEIGHBOURS);
  c.have_rep = false;
  c.rev_valid = false;
  if (c.n_atoms == 0) {
    // empty batch: only the species list (feature layout) is defined
    c.pl.n_atoms = c.pl.n_pairs = 0;
    int64_t* off = c.pl.centre_off.reserve(1);
    RGPU_HIP(hipMemsetAsync(off, 0, sizeof(int64_t), c.stream));
    return;
  }
  build_pair_list(c.pl, c.n_struct, c.atom_off.data(), positions, c.cells.data(), c.pbc.data(),
                  c.hp.cutoff, c.stream, defer, c.use_graphs);
  // cell volumes for the stress (uploaded when the cells change)
  std::vector<double> iv(static_cast<size_t>(c.n_struct));
  for (int s = 0; s < c.n_struct; ++s) iv[size_t(s)] = 1.0 / cell_volume(&c.cells[9 * s]);
  const double* before = c.d_inv_vol.get();
  double* d = c.d_inv_vol.reserve(iv.size());
  if (d != before || iv != c.inv_vol) {
    RGPU_HIP(hipMemcpy(d, iv.data(), sizeof(double) * iv.size(), hipMemcpyHostToDevice));
    c.inv_vol = iv;
  }
}

PairView pair_view(Context& c) {
  PairView pv;
  pv.n_centres = c.n_atoms;
  pv.centre_off = c.pl.centre_off.get();
  pv.pair_j = c.pl.pair_j.get();
  pv.pair_vec = c.pl.pair_vec.get();
  pv.pair_dist = c.pl.pair_dist.get();
  pv.species_of = c.d_species_of.get();
  pv.pair_fc = c.pair_fc.get();
  pv.pair_dfc = c.pair_dfc.get();
  return pv;
}

// Room for the reverse pair lists of the current list (the same buffers on
// every call for one list)
void reverse_buffers(Context& c, ReverseLists* rl) {
  const int64_t n = c.n_atoms, cap = std::max<int64_t>(n > 0 ? c.pl.cap_pairs : 0, 1);
  const size_t count_cap = c.rev_count.capacity();
  rl->count = c.rev_count.reserve(size_t(std::max<int64_t>(n, 1)));
  if (c.rev_count.capacity() != count_cap) c.rev_count_zero = 0;  // fresh memory
  rl->roff = c.rev_off.reserve(size_t(n + 1));
  rl->rlist = c.rev_list.reserve(size_t(cap));
}

// Builds the reverse pair lists of the current list on stream s (after the
// list itself on that stream).
void build_reverse_lists(Context& c, const PairView& pv, const ReverseLists& rl, hipStream_t s) {
  // the fill takes every count it raised back to zero (a launch that throws
  // leaves the counts unknown)
  const int64_t n = c.n_atoms, zero = c.rev_count_zero;
  c.rev_count_zero = 0;
  launch_reverse_pairs(pv, rl, c.rev_scan, s, n > zero);
  c.rev_count_zero = std::max(zero, n);
  c.rev_valid = true;
}

// Starts the reverse pair lists of the current list on the side stream,
// after the work already on the context stream, so that they are built
// while the representation stage runs; backward_view makes the context
// stream wait for them.
void prefetch_reverse_lists(Context& c) {
  if (c.rev_valid || c.n_atoms == 0) return;
  if (!c.side) {
    RGPU_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    RGPU_HIP(hipEventCreateWithFlags(&c.ev_list, hipEventDisableTiming));
    RGPU_HIP(hipEventCreateWithFlags(&c.ev_rev, hipEventDisableTiming));
  }
  ReverseLists rl;
  reverse_buffers(c, &rl);
  exclusive_scan_reserve(c.n_atoms, c.rev_scan);  // no allocation on the side stream's way
  RGPU_HIP(hipEventRecord(c.ev_list, c.stream));
  RGPU_HIP(hipStreamWaitEvent(c.side, c.ev_list, 0));
  build_reverse_lists(c, pair_view(c), rl, c.side);
  RGPU_HIP(hipEventRecord(c.ev_rev, c.side));
  c.rev_pending = true;
}

// The pair view of an adjoint launch: room for its pair and centre forces,
// and the reverse pair lists of the current list (built once per list)
PairView backward_view(Context& c, ReverseLists* rl) {
  PairView pv = pair_view(c);
  const int64_t n = c.n_atoms, cap = std::max<int64_t>(n > 0 ? c.pl.cap_pairs : 0, 1);
  pv.pair_force = c.pair_force.reserve(size_t(3 * cap));
  pv.centre_force = c.centre_force.reserve(size_t(std::max<int64_t>(3 * n, 1)));
  reverse_buffers(c, rl);
  if (!c.rev_valid) build_reverse_lists(c, pv, *rl, c.stream);
  c.join_side();
  return pv;
}

template <class T>
DeviceBuffer<T>& cbuf(Context& c);
template <>
DeviceBuffer<double>& cbuf<double>(Context& c) {
  return c.c64;
}
template <>
DeviceBuffer<float>& cbuf<float>(Context& c) {
  return c.c32;
}
template <class T>
DeviceBuffer<T>& pbuf(Context& c);
template <>
DeviceBuffer<double>& pbuf<double>(Context& c) {
  return c.p64;
}
template <>
DeviceBuffer<float>& pbuf<float>(Context& c) {
  return c.p32;
}

void check_half_list(const Context& c);

template <class T>
void compute_rep(Context& c) {
  const int S = int(c.species.size());
  const int64_t n = c.n_atoms;
  const int64_t E = int64_t(S) * c.rd.N * c.rd.LL;
  const int64_t D = c.D_full();
  const bool ps = c.hp.kind == SoapHypers::kPowerSpectrum;
  const int64_t cap = n > 0 ? c.pl.cap_pairs : 0;
  constexpr bool split = std::is_same<T, float>::value;
  if (split && c.sparsified())
    throw InvalidArgument("coefficient_subselection is computed in FP64 only (set RGPU_FP64)");
  T* cc = cbuf<T>(c).reserve(size_t(std::max<int64_t>(n * E, 1)));
  unsigned long long* km = c.keymask.reserve(size_t(std::max<int64_t>(n, 1)));
  double* fc = c.pair_fc.reserve(size_t(std::max<int64_t>(cap, 1)));
  double* dfc = c.pair_dfc.reserve(size_t(std::max<int64_t>(cap, 1)));
  // FP64 keeps the invariants in FP64; the mixed path stores them as the
  // BF16 hi/lo operand planes of the matrix-core GEMMs
  T* pp = ps && !split ? pbuf<T>(c).reserve(size_t(std::max<int64_t>(n * D, 1))) : nullptr;
  const int64_t Dp = c.Dp();
  unsigned short* ph = ps && split ? c.p_hi.reserve(size_t(std::max<int64_t>(n * Dp, 1))) : nullptr;
  unsigned short* pl = ps && split ? c.p_lo.reserve(size_t(std::max<int64_t>(n * Dp, 1))) : nullptr;
  double* n2 = ps ? c.norm2.reserve(size_t(std::max<int64_t>(n, 1))) : nullptr;
  // sparsified: the selected coefficients and their norms (p keeps S^T xs)
  double* xs = split || !c.sparsified() ? nullptr
                                        : c.ps64.reserve(size_t(std::max<int64_t>(n * c.D(), 1)));
  c.p32_valid = false;
  const int64_t* n_pairs_dev = c.pl.centre_off.get() + n;
  // the half list's buffers are sized for the full list, before any capture
  HalfList hl{};
  if (c.tile().half_nl && n > 0) {
    check_half_list(c);
    const size_t cp = size_t(std::max<int64_t>(cap, 1)), B = size_t(c.rd.N) * c.rd.LL;
    hl.count = c.half_count.reserve(size_t(n));
    hl.off = c.half_off.reserve(size_t(n + 1));
    hl.src = c.half_src.reserve(cp);
    hl.j = c.half_j.reserve(cp);
    hl.vec = c.half_vec.reserve(3 * cp);
    hl.dist = c.half_dist.reserve(cp);
    hl.fc = c.half_fc.reserve(cp);
    hl.dfc = c.half_dfc.reserve(cp);
    if constexpr (std::is_same<T, float>::value)
      hl.rows = c.half_rows32.reserve(cp * B);
    else
      hl.rows = c.half_rows64.reserve(cp * B);
    reserve_half_list_scan(n, c.half_scan);
    hl.scan_scratch = &c.half_scan;
  }
  // expansion and invariants in one persistent launch (c kept in LDS
  // between them) unless the half list or separate launches are selected
  const bool fused = ps && !hl.rows && c.tile().fused == 1;
  auto cutoff_and_expansion = [&] {
    launch_pair_cutoff(c.rd, n_pairs_dev, cap, c.pl.pair_dist.get(), fc, dfc, c.stream);
    if (fused)
      launch_expansion_ps<T>(c.rd, S, pair_view(c), cc, km, pp, ph, pl, n2, c.stream, c.tile());
    else if (!hl.rows && c.tile().fwd_lds == 1)
      launch_expansion_ps<T>(c.rd, S, pair_view(c), cc, km, nullptr, nullptr, nullptr, nullptr,
                             c.stream, c.tile());
    else if (hl.rows)
      launch_expansion_half<T>(c.rd, S, pair_view(c), c.pl.pos.get(), hl, cc, km, c.stream, c.tile());
    else
      launch_expansion<T>(c.rd, S, pair_view(c), cc, km, c.stream, c.tile());
  };
  auto invariants = [&] {
    if (!ps) return;
    if (fused) {
      // (the invariants came with the expansion: the selection only)
      if constexpr (!split)
        if (xs) launch_sparsify(n, D, c.D(), c.selection(), pp, xs, n2, c.stream);
      return;
    }
    if constexpr (split) {
      launch_power_spectrum_split(c.rd, S, n, cc, ph, pl, n2, c.stream, c.tile());
    } else {
      launch_power_spectrum<T>(c.rd, S, n, cc, pp, n2, c.stream, c.tile());
      if (xs) launch_sparsify(n, D, c.D(), c.selection(), pp, xs, n2, c.stream);
    }
  };
  std::array<uintptr_t, 29> sig{};
  bool replay = c.use_graphs && !c.timing && n > 0;
  if (replay) {
    sig = {
        uintptr_t(c.precision), uintptr_t(S), uintptr_t(n), uintptr_t(cap), uintptr_t(cc),
        uintptr_t(km), uintptr_t(fc), uintptr_t(dfc), uintptr_t(pp), uintptr_t(ph), uintptr_t(xs),
        uintptr_t(pl), uintptr_t(n2), uintptr_t(c.pl.centre_off.get()),
        uintptr_t(c.pl.pair_j.get()), uintptr_t(c.pl.pair_vec.get()), uintptr_t(c.pl.pair_dist.get()),
        uintptr_t(c.d_species_of.get()), uintptr_t(c.tile().pb_fwd), uintptr_t(c.tile().kc),
        uintptr_t(c.d_sel_col.get()), uintptr_t(c.tile().half_nl), uintptr_t(hl.src),
        uintptr_t(hl.rows), uintptr_t(c.tile().fused), uintptr_t(c.tile().fwd_lds),
        uintptr_t(c.tile().fwd_group), uintptr_t(c.d_sel_inv_off.get()), uintptr_t(c.d_sel_inv_k.get())};
    // (instantiating a graph allocates: the allocation-free path launches
    // the stage eagerly instead of capturing it)
    if ((!c.rep_exec || sig != c.rep_sig) && g_no_alloc) replay = false;
  }
  if (replay) {
    // replay a captured graph of the stage while every buffer, size and
    // precision it was captured with is unchanged
    if (!c.rep_exec || sig != c.rep_sig) {
      c.drop_graph();
      hipGraph_t g = nullptr;
      RGPU_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      try {
        cutoff_and_expansion();
        invariants();
      } catch (...) {
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      RGPU_HIP(hipStreamEndCapture(c.stream, &g));
      const hipError_t e = hipGraphInstantiate(&c.rep_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      RGPU_HIP(e);
      c.rep_sig = sig;
      ++c.graph_captures;
    }
    RGPU_HIP(hipGraphLaunch(c.rep_exec, c.stream));
    ++c.graph_launches;
  } else {
    {
      StageTimer t(c, RGPU_STAGE_EXPANSION);
      cutoff_and_expansion();
    }
    {
      StageTimer t(c, RGPU_STAGE_POWER_SPECTRUM);
      invariants();
    }
  }
  c.have_rep = true;
}

// (callers other than the prediction have settled the neighbour list)
void compute(Context& c);

// The allocation-free GAP energy keeps its per-centre energies and the
// energy GEMM's partial sums in this file-scope device array (each device
// holds its own copy until the process ends).  Host code never uses the array's name as an address: the address
// on each device comes from hipGetSymbolAddress, once per device, with that
// device current.
constexpr int64_t kScratchDoubles = int64_t(1) << 23;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

double* scratch_on_current_device() {
  static std::mutex mu;
  static double* ptr[64] = {};
  int dev = 0;
  RGPU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw RuntimeError("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    RGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = static_cast<double*>(p);
  }
  return ptr[dev];
}

void compute(Context& c) {
  require_batch_unsettled(c);
  if (c.precision == RGPU_FP64)
    compute_rep<double>(c);
  else
    compute_rep<float>(c);
}

void ensure_rep(Context& c) {
  if (!c.have_rep) compute(c);
}

void require_invariants(const Context& c) {
  if (c.hp.kind != SoapHypers::kPowerSpectrum)
    throw InvalidArgument("this operation needs SphericalInvariants (PowerSpectrum) hypers");
}

// calculator_spherical_expansion.hh:2159-2185: the half list needs every
// atom a centre and periodic cell vectors longer than the expansion's
// diameter
void check_half_list(const Context& c) {
  if (c.masked)
    throw RuntimeError(
        "Half neighbor list should only be used when all the atoms inside the unit cell are "
        "centers, i.e. center_atoms_mask should not mask atoms.");
  for (int s = 0; s < c.n_struct; ++s) {
    const double* C = &c.cells[9 * size_t(s)];
    double len[3];
    bool small = false;
    for (int k = 0; k < 3; ++k) {
      len[k] = std::sqrt(C[k] * C[k] + C[3 + k] * C[3 + k] + C[6 + k] * C[6 + k]);
      if (c.pbc[3 * size_t(s) + k] && len[k] < 2. * c.hp.cutoff) small = true;
    }
    if (small)
      throw RuntimeError(
          "Half neighbor list should only be used when the diameter of the spherical expansion is "
          "smaller than the unit cell in periodic directions: [" + std::to_string(len[0]) + " " +
          std::to_string(len[1]) + " " + std::to_string(len[2]) + "] > " +
          std::to_string(2 * c.hp.cutoff));
  }
}

// calculator_spherical_expansion.hh:2155-2158
void require_unmasked(const Context& c) {
  if (c.masked)
    throw InvalidArgument("Can't compute spherical expansion gradients with masked center atoms");
}

// GAP species blocks of the current batch (the centres' rows)
void build_rows(Context& c) {
  const int S = int(c.species.size());
  c.row_off.assign(S + 1, 0);
  std::vector<int> rows;
  rows.reserve(size_t(c.n_atoms));
  for (int s = 0; s < S; ++s) {
    for (int64_t i : c.centre_idx)
      if (c.species_of[i] == s) rows.push_back(int(i));
    c.row_off[s + 1] = int64_t(rows.size());
  }
  int* d = c.d_rows.reserve(std::max<size_t>(rows.size(), 1));
  if (!rows.empty())
    RGPU_HIP(hipMemcpyAsync(d, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, c.stream));
}

template <class T>
GapBlock<T> gap_block(Context& c, int s, const T* X, const T* Tm) {
  GapBlock<T> b;
  b.n_rows = c.row_off[s + 1] - c.row_off[s];
  b.row_idx = c.d_rows.get() + c.row_off[s];
  b.n_cols = c.sp_off[s + 1] - c.sp_off[s];
  b.col0 = c.sp_off[s];
  b.X = X;
  b.norm2 = c.norm2.get();
  b.normalize = c.hp.normalize ? 1 : 0;
  b.Tm = Tm;
  b.alpha = c.have_weights ? c.alpha.get() : nullptr;
  b.zeta = c.zeta;
  b.D = c.D();
  return b;
}

template <class T>
const T* sparse_values(Context& c);
template <>
const double* sparse_values<double>(Context& c) {
  return c.T64.get();
}
template <>
const float* sparse_values<float>(Context& c) {
  return c.T32.get();
}
template <class T>
DeviceBuffer<T>& Gbuf(Context& c);
template <>
DeviceBuffer<double>& Gbuf<double>(Context& c) {
  return c.G64;
}
template <>
DeviceBuffer<float>& Gbuf<float>(Context& c) {
  return c.G32;
}
// Backward from G (device, [n][D] in precision T) to grad/virial (device).
// xg (may be null): X_i . G_i per centre, times xg_scale.
// (FP64, sparsified: G holds the selected columns and is scattered into the
// full layout first)
template <class T>
void backward_device(Context& c, const T* G, double* grad, double* cvir, const double* xg = nullptr,
                     double xg_scale = 1.0) {
  const int S = int(c.species.size());
  StageTimer t(c, RGPU_STAGE_BACKWARD);
  ReverseLists rl;
  const PairView pv = backward_view(c, &rl);
  if constexpr (std::is_same<T, double>::value) {
    if (c.sparsified()) {
      double* gf = c.gfull.reserve(size_t(std::max<int64_t>(c.n_atoms * c.D_full(), 1)));
      launch_scatter_columns(c.n_atoms, c.D(), c.selection(), G, c.D_full(), gf, c.stream);
      G = gf;
    }
  }
  if constexpr (std::is_same<T, float>::value) {
    float* work = nullptr;
    if (c.tile().adj_kind == 3)
      work = c.adj_work.reserve(size_t(std::max<int64_t>(c.n_atoms * S * c.rd.N * c.rd.LL, 1)));
    launch_backward_split(c.rd, S, pv, c.c32.get(), c.p_hi.get(), c.p_lo.get(), c.norm2.get(),
                          c.hp.normalize, xg, xg_scale, G, cvir, c.stream, c.tile(), work);
  } else {
    launch_backward<T>(c.rd, S, pv, cbuf<T>(c).get(), pbuf<T>(c).get(), c.norm2.get(),
                       c.hp.normalize, xg, xg_scale, G, cvir, c.stream, c.tile());
  }
  launch_force_gather(pv, rl, grad, c.stream);
}

// FP32 invariants of the mixed path in the full layout (for export and
// K_NM), formed on demand from the stored expansion (the norms of the packed
// pass are kept; this pass's equal them to rounding)
const float* ensure_p32(Context& c) {
  if (!c.p32_valid) {
    const int64_t n = c.n_atoms;
    float* p = c.p32.reserve(size_t(std::max<int64_t>(n * c.D(), 1)));
    double* n2 = c.tmp_norm2.reserve(size_t(std::max<int64_t>(n, 1)));
    if (n > 0)
      launch_power_spectrum<float>(c.rd, int(c.species.size()), n, c.c32.get(), p, n2, c.stream,
                                   c.tile());
    c.p32_valid = true;
  }
  return c.p32.get();
}

template <class T>
const T* invariants_of(Context& c);
template <>
const double* invariants_of<double>(Context& c) {
  return c.sparsified() ? c.ps64.get() : c.p64.get();
}
template <>
const float* invariants_of<float>(Context& c) {
  return ensure_p32(c);
}


// Outputs of a prediction: per-structure energies from per-centre energies
// e_centre (device, may be null), gradient from device grad [n_atoms][3],
// neg_stress from per-centre virials (summed per structure and divided by
// the cell volume on the device).  One synchronisation at the end.
// true for device (or managed) memory; host memory, registered or not, is
// false
bool is_device_memory(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable host memory is unknown to the runtime
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void emit_outputs(Context& c, const double* e_centre, double* energy, double* grad_d,
                  double* cvir_d, double* gradient, double* neg_stress) {
  const int ns = c.n_struct;
  // the reductions and the gradient copy write straight into page-locked
  // staging ([energy | gradient | stress], zeroed here: the host owns it
  // between predictions); one wait, then the caller's arrays (host arrays
  // by memcpy, device arrays by a copy enqueued before the wait)
  const size_t ne = energy && e_centre ? size_t(ns) : 0,
               ng = gradient && grad_d ? size_t(3 * c.n_atoms) : 0, nst = neg_stress ? size_t(6 * ns) : 0;
  double* h = c.out_stage.reserve(std::max<size_t>(ne + ng + nst, 1));
  double* hd = c.out_stage.device();
  std::memset(h, 0, sizeof(double) * (ne + ng + nst));
  if (ne && nst)
    launch_segment_sum2(e_centre, 1, hd, nullptr, cvir_d, 6, hd + ne + ng, c.d_inv_vol.get(), c.d_atom_off.get(), ns,
                        c.stream);
  else if (ne)
    launch_segment_sum(e_centre, 1, c.d_atom_off.get(), ns, hd, c.stream);
  else if (nst)
    launch_segment_sum(cvir_d, 6, c.d_atom_off.get(), ns, hd + ne + ng, c.stream, c.d_inv_vol.get());
  if (ng) launch_strided_copy(grad_d, int64_t(ng), hd + ne, 1, c.stream);
  struct Out {
    double* dst;
    size_t off, n;
    bool device;
  } outs[3] = {{energy, 0, ne, ne && is_device_memory(energy)},
               {gradient, ne, ng, ng && is_device_memory(gradient)},
               {neg_stress, ne + ng, nst, nst && is_device_memory(neg_stress)}};
  for (const Out& o : outs)
    if (o.n && o.device)
      RGPU_HIP(hipMemcpyAsync(o.dst, hd + o.off, sizeof(double) * o.n, hipMemcpyDefault, c.stream));
  RGPU_HIP(hipStreamSynchronize(c.stream));
  for (const Out& o : outs)
    if (o.n && !o.device) std::memcpy(o.dst, h + o.off, sizeof(double) * o.n);
}

// column tiles of the widest species block in the energy GEMM of precision
// T (the partial row sums it keeps per centre)
template <class T>
int energy_tiles(const Context& c) {
  int t = 0;
  for (size_t s = 0; s + 1 < c.sp_off.size(); ++s) {
    const int64_t m = c.sp_off[s + 1] - c.sp_off[s];
    t = std::max(t, std::is_same<T, float>::value ? gap_mixed_energy_tiles(m) : gap_f64_energy_tiles(m));
  }
  return t;
}

template <class T>
void gap_predict_t(Context& c, double* energy, double* gradient, double* neg_stress) {
  const int S = int(c.species.size());
  // the mixed path works on the packed invariants
  const int64_t n = c.n_atoms, D = std::is_same<T, float>::value ? c.Dp() : c.D();
  const bool want_grad = gradient || neg_stress;
  if (want_grad) require_unmasked(c);
  double* ec = c.e_centre.reserve(size_t(std::max<int64_t>(n, 1)));
  double* part = c.e_part.reserve(size_t(std::max<int64_t>(n * energy_tiles<T>(c), 1)));
  RGPU_HIP(hipMemsetAsync(ec, 0, sizeof(double) * n, c.stream));
  const T* X = nullptr;
  if constexpr (std::is_same<T, double>::value) X = invariants_of<double>(c);
  std::vector<GapBlock<T>> blks;
  std::vector<int64_t> wt_off;
  int64_t wt_total = 0;
  bool rows_without_points = false;
  for (int s = 0; s < S; ++s) {
    blks.push_back(gap_block<T>(c, s, X, sparse_values<T>(c)));
    blks.back().D = D;
    wt_off.push_back(wt_total);
    // blocks start on 8-element (16-byte) boundaries
    wt_total += (blks.back().n_rows * blks.back().n_cols + 7) / 8 * 8;
    if (blks.back().n_rows > 0 && blks.back().n_cols == 0) rows_without_points = true;
  }
  T* G = nullptr;
  if (want_grad) {
    // the force product writes every row of a species with sparse points
    G = Gbuf<T>(c).reserve(size_t(std::max<int64_t>(n * D, 1)));
    if (rows_without_points) RGPU_HIP(hipMemsetAsync(G, 0, sizeof(T) * n * D, c.stream));
  }
  if constexpr (std::is_same<T, float>::value) {
    std::vector<SplitPlanes> wts(static_cast<size_t>(S));
    if (want_grad) {
      unsigned short* wh = c.wth.reserve(size_t(std::max<int64_t>(wt_total, 1)));
      unsigned short* wl = c.wtl.reserve(size_t(std::max<int64_t>(wt_total, 1)));
      for (int s = 0; s < S; ++s) wts[s] = SplitPlanes{wh + wt_off[s], wl + wt_off[s]};
    }
    const SplitPlanes xs{c.p_hi.get(), c.p_lo.get()};
    {
      StageTimer t(c, RGPU_STAGE_GAP_KERNEL);
      run_gap_mixed_kernel(blks.data(), S, c.mixed_sparse, xs, ec, part, n, want_grad ? wts.data() : nullptr,
                           c.stream);
    }
    if (want_grad) {
      StageTimer t(c, RGPU_STAGE_GAP_FORCE);
      run_gap_mixed_force(blks.data(), S, c.mixed_sparse, wts.data(), G, c.stream);
    }
  } else {
    double* wt_all = want_grad ? c.wt64.reserve(size_t(std::max<int64_t>(wt_total, 1))) : nullptr;
    std::vector<double*> wts(static_cast<size_t>(S), nullptr);
    if (want_grad)
      for (int s = 0; s < S; ++s) wts[s] = wt_all + wt_off[s];
    {
      StageTimer t(c, RGPU_STAGE_GAP_KERNEL);
      run_gap_f64_kernel(blks.data(), S, ec, part, n, want_grad ? wts.data() : nullptr, c.stream);
    }
    if (want_grad) {
      StageTimer t(c, RGPU_STAGE_GAP_FORCE);
      run_gap_f64_force(blks.data(), S, wts.data(), G, c.stream);
    }
  }
  double *gd = nullptr, *cv = nullptr;
  if (want_grad) {
    gd = c.grad.reserve(size_t(std::max<int64_t>(3 * n, 1)));
    cv = c.cvir.reserve(size_t(std::max<int64_t>(6 * n, 1)));
    // (the force gather writes every atom's gradient)
    // the GAP energy is homogeneous of degree zeta in X, so X_i . G_i = zeta E_i
    backward_device<T>(c, G, gd, cv, ec, double(c.zeta));
  }
  emit_outputs(c, ec, energy, gd, cv, gradient, neg_stress);
}

// Per-structure GAP energies into energy (device memory) with the
// per-centre energies in ec and the GEMM partials in part (both in the
// scratch array): no buffer of the context grows and nothing is allocated
// (the caller has set the scope).
template <class T>
void gap_energy_noalloc_t(Context& c, double* ec, double* part, double* energy) {
  const int S = int(c.species.size());
  const int64_t n = c.n_atoms, D = std::is_same<T, float>::value ? c.Dp() : c.D();
  launch_fill(ec, n, 0.0, c.stream);
  const T* X = nullptr;
  if constexpr (std::is_same<T, double>::value) X = invariants_of<double>(c);
  std::array<GapBlock<T>, kMaxSpecies> blks{};
  for (int s = 0; s < S; ++s) {
    blks[size_t(s)] = gap_block<T>(c, s, X, sparse_values<T>(c));
    blks[size_t(s)].D = D;
  }
  if constexpr (std::is_same<T, float>::value)
    run_gap_mixed_kernel(blks.data(), S, c.mixed_sparse, SplitPlanes{c.p_hi.get(), c.p_lo.get()}, ec,
                         part, n, nullptr, c.stream);
  else
    run_gap_f64_kernel(blks.data(), S, ec, part, n, nullptr, c.stream);
  launch_fill(energy, c.n_struct, 0.0, c.stream);
  launch_segment_sum(ec, 1, c.d_atom_off.get(), c.n_struct, energy, c.stream);
}

template <class T>
void gap_knm_t(Context& c, int target, double* knm) {
  const int S = int(c.species.size());
  const int64_t n = c.n_atoms;
  const int64_t rows = target == 0 ? c.n_struct : n;
  // rows per centre; the Structure rows are their sums over each
  // structure's centres, in centre order (no atomics: the same on every run)
  double* ka = c.tmp64.reserve(size_t(std::max<int64_t>(n * c.M, 1)));
  double* k = target == 0 ? c.out64.reserve(size_t(std::max<int64_t>(rows * c.M, 1))) : ka;
  RGPU_HIP(hipMemsetAsync(ka, 0, sizeof(double) * n * c.M, c.stream));
  const T* X = invariants_of<T>(c);
  for (int s = 0; s < S; ++s) {
    GapBlock<T> blk = gap_block<T>(c, s, X, sparse_values<T>(c));
    blk.alpha = nullptr;
    if (blk.n_rows == 0 || blk.n_cols == 0) continue;
    launch_gap_kernel<T>(blk, c.M, nullptr, nullptr, ka, c.stream);
  }
  if (target == 0) launch_rows_segment_sum(ka, c.M, c.d_atom_off.get(), c.n_struct, k, c.stream);
  // columns back to the order the sparse points were given; Atom rows of
  // the centres only
  std::vector<double> h(static_cast<size_t>(rows * c.M));
  RGPU_HIP(hipMemcpyAsync(h.data(), k, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
  RGPU_HIP(hipStreamSynchronize(c.stream));
  const int64_t out_rows = target == 0 ? rows : c.n_centres;
  std::vector<double> o(static_cast<size_t>(out_rows * c.M));
  for (int64_t r = 0; r < out_rows; ++r) {
    const int64_t src = target == 0 ? r : c.centre_idx[size_t(r)];
    for (int64_t m = 0; m < c.M; ++m) o[size_t(r * c.M + c.sp_order[m])] = h[size_t(src * c.M + m)];
  }
  if (!o.empty()) RGPU_HIP(hipMemcpy(knm, o.data(), sizeof(double) * o.size(), hipMemcpyDefault));
}

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return RGPU_OK;
  } catch (const InvalidArgument& e) {
    return fail(RGPU_ERR_INVALID, e.what());
  } catch (const RuntimeError& e) {
    return fail(RGPU_ERR_RUNTIME, e.what());
  } catch (const HipError& e) {
    return fail(RGPU_ERR_HIP, e.what());
  } catch (const std::exception& e) {
    return fail(RGPU_ERR_RUNTIME, e.what());
  }
}

}  // namespace
}  // namespace rgpu

using namespace rgpu;

// Autotuner: times the candidate work partitions of the expansion, the
// power spectrum and the adjoint on the current batch and precision (HIP
// events, best of 3 after a warm-up launch) and keeps the fastest.
template <class T>
void autotune_t(Context& c) {
  const int S = int(c.species.size());
  const int64_t n = c.n_atoms;
  if (n == 0) return;
  const bool ps = c.hp.kind == SoapHypers::kPowerSpectrum;
  compute_rep<T>(c);  // buffers and pair cutoffs in place
  hipEvent_t a = c.event(), b = c.event();
  auto time_ms = [&](auto&& launch) {
    launch();
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      RGPU_HIP(hipEventRecord(a, c.stream));
      launch();
      RGPU_HIP(hipEventRecord(b, c.stream));
      RGPU_HIP(hipEventSynchronize(b));
      float ms = 0;
      RGPU_HIP(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    return best;
  };
  SoapTiles& t = c.tile();
  const PairView pv = pair_view(c);
  T* cc = cbuf<T>(c).get();
  unsigned long long* km = c.keymask.get();
  {
    float best = 1e30f;
    int pick = 0;
    for (int pb : {8, 16, 32, 48, 64}) {
      SoapTiles cand = t;
      cand.pb_fwd = pb;
      cand.fwd_group = 0;  // pairs per batch of the per-entry expansion
      cand.fwd_lds = 0;
      float ms;
      try {
        ms = time_ms([&] { launch_expansion<T>(c.rd, S, pv, cc, km, c.stream, cand); });
      } catch (const InvalidArgument&) {
        continue;  // needs more LDS than the device provides
      }
      if (ms < best) best = ms, pick = pb;
    }
    t.pb_fwd = pick;
    // the pair-group expansion against the per-entry one, then the
    // expansion writing c once from LDS
    {
      SoapTiles reg = t, lds = t, ent = t;
      reg.fwd_lds = 0;
      reg.fwd_group = 1;
      ent.fwd_lds = 0;
      ent.fwd_group = 0;
      lds.fwd_lds = 1;
      float ms_reg = time_ms([&] { launch_expansion<T>(c.rd, S, pv, cc, km, c.stream, reg); });
      const float ms_ent = time_ms([&] { launch_expansion<T>(c.rd, S, pv, cc, km, c.stream, ent); });
      if (ms_ent < ms_reg) reg = ent, ms_reg = ms_ent;
      t = reg;
      float ms_lds = 1e30f;
      for (int pb : {16, 32}) {
        SoapTiles cand = lds;
        cand.pb_fwd = pb;
        try {
          const float ms = time_ms([&] {
            launch_expansion_ps<T>(c.rd, S, pv, cc, km, nullptr, nullptr, nullptr, nullptr, c.stream,
                                   cand);
          });
          if (ms < ms_lds) ms_lds = ms, lds = cand;
        } catch (const InvalidArgument&) {
        }
      }
      if (ms_lds < ms_reg) t = lds;
    }
    // the expansion over the half neighbour list, where it applies
    bool half_ok = true;
    try {
      check_half_list(c);
    } catch (const RuntimeError&) {
      half_ok = false;
    }
    if (half_ok) {
      const double* pos = c.pl.pos.get();
      const size_t cp = size_t(std::max<int64_t>(c.pl.cap_pairs, 1)), B = size_t(c.rd.N) * c.rd.LL;
      HalfList hl{};
      hl.count = c.half_count.reserve(size_t(n));
      hl.off = c.half_off.reserve(size_t(n + 1));
      hl.src = c.half_src.reserve(cp);
      hl.j = c.half_j.reserve(cp);
      hl.vec = c.half_vec.reserve(3 * cp);
      hl.dist = c.half_dist.reserve(cp);
      hl.fc = c.half_fc.reserve(cp);
      hl.dfc = c.half_dfc.reserve(cp);
      if constexpr (std::is_same<T, float>::value)
        hl.rows = c.half_rows32.reserve(cp * B);
      else
        hl.rows = c.half_rows64.reserve(cp * B);
      reserve_half_list_scan(n, c.half_scan);
      hl.scan_scratch = &c.half_scan;
      SoapTiles full = t, half = t;
      full.half_nl = 0;
      half.half_nl = 1;
      const float ms_full =
          time_ms([&] { launch_expansion<T>(c.rd, S, pv, cc, km, c.stream, full); });
      const float ms_half = time_ms(
          [&] { launch_expansion_half<T>(c.rd, S, pv, pos, hl, cc, km, c.stream, half); });
      t.half_nl = ms_half < ms_full ? 1 : 0;
    } else {
      t.half_nl = 0;
    }
  }
  if (ps) {
    const int K = S * (S + 1) / 2;
    constexpr bool split = std::is_same<T, float>::value;
    T* pp = split ? nullptr : pbuf<T>(c).get();
    unsigned short* ph = c.p_hi.get();
    unsigned short* pl = c.p_lo.get();
    double* n2 = c.norm2.get();
    float best = 1e30f;
    int pick = 0;
    for (int kc = 1; kc <= std::min(K, 12); ++kc) {
      SoapTiles cand = t;
      cand.kc = kc;
      float ms;
      try {
        ms = time_ms([&] {
          if constexpr (split)
            launch_power_spectrum_split(c.rd, S, n, cc, ph, pl, n2, c.stream, cand);
          else
            launch_power_spectrum<T>(c.rd, S, n, cc, pp, n2, c.stream, cand);
        });
      } catch (const InvalidArgument&) {
        continue;
      }
      if (ms < best) best = ms, pick = kc;
    }
    t.kc = pick;
    // the fused expansion + invariants launch against the two launches
    {
      SoapTiles sep = t, fu = t;
      sep.fused = 0;
      fu.fused = 1;
      const float ms_sep = time_ms([&] {
        launch_expansion<T>(c.rd, S, pv, cc, km, c.stream, sep);
        if constexpr (split)
          launch_power_spectrum_split(c.rd, S, n, cc, ph, pl, n2, c.stream, sep);
        else
          launch_power_spectrum<T>(c.rd, S, n, cc, pp, n2, c.stream, sep);
      });
      float ms_fu = 1e30f;
      try {
        ms_fu = time_ms([&] {
          launch_expansion_ps<T>(c.rd, S, pv, cc, km, pp, ph, pl, n2, c.stream, fu);
        });
      } catch (const InvalidArgument&) {
      }
      t.fused = ms_fu < ms_sep ? 1 : 0;
    }
    // adjoint on a zero dL/dX (its time does not depend on the values)
    const int64_t D = std::max(c.D(), c.Dp());
    DeviceBuffer<T> G;
    DeviceBuffer<double> cv;
    T* g = G.reserve(size_t(n * D));
    RGPU_HIP(hipMemsetAsync(g, 0, sizeof(T) * n * D, c.stream));
    double* cvir = cv.reserve(size_t(6 * n));
    ReverseLists rl;
    const PairView bpv = backward_view(c, &rl);
    float* work = split ? c.adj_work.reserve(size_t(std::max<int64_t>(n * S * c.rd.N * c.rd.LL, 1)))
                        : nullptr;
    auto time_adjoint = [&](const SoapTiles& cand) {
      try {
        return time_ms([&] {
          if constexpr (split)
            launch_backward_split(c.rd, S, bpv, cc, ph, pl, n2, c.hp.normalize, nullptr, 1.0, g,
                                  cvir, c.stream, cand, cand.adj_kind == 3 ? work : nullptr);
          else
            launch_backward<T>(c.rd, S, bpv, cc, pp, n2, c.hp.normalize, nullptr, 1.0, g, cvir,
                               c.stream, cand);
        });
      } catch (const InvalidArgument&) {
        return 1e30f;  // needs more LDS than the device provides
      }
    };
    // key blocks per pass of the FP32 adjoint (the built-in choice fits an
    // LDS budget), then pairs per batch
    if (split) {
      // the staged kernel (adj_kind 0) with its key chunks
      SoapTiles bestv = t;
      bestv.adj_kind = 0;
      bestv.pb_adj = 0;
      bestv.keys_adj = 0;
      best = time_adjoint(bestv);
      int last = -1;
      for (int div = 1; div <= 4; ++div) {
        const int kch = (K + div - 1) / div;
        if (kch == last) continue;
        last = kch;
        SoapTiles cand = bestv;
        cand.keys_adj = kch;
        const float ms = time_adjoint(cand);
        if (ms < best) best = ms, bestv = cand;
      }
      // the matrix-core kernel (adj_kind 1) with its own key chunks
      for (int div = 1; div <= 3; ++div) {
        SoapTiles cand = t;
        cand.adj_kind = 1;
        cand.pb_adj = 0;
        cand.keys_adj = div == 1 ? 0 : (K + div - 1) / div;
        const float ms = time_adjoint(cand);
        if (ms < best) best = ms, bestv = cand;
      }
      // the register-tile kernels (adj_kind 2, and 3 with the pair phase
      // as a kernel of its own): their chunks of key rounds
      for (int kind : {2, 3})
        for (int kch : {0, S + (S + 1) / 2, S}) {
          SoapTiles cand = t;
          cand.adj_kind = kind;
          cand.pb_adj = 0;
          cand.keys_adj = kch;
          const float ms = time_adjoint(cand);
          if (ms < best) best = ms, bestv = cand;
        }
      t.keys_adj = bestv.keys_adj;
      t.adj_kind = bestv.adj_kind;
    }
    // pairs per batch of the pair phase in the adjoint's workgroup (0: the
    // most that fit; adj_kind 3 has a kernel of its own for it)
    best = 1e30f;
    pick = 0;
    if (t.adj_kind != 3)
      for (int pb : {0, 2, 4, 6, 8, 12, 16, 24, 32}) {
        SoapTiles cand = t;
        cand.pb_adj = pb;
        const float ms = time_adjoint(cand);
        if (ms < best) best = ms, pick = pb;
      }
    t.pb_adj = pick;
    RGPU_HIP(hipStreamSynchronize(c.stream));
  }
  c.ev_pool.push_back(a);
  c.ev_pool.push_back(b);
  c.have_rep = false;
  c.drop_graph();
}

struct rgpu_context {
  Context c;
};

extern "C" {

const char* rgpu_last_error(void) { return g_last_error.c_str(); }

int rgpu_context_create(const char* hypers_json, rgpu_context** out) {
  return guarded([&] {
    if (!hypers_json || !out) throw InvalidArgument("null argument");
    *out = nullptr;
    // hypers are validated before any device call, so bad JSON is reported
    // as an argument error on any host
    SoapHypers hp = parse_soap_hypers(hypers_json);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw HipError("no HIP device available: the MI355X library needs a GPU");
    std::unique_ptr<rgpu_context> ctx(new rgpu_context());
    ctx->c.hp = hp;
    build_tables(ctx->c);
    RGPU_HIP(hipDeviceSynchronize());
    *out = ctx.release();
  });
}

void rgpu_context_destroy(rgpu_context* ctx) { delete ctx; }

int rgpu_set_stream(rgpu_context* ctx, void* s) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    // a deferred build is checked on the stream it was enqueued on
    if (ctx->c.have_batch) settle_nl(ctx->c);
    ctx->c.stream = static_cast<hipStream_t>(s);
  });
}

int rgpu_autotune(rgpu_context* ctx) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch(c);
    if (c.precision == RGPU_FP64)
      autotune_t<double>(c);
    else
      autotune_t<float>(c);
  });
}

int rgpu_set_tiles(rgpu_context* ctx, int pairs_fwd, int keys_ps, int pairs_adj, int keys_adj,
                   int adj_kind, int fused) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    if (pairs_fwd > 64 || pairs_adj > 64) throw InvalidArgument("at most 64 pairs per batch");
    Context& c = ctx->c;
    SoapTiles& t = c.tile();
    t.pb_fwd = std::max(pairs_fwd, 0);
    t.kc = std::max(keys_ps, 0);
    t.pb_adj = std::max(pairs_adj, 0);
    t.keys_adj = std::max(keys_adj, 0);
    if (adj_kind < 0 || adj_kind > 3) throw InvalidArgument("adj_kind must be 0, 1, 2 or 3");
    t.adj_kind = adj_kind;
    if (fused > 3) throw InvalidArgument("fused must be -1, 0, 1, 2 or 3");
    t.fused = fused < 0 ? -1 : (fused == 1 ? 1 : 0);
    t.fwd_lds = fused == 2 ? 1 : 0;
    t.fwd_group = fused == 3 ? 1 : 0;
    c.have_rep = false;
    c.drop_graph();
  });
}

int rgpu_enable_graphs(rgpu_context* ctx, int enable) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    c.use_graphs = enable != 0;
    if (!c.use_graphs) {
      c.drop_graph();
      return;
    }
    if (!c.stream) {  // capture needs a stream other than the null stream
      hipStream_t s = nullptr;
      RGPU_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      c.own_stream = s;
      c.stream = s;
    }
  });
}

int rgpu_set_half_list(rgpu_context* ctx, int enable) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    if (enable && c.have_batch) check_half_list(c);
    c.tiles[0].half_nl = c.tiles[1].half_nl = enable ? 1 : 0;
    c.have_rep = false;
    c.drop_graph();
  });
}

int rgpu_set_precision(rgpu_context* ctx, int precision) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    if (precision != RGPU_FP64 && precision != RGPU_MIXED) throw InvalidArgument("unknown precision");
    Context& c = ctx->c;
    if (precision == RGPU_MIXED && c.sparsified())
      throw InvalidArgument("coefficient_subselection is computed in FP64 only (set RGPU_FP64)");
    if (c.precision != precision) c.have_rep = false;
    c.precision = precision;
    if (c.have_sparse && precision == RGPU_MIXED && c.T32.get() == nullptr)
      throw InvalidArgument("internal: sparse points not prepared for mixed precision");
  });
}

int rgpu_get_info(rgpu_context* ctx, int key, int64_t* value) {
  return guarded([&] {
    if (!ctx || !value) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    switch (key) {
      case RGPU_INFO_MAX_RADIAL: *value = c.hp.max_radial; break;
      case RGPU_INFO_MAX_ANGULAR: *value = c.hp.max_angular; break;
      case RGPU_INFO_N_SPECIES: *value = int64_t(c.species.size()); break;
      case RGPU_INFO_FEATURE_SIZE: *value = c.D(); break;
      case RGPU_INFO_N_CENTRES: *value = c.n_centres; break;
      case RGPU_INFO_N_ATOMS: *value = c.n_atoms; break;
      case RGPU_INFO_N_PAIRS:
        if (c.have_batch) settle_nl(c);
        *value = c.have_batch ? c.pl.n_pairs : 0;
        break;
      case RGPU_INFO_N_STRUCT: *value = c.n_struct; break;
      case RGPU_INFO_SPLINE_GRID: *value = c.rd.use_spline ? c.rd.grid_n : 0; break;
      case RGPU_INFO_N_SPARSE: *value = c.M; break;
      case RGPU_INFO_IS_INVARIANTS: *value = c.hp.kind == SoapHypers::kPowerSpectrum; break;
      case RGPU_INFO_PRECISION: *value = c.precision; break;
      case RGPU_INFO_GRAPH_CAPTURES: *value = c.graph_captures; break;
      case RGPU_INFO_GRAPH_LAUNCHES: *value = c.graph_launches; break;
      case RGPU_INFO_NL_GRAPH_LAUNCHES: *value = c.pl.graph_launches; break;
      case RGPU_INFO_TILE_PAIRS_FWD: *value = c.tile().pb_fwd; break;
      case RGPU_INFO_TILE_KEYS_PS: *value = c.tile().kc; break;
      case RGPU_INFO_TILE_PAIRS_ADJ: *value = c.tile().pb_adj; break;
      case RGPU_INFO_TILE_KEYS_ADJ: *value = c.tile().keys_adj; break;
      case RGPU_INFO_TILE_ADJ_KIND: *value = c.tile().adj_kind; break;
      case RGPU_INFO_HALF_LIST: *value = c.tile().half_nl; break;
      case RGPU_INFO_FUSED: *value = c.tile().fused == 1; break;
      case RGPU_INFO_FWD_LDS: *value = c.tile().fwd_lds; break;
      case RGPU_INFO_FWD_GROUP: *value = c.tile().fwd_group; break;
      default: throw InvalidArgument("unknown info key");
    }
  });
}

int rgpu_set_structures(rgpu_context* ctx, int n_struct, const int64_t* atom_offsets,
                        const double* positions, const int32_t* types, const double* cells,
                        const int32_t* pbc, int n_species, const int32_t* species,
                        const uint8_t* centre_mask) {
  return guarded([&] {
    if (!ctx || !atom_offsets || n_struct < 0 || n_species < 0) throw InvalidArgument("bad arguments");
    Context& c = ctx->c;
    c.have_batch = c.have_rep = false;
    const int64_t n_atoms = atom_offsets[n_struct];
    if (atom_offsets[0] != 0) throw InvalidArgument("atom_offsets[0] must be 0");
    if (n_atoms > 0 && (!positions || !types)) throw InvalidArgument("null positions/types");
    if (n_struct > 0 && (!cells || !pbc)) throw InvalidArgument("null cells/pbc");
    if (n_species > kMaxSpecies) throw InvalidArgument("at most 64 species");
    std::vector<int> sp(species, species + n_species);
    for (int k = 1; k < n_species; ++k)
      if (sp[k] <= sp[k - 1]) throw InvalidArgument("species must be sorted and unique");
    std::vector<int> sof(static_cast<size_t>(n_atoms));
    for (int64_t a = 0; a < n_atoms; ++a) {
      const int t = types[a];
      if (t < 0 || t >= kMaxChemElements)
        throw RuntimeError("atom types (or atomic numbers) should be in the range [0, 130].");
      auto it = std::lower_bound(sp.begin(), sp.end(), t);
      if (it == sp.end() || *it != t) throw InvalidArgument("atom type missing from the species list");
      sof[a] = int(it - sp.begin());
    }
    // centres: the atoms of the mask (every atom without one)
    std::vector<int64_t> cidx;
    for (int64_t a = 0; a < n_atoms; ++a)
      if (!centre_mask || centre_mask[a]) cidx.push_back(a);
    const std::string& method = c.hp.expansion_by_species;
    if (method == "user defined") {
      // calculator_spherical_expansion.hh:2520-2545: every centre species is
      // a global species; the layout is the global species list
      std::vector<int> missing;
      for (int64_t a : cidx) {
        const int t = types[a];
        if (std::find(c.hp.global_species.begin(), c.hp.global_species.end(), t) ==
                c.hp.global_species.end() &&
            std::find(missing.begin(), missing.end(), t) == missing.end())
          missing.push_back(t);
      }
      if (!missing.empty()) {
        std::string m = "global_species is missing at least these species: '";
        for (int t : missing) m += std::to_string(t) + ", ";
        throw RuntimeError(m + "'.");
      }
      for (int z : c.hp.global_species)
        if (!std::binary_search(sp.begin(), sp.end(), z))
          throw InvalidArgument("the species list must contain every global species (" +
                                std::to_string(z) + " is missing)");
    }
    // coefficient_subselection: full-layout column of each selected
    // coefficient over this species list
    std::vector<int64_t> scol;
    if (c.hp.kind == SoapHypers::kPowerSpectrum && c.hp.sparsified) {
      const int N = c.hp.max_radial, L1 = c.hp.max_angular + 1, S = int(sp.size());
      for (size_t k = 0; k < c.hp.sel_a.size(); ++k) {
        auto ia = std::lower_bound(sp.begin(), sp.end(), c.hp.sel_a[k]);
        auto ib = std::lower_bound(sp.begin(), sp.end(), c.hp.sel_b[k]);
        if (ia == sp.end() || *ia != c.hp.sel_a[k] || ib == sp.end() || *ib != c.hp.sel_b[k]) {
          scol.push_back(-1);
          continue;
        }
        const int a = int(ia - sp.begin()), b = int(ib - sp.begin());
        const int64_t key = int64_t(a) * S - int64_t(a) * (a - 1) / 2 + (b - a);
        scol.push_back((key * N * N + int64_t(c.hp.sel_n1[k]) * N + c.hp.sel_n2[k]) * L1 +
                       c.hp.sel_l[k]);
      }
    }
    // RadialDimReduction: the spline table of each batch species (a species
    // without projections is an error, calculator_spherical_expansion.hh
    // :1339-1371)
    std::vector<int> sp_tab;
    if (!c.rh->tab_species.empty()) {
      for (int z : sp) {
        auto it = std::find(c.rh->tab_species.begin(), c.rh->tab_species.end(), z);
        if (it == c.rh->tab_species.end())
          throw RuntimeError("RadialDimReduction is missing projection matrices at least for species '" +
                             std::to_string(z) + "'");
        sp_tab.push_back(int(it - c.rh->tab_species.begin()));
      }
    }
    // sparse points are laid out over the species list they were set with
    // (columns per key, rows grouped by species): any other list drops them
    const bool species_changed = c.species != sp;
    if (species_changed) c.have_sparse = false;
    c.n_struct = n_struct;
    c.n_atoms = n_atoms;
    c.atom_off.assign(atom_offsets, atom_offsets + n_struct + 1);
    c.cells.assign(cells, cells + 9 * n_struct);
    c.pbc.assign(pbc, pbc + 3 * n_struct);
    c.types.assign(types, types + n_atoms);
    c.species = sp;
    c.species_of = sof;
    c.masked = int64_t(cidx.size()) != n_atoms;
    c.n_centres = int64_t(cidx.size());
    c.centre_idx = cidx;
    int64_t* dci = c.d_centre_idx.reserve(std::max<size_t>(cidx.size(), 1));
    if (!cidx.empty())
      RGPU_HIP(hipMemcpyAsync(dci, cidx.data(), sizeof(int64_t) * cidx.size(), hipMemcpyHostToDevice,
                              c.stream));
    c.sel_col = scol;
    if (!scol.empty()) {
      int64_t* dsc = c.d_sel_col.reserve(scol.size());
      RGPU_HIP(hipMemcpyAsync(dsc, scol.data(), sizeof(int64_t) * scol.size(), hipMemcpyHostToDevice,
                              c.stream));
      // inverse: the selections naming each full-layout column, ascending
      const int64_t Df = c.D_full();
      std::vector<int64_t> inv_off(size_t(Df + 1), 0), inv_k;
      for (int64_t k : scol)
        if (k >= 0) ++inv_off[size_t(k + 1)];
      for (int64_t o = 0; o < Df; ++o) inv_off[size_t(o + 1)] += inv_off[size_t(o)];
      inv_k.resize(size_t(std::max<int64_t>(inv_off[size_t(Df)], 1)));
      std::vector<int64_t> fill(inv_off.begin(), inv_off.end() - 1);
      for (size_t k = 0; k < scol.size(); ++k)
        if (scol[k] >= 0) inv_k[size_t(fill[size_t(scol[k])]++)] = int64_t(k);
      int64_t* dio = c.d_sel_inv_off.reserve(inv_off.size());
      int64_t* dik = c.d_sel_inv_k.reserve(inv_k.size());
      RGPU_HIP(hipStreamSynchronize(c.stream));  // no kernel still reads the old map
      RGPU_HIP(hipMemcpy(dio, inv_off.data(), sizeof(int64_t) * inv_off.size(), hipMemcpyHostToDevice));
      RGPU_HIP(hipMemcpy(dik, inv_k.data(), sizeof(int64_t) * inv_k.size(), hipMemcpyHostToDevice));
    }
    c.rd.sp_tab = nullptr;
    if (!sp_tab.empty()) {
      int* dt = c.d_sp_tab.reserve(sp_tab.size());
      RGPU_HIP(hipMemcpyAsync(dt, sp_tab.data(), sizeof(int) * sp_tab.size(), hipMemcpyHostToDevice,
                              c.stream));
      c.rd.sp_tab = dt;
    }
    int* d = c.d_species_of.reserve(std::max<size_t>(sof.size(), 1));
    if (n_atoms) RGPU_HIP(hipMemcpyAsync(d, sof.data(), sizeof(int) * n_atoms, hipMemcpyHostToDevice, c.stream));
    int64_t* dof = c.d_atom_off.reserve(size_t(n_struct + 1));
    RGPU_HIP(hipMemcpyAsync(dof, c.atom_off.data(), sizeof(int64_t) * (n_struct + 1), hipMemcpyHostToDevice,
                            c.stream));
    build_batch_nl(c, positions);
    build_rows(c);
    RGPU_HIP(hipStreamSynchronize(c.stream));
    c.have_batch = true;
  });
}

int rgpu_update_positions(rgpu_context* ctx, const double* positions, const double* cells) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch_unsettled(c);
    if (cells) c.cells.assign(cells, cells + 9 * c.n_struct);
    // the size check of the list waits for the next call that needs it (a
    // prediction checks after its own work, so the step has no extra wait)
    build_batch_nl(c, positions, true);
  });
}

int rgpu_get_neighbours(rgpu_context* ctx, int64_t* centre_offsets, int32_t* pair_j, double* pair_vec) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch(c);
    if (centre_offsets)
      RGPU_HIP(hipMemcpyAsync(centre_offsets, c.pl.centre_off.get(), sizeof(int64_t) * (c.n_atoms + 1),
                              hipMemcpyDefault, c.stream));
    if (pair_j && c.pl.n_pairs)
      RGPU_HIP(hipMemcpyAsync(pair_j, c.pl.pair_j.get(), sizeof(int) * c.pl.n_pairs, hipMemcpyDefault,
                              c.stream));
    if (pair_vec && c.pl.n_pairs)
      RGPU_HIP(hipMemcpyAsync(pair_vec, c.pl.pair_vec.get(), sizeof(double) * 3 * c.pl.n_pairs,
                              hipMemcpyDefault, c.stream));
    RGPU_HIP(hipStreamSynchronize(c.stream));
  });
}

int rgpu_compute(rgpu_context* ctx) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    require_batch(ctx->c);
    compute(ctx->c);
    RGPU_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int rgpu_get_expansion(rgpu_context* ctx, double* coeffs, uint64_t* key_mask) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch(c);
    ensure_rep(c);
    const int64_t E = int64_t(c.species.size()) * c.rd.N * c.rd.LL, nc = c.n_centres;
    const int64_t* rows = c.masked ? c.d_centre_idx.get() : nullptr;
    if (coeffs && nc) {
      if (c.precision == RGPU_FP64 && !rows) {
        RGPU_HIP(hipMemcpyAsync(coeffs, c.c64.get(), sizeof(double) * nc * E, hipMemcpyDefault,
                                c.stream));
      } else {
        double* o = c.out64.reserve(size_t(nc * E));
        if (c.precision == RGPU_FP64)
          launch_normalise<double>(nc, E, c.c64.get(), nullptr, false, o, c.stream, rows);
        else
          launch_normalise<float>(nc, E, c.c32.get(), nullptr, false, o, c.stream, rows);
        RGPU_HIP(hipMemcpyAsync(coeffs, o, sizeof(double) * nc * E, hipMemcpyDefault, c.stream));
      }
    }
    if (key_mask && nc) {
      // per-atom masks of the expansion (neighbour species and the centre's
      // own); the species-key method decides the keys of each centre
      // (calculator_spherical_expansion.hh:2414-2570)
      std::vector<uint64_t> km(static_cast<size_t>(c.n_atoms));
      RGPU_HIP(hipMemcpyAsync(km.data(), c.keymask.get(), sizeof(uint64_t) * km.size(),
                              hipMemcpyDeviceToHost, c.stream));
      RGPU_HIP(hipStreamSynchronize(c.stream));
      const std::string& method = c.hp.expansion_by_species;
      std::vector<uint64_t> out(static_cast<size_t>(nc));
      if (method == "user defined") {
        uint64_t all = 0;
        for (int z : c.hp.global_species) {
          auto it = std::lower_bound(c.species.begin(), c.species.end(), z);
          all |= 1ull << (it - c.species.begin());
        }
        std::fill(out.begin(), out.end(), all);
      } else if (method == "structure wise") {
        // every centre of a structure has the species of all of its centres
        // and their neighbours
        std::vector<uint64_t> per(static_cast<size_t>(c.n_struct), 0);
        std::vector<int> st(static_cast<size_t>(nc));
        for (int64_t r = 0, s = 0; r < nc; ++r) {
          const int64_t a = c.centre_idx[size_t(r)];
          while (c.atom_off[size_t(s) + 1] <= a) ++s;
          st[size_t(r)] = int(s);
          per[size_t(s)] |= km[size_t(a)];
        }
        for (int64_t r = 0; r < nc; ++r) out[size_t(r)] = per[size_t(st[size_t(r)])];
      } else {
        for (int64_t r = 0; r < nc; ++r) out[size_t(r)] = km[size_t(c.centre_idx[size_t(r)])];
      }
      RGPU_HIP(hipMemcpy(key_mask, out.data(), sizeof(uint64_t) * nc, hipMemcpyDefault));
    }
    RGPU_HIP(hipStreamSynchronize(c.stream));
  });
}

int rgpu_get_features(rgpu_context* ctx, double* features, double* inv_norms) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch(c);
    require_invariants(c);
    ensure_rep(c);
    const int64_t D = c.D(), nc = c.n_centres;
    const int64_t* rows = c.masked ? c.d_centre_idx.get() : nullptr;
    if (features && nc) {
      double* o = c.out64.reserve(size_t(nc * D));
      if (c.precision == RGPU_FP64)
        launch_normalise<double>(nc, D, invariants_of<double>(c), c.norm2.get(), c.hp.normalize, o,
                                 c.stream, rows);
      else
        launch_normalise<float>(nc, D, ensure_p32(c), c.norm2.get(), c.hp.normalize, o, c.stream,
                                rows);
      RGPU_HIP(hipMemcpyAsync(features, o, sizeof(double) * nc * D, hipMemcpyDefault, c.stream));
    }
    if (inv_norms && nc) {
      std::vector<double> h(static_cast<size_t>(c.n_atoms)), o(static_cast<size_t>(nc));
      RGPU_HIP(hipMemcpyAsync(h.data(), c.norm2.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                              c.stream));
      RGPU_HIP(hipStreamSynchronize(c.stream));
      for (int64_t r = 0; r < nc; ++r) {
        const double v = h[size_t(c.centre_idx[size_t(r)])];
        o[size_t(r)] = (c.hp.normalize && v > 0) ? 1.0 / std::sqrt(v) : 1.0;
      }
      RGPU_HIP(hipMemcpyAsync(inv_norms, o.data(), sizeof(double) * nc, hipMemcpyDefault, c.stream));
    }
    RGPU_HIP(hipStreamSynchronize(c.stream));
  });
}

int rgpu_get_features_gradient(rgpu_context* ctx, double* out) {
  return guarded([&] {
    if (!ctx || !out) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    require_batch(c);
    require_invariants(c);
    require_unmasked(c);
    if (c.precision != RGPU_FP64)
      throw InvalidArgument("the feature gradients are computed in FP64 (set RGPU_FP64)");
    ensure_rep(c);
    const int S = int(c.species.size());
    const int64_t n = c.n_atoms, np = c.pl.n_pairs, D = c.D();
    const int64_t B = int64_t(c.rd.N) * c.rd.LL;
    if (np == 0 || D == 0) return;
    double* dcp = c.dcp.reserve(size_t(np * 3 * B));
    double* dcs = c.dcs.reserve(size_t(std::max<int64_t>(n * 3 * S * B, 1)));
    launch_pair_gradients(c.rd, S, pair_view(c), dcp, dcs, c.stream);
    double* o = is_device_memory(out) ? out : c.tmp64.reserve(size_t(3 * np * D));
    launch_feature_gradient_rows(c.rd.N, c.rd.L1, S, pair_view(c), np, c.c64.get(),
                                 invariants_of<double>(c), c.norm2.get(), c.hp.normalize, dcp, dcs,
                                 D, c.sparsified() ? c.d_sel_col.get() : nullptr, o, c.stream);
    if (o != out)
      RGPU_HIP(hipMemcpyAsync(out, o, sizeof(double) * 3 * np * D, hipMemcpyDeviceToHost, c.stream));
    RGPU_HIP(hipStreamSynchronize(c.stream));
  });
}

int rgpu_features_backward(rgpu_context* ctx, const double* dL_dX, double* dL_dr, double* neg_stress) {
  return guarded([&] {
    if (!ctx || !dL_dX) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    require_batch(c);
    require_invariants(c);
    require_unmasked(c);
    ensure_rep(c);
    const int64_t D = c.D(), n = c.n_atoms;
    double* gd = c.grad.reserve(size_t(std::max<int64_t>(3 * n, 1)));
    double* cv = c.cvir.reserve(size_t(std::max<int64_t>(6 * n, 1)));
    double* g64 = c.tmp64.reserve(size_t(std::max<int64_t>(n * D, 1)));
    RGPU_HIP(hipMemcpyAsync(g64, dL_dX, sizeof(double) * n * D, hipMemcpyDefault, c.stream));
    if (c.precision == RGPU_FP64) {
      backward_device<double>(c, g64, gd, cv);
    } else {
      // dL/dX in the packed layout of the mixed path
      float* g32 = c.G32.reserve(size_t(std::max<int64_t>(n * c.Dp(), 1)));
      launch_pack_gradient(int(c.species.size()), c.rd.N, c.rd.L1, n, g64, g32, c.stream);
      backward_device<float>(c, g32, gd, cv);
    }
    emit_outputs(c, nullptr, nullptr, gd, cv, dL_dr, neg_stress);
  });
}

int rgpu_set_sparse_points(rgpu_context* ctx, int zeta, int64_t n_sparse, const int32_t* species,
                           const double* values, const double* weights) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch(c);
    require_invariants(c);
    if (zeta < 1) throw InvalidArgument("zeta should be specified for the GAP kernel (>= 1)");
    if (n_sparse < 0 || (n_sparse > 0 && (!species || !values))) throw InvalidArgument("bad sparse points");
    const int S = int(c.species.size());
    const int64_t D = c.D();
    // species-sorted (stable) order
    std::vector<int64_t> order(static_cast<size_t>(n_sparse));
    for (int64_t m = 0; m < n_sparse; ++m) order[m] = m;
    std::vector<int> sidx(static_cast<size_t>(n_sparse));
    for (int64_t m = 0; m < n_sparse; ++m) {
      auto it = std::lower_bound(c.species.begin(), c.species.end(), species[m]);
      sidx[m] = (it == c.species.end() || *it != species[m]) ? S : int(it - c.species.begin());
    }
    // species-sorted: batch species first (in the context's order), then any
    // other species grouped by value (they contribute no K_NM columns)
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      if (sidx[a] != sidx[b]) return sidx[a] < sidx[b];
      return sidx[a] == S && species[a] < species[b];
    });
    c.sp_order = order;
    c.sp_species.resize(size_t(n_sparse));
    c.sp_off.assign(S + 1, 0);
    for (int64_t m = 0; m < n_sparse; ++m) {
      c.sp_species[m] = species[order[m]];
      if (sidx[order[m]] < S) c.sp_off[sidx[order[m]] + 1]++;
    }
    for (int s = 0; s < S; ++s) c.sp_off[s + 1] += c.sp_off[s];
    std::vector<double> T(size_t(n_sparse * D)), a(size_t(std::max<int64_t>(n_sparse, 1)), 0.0);
    for (int64_t m = 0; m < n_sparse; ++m) {
      std::copy(values + order[m] * D, values + (order[m] + 1) * D, T.begin() + m * D);
      if (weights) a[m] = weights[order[m]];
    }
    double* dT = c.T64.reserve(std::max<size_t>(T.size(), 1));
    if (!T.empty()) RGPU_HIP(hipMemcpy(dT, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    std::vector<float> Tf(T.begin(), T.end());
    float* dTf = c.T32.reserve(std::max<size_t>(Tf.size(), 1));
    if (!Tf.empty()) RGPU_HIP(hipMemcpy(dTf, Tf.data(), sizeof(float) * Tf.size(), hipMemcpyHostToDevice));
    double* da = c.alpha.reserve(a.size());
    RGPU_HIP(hipMemcpy(da, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice));
    c.zeta = zeta;
    c.M = n_sparse;
    c.have_weights = weights != nullptr;
    // (the mixed path's packed operands; a sparsified layout is FP64 only)
    if (!c.sparsified()) prepare_mixed_sparse(c.mixed_sparse, T, n_sparse, S, c.rd.N, c.rd.L1, c.stream);
    c.have_sparse = true;
  });
}

int rgpu_gram_f64(int64_t rows, int64_t cols, const double* A, double* out) {
  return guarded([&] {
    if (rows < 0 || cols < 0) throw InvalidArgument("negative matrix size");
    if ((rows > 0 && cols > 0 && !A) || (cols > 0 && !out)) throw InvalidArgument("null argument");
    if ((rows + 31) / 32 > 65535 || (cols + 31) / 32 > 65535)
      throw InvalidArgument("matrix too large for rgpu_gram_f64 (at most 2097120 rows and columns)");
    const size_t nc = size_t(cols) * size_t(cols);
    if (rows == 0 || cols == 0) {
      if (cols > 0) std::fill(out, out + nc, 0.0);
      return;
    }
    DeviceBuffer<double> dA, dAt, dC;
    DeviceBuffer<int> dI;
    std::vector<int> iota(static_cast<size_t>(cols));
    for (int64_t k = 0; k < cols; ++k) iota[size_t(k)] = int(k);
    const size_t na = size_t(rows) * size_t(cols);
    double* a = dA.reserve(na);
    double* at = dAt.reserve(na);
    double* cm = dC.reserve(nc);
    int* io = dI.reserve(iota.size());
    RGPU_HIP(hipMemcpy(a, A, sizeof(double) * na, hipMemcpyHostToDevice));
    RGPU_HIP(hipMemcpy(io, iota.data(), sizeof(int) * iota.size(), hipMemcpyHostToDevice));
    run_gram_f64(rows, cols, a, at, io, cm, nullptr);
    RGPU_HIP(hipMemcpy(out, cm, sizeof(double) * nc, hipMemcpyDeviceToHost));
  });
}

int rgpu_gap_kmm(rgpu_context* ctx, double* kmm) {
  return guarded([&] {
    if (!ctx || !kmm) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    if (!c.have_sparse) throw InvalidArgument(kNoSparse);
    const int64_t M = c.M, D = c.D();
    std::vector<int> iota(static_cast<size_t>(std::max<int64_t>(M, 1)));
    for (int64_t m = 0; m < M; ++m) iota[size_t(m)] = int(m);
    int* rows = c.d_sp_rows.reserve(iota.size());
    RGPU_HIP(hipMemcpyAsync(rows, iota.data(), sizeof(int) * iota.size(), hipMemcpyHostToDevice, c.stream));
    double* k = c.tmp64.reserve(size_t(std::max<int64_t>(M * M, 1)));
    RGPU_HIP(hipMemsetAsync(k, 0, sizeof(double) * M * M, c.stream));
    // one block per run of equal species (sparse points are species-sorted)
    for (int64_t m0 = 0; m0 < M;) {
      int64_t m1 = m0 + 1;
      while (m1 < M && c.sp_species[size_t(m1)] == c.sp_species[size_t(m0)]) ++m1;
      GapBlock<double> b;
      b.n_rows = b.n_cols = m1 - m0;
      b.row_idx = rows + m0;
      b.col0 = m0;
      b.X = c.T64.get();
      b.norm2 = nullptr;
      b.normalize = 0;
      b.Tm = c.T64.get();
      b.alpha = nullptr;
      b.zeta = c.zeta;
      b.D = D;
      launch_gap_kernel<double>(b, M, nullptr, nullptr, k, c.stream);
      m0 = m1;
    }
    std::vector<double> h(static_cast<size_t>(M * M)), o(h.size());
    RGPU_HIP(hipMemcpyAsync(h.data(), k, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    RGPU_HIP(hipStreamSynchronize(c.stream));
    for (int64_t r = 0; r < M; ++r)
      for (int64_t m = 0; m < M; ++m)
        o[size_t(c.sp_order[r] * M + c.sp_order[m])] = h[size_t(r * M + m)];
    RGPU_HIP(hipMemcpy(kmm, o.data(), sizeof(double) * o.size(), hipMemcpyDefault));
  });
}

int rgpu_gap_knm_gradient(rgpu_context* ctx, int neg_stress, double* out) {
  return guarded([&] {
    if (!ctx || !out) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    require_batch(c);
    require_invariants(c);
    if (!c.have_sparse) throw InvalidArgument(kNoSparse);
    if (c.precision != RGPU_FP64)
      throw InvalidArgument("the K_NM gradient rows are computed in FP64 (set RGPU_FP64)");
    require_unmasked(c);
    ensure_rep(c);
    const int S = int(c.species.size());
    const int64_t n = c.n_atoms, M = c.M, D = c.D(), ns = c.n_struct;
    const int64_t rows = 3 * n + (neg_stress ? 6 * ns : 0);
    double* o = c.knm_grad.reserve(size_t(std::max<int64_t>(rows * M, 1)));
    if (rows * M == 0) return;
    // d[i][m] = X_i . T_m for the centre's species (the K_NM of zeta 1)
    double* d = c.tmp64.reserve(size_t(std::max<int64_t>(n * M, 1)));
    RGPU_HIP(hipMemsetAsync(d, 0, sizeof(double) * n * M, c.stream));
    for (int s = 0; s < S; ++s) {
      GapBlock<double> blk = gap_block<double>(c, s, invariants_of<double>(c), c.T64.get());
      blk.alpha = nullptr;
      blk.zeta = 1;
      if (blk.n_rows == 0 || blk.n_cols == 0) continue;
      launch_gap_kernel<double>(blk, M, nullptr, nullptr, d, c.stream);
    }
    double* gs = c.knm_scale.reserve(size_t(std::max<int64_t>(n, 1)));
    double* gd = c.grad.reserve(size_t(std::max<int64_t>(3 * n, 1)));
    double* cv = c.cvir.reserve(size_t(std::max<int64_t>(6 * n, 1)));
    double* st = c.out64.reserve(size_t(6 * std::max<int64_t>(ns, 1)));
    // a sparsified row enters the adjoint scattered into the full layout
    double* grow = c.sparsified() ? c.gfull.reserve(size_t(c.D_full())) : nullptr;
    ReverseLists rl;
    const PairView bpv = backward_view(c, &rl);
    // column m (species-sorted): dL/dX_i = zeta d_im^(zeta-1) T_m for the
    // centres of m's species, through the adjoint; rows written in the
    // caller's sparse-point order
    for (int s = 0; s < S; ++s)
      for (int64_t m = c.sp_off[s]; m < c.sp_off[s + 1]; ++m) {
        launch_knm_gradient_scale(n, d + m, M, c.d_species_of.get(), s, c.zeta, gs, c.stream);
        const double* row = c.T64.get() + m * D;
        if (grow) {
          launch_scatter_columns(1, D, c.selection(), row, c.D_full(), grow, c.stream);
          row = grow;
        }
        launch_backward_rank1(c.rd, S, bpv, c.c64.get(), c.p64.get(), c.norm2.get(),
                              c.hp.normalize, row, gs, cv, c.stream, c.tile());
        launch_force_gather(bpv, rl, gd, c.stream);
        const int64_t col = c.sp_order[size_t(m)];
        launch_strided_copy(gd, 3 * n, o + col, M, c.stream);
        if (neg_stress) {
          RGPU_HIP(hipMemsetAsync(st, 0, sizeof(double) * 6 * ns, c.stream));
          launch_segment_sum(cv, 6, c.d_atom_off.get(), int(ns), st, c.stream, c.d_inv_vol.get());
          launch_strided_copy(st, 6 * ns, o + 3 * n * M + col, M, c.stream);
        }
      }
    RGPU_HIP(hipMemcpyAsync(out, o, sizeof(double) * rows * M, hipMemcpyDefault, c.stream));
    RGPU_HIP(hipStreamSynchronize(c.stream));
  });
}

int rgpu_gap_knm(rgpu_context* ctx, int target, double* knm) {
  return guarded([&] {
    if (!ctx || !knm) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    require_batch(c);
    if (!c.have_sparse) throw InvalidArgument(kNoSparse);
    if (target != 0 && target != 1) throw InvalidArgument("target must be 0 (Structure) or 1 (Atom)");
    ensure_rep(c);
    if (c.precision == RGPU_FP64)
      gap_knm_t<double>(c, target, knm);
    else
      gap_knm_t<float>(c, target, knm);
  });
}

int rgpu_gap_predict(rgpu_context* ctx, double* energy, double* gradient, double* neg_stress) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    Context& c = ctx->c;
    require_batch_unsettled(c);
    if (!c.have_sparse) throw InvalidArgument(kNoSparse);
    if (!c.have_weights) throw InvalidArgument("sparse points were set without weights");
    // runs on an unchecked neighbour list, then checks it: a list whose
    // capacity was short held no pairs, so it is built again and the
    // prediction repeated
    for (int attempt = 0; attempt < 2; ++attempt) {
      // the reverse pair lists of the forces are built beside the
      // representation stage
      if ((gradient || neg_stress) && !c.masked) prefetch_reverse_lists(c);
      ensure_rep(c);
      if (c.precision == RGPU_FP64)
        gap_predict_t<double>(c, energy, gradient, neg_stress);
      else
        gap_predict_t<float>(c, energy, gradient, neg_stress);
      if (settle_nl(c)) break;
    }
  });
}

int rgpu_gap_energy_noalloc(rgpu_context* ctx, double* energy) {
  return guarded([&] {
    if (!ctx || !energy) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    require_batch_unsettled(c);
    if (!c.have_sparse) throw InvalidArgument(kNoSparse);
    if (!c.have_weights) throw InvalidArgument("sparse points were set without weights");
    // scratch: the per-centre energies, then the energy GEMM's partial sums
    // (one per centre and column tile of the widest species block)
    const int64_t n = c.n_atoms, lead = (n + 31) / 32 * 32;
    const int tiles = c.precision == RGPU_FP64 ? energy_tiles<double>(c) : energy_tiles<float>(c);
    if (lead + n * tiles > kScratchDoubles)
      throw InvalidArgument("rgpu_gap_energy_noalloc: the batch has " + std::to_string(n) + " centres x (1 + " +
                            std::to_string(tiles) + " column tiles) = " + std::to_string(lead + n * tiles) +
                            " doubles of scratch, more than the " + std::to_string(kScratchDoubles) +
                            " it holds (use rgpu_gap_predict)");
    if (c.n_struct > 0 && !is_device_memory(energy))
      throw InvalidArgument("rgpu_gap_energy_noalloc writes device memory: energy must be a "
                            "device pointer");
    if (c.timing)
      throw InvalidArgument("stage timing creates events: disable it for rgpu_gap_energy_noalloc");
    double* ec = scratch_on_current_device();
    NoAllocScope no_alloc;
    // a deferred size check that finds the list too small needs a rebuild,
    // which allocates: leave it to the next allocating call
    if (c.nl_stale || !finish_pair_list(c.pl, c.stream)) {
      c.nl_stale = true;
      c.have_rep = false;
      throw RuntimeError(
          "the neighbour list outgrew its buffers: run rgpu_gap_predict (it rebuilds the list) "
          "before rgpu_gap_energy_noalloc");
    }
    ensure_rep(c);
    if (c.precision == RGPU_FP64)
      gap_energy_noalloc_t<double>(c, ec, ec + lead, energy);
    else
      gap_energy_noalloc_t<float>(c, ec, ec + lead, energy);
  });
}

int rgpu_enable_timing(rgpu_context* ctx, int enable) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    ctx->c.timing = enable != 0;
  });
}

int rgpu_get_timing(rgpu_context* ctx, int stage, double* total_ms, int64_t* launches) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    if (stage < 0 || stage >= RGPU_N_STAGES) throw InvalidArgument("unknown stage");
    ctx->c.collect();
    if (total_ms) *total_ms = ctx->c.stage_ms[stage];
    if (launches) *launches = ctx->c.stage_n[stage];
  });
}

int rgpu_reset_timing(rgpu_context* ctx) {
  return guarded([&] {
    if (!ctx) throw InvalidArgument("null context");
    ctx->c.collect();
    for (int k = 0; k < RGPU_N_STAGES; ++k) {
      ctx->c.stage_ms[k] = 0;
      ctx->c.stage_n[k] = 0;
    }
  });
}

int rgpu_eval_radial(rgpu_context* ctx, int64_t n, const double* r, double* R, double* dR) {
  return guarded([&] {
    if (!ctx || !r || !R) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    const int64_t F = int64_t(c.rd.N) * c.rd.L1;
    DeviceBuffer<double> dr, dv, dd;
    double* pr = dr.reserve(size_t(std::max<int64_t>(n, 1)));
    RGPU_HIP(hipMemcpy(pr, r, sizeof(double) * n, hipMemcpyDefault));
    double* pv = dv.reserve(size_t(std::max<int64_t>(n * F, 1)));
    double* pd = dR ? dd.reserve(size_t(std::max<int64_t>(n * F, 1))) : nullptr;
    if (c.rd.use_spline) {
      if (c.precision == RGPU_FP64)
        launch_radial_spline_eval<double>(c.rd, pr, n, pv, pd, c.stream);
      else
        launch_radial_spline_eval<float>(c.rd, pr, n, pv, pd, c.stream);
    } else {
      launch_radial_direct(c.rd, pr, n, pv, pd, c.stream);
    }
    RGPU_HIP(hipStreamSynchronize(c.stream));
    RGPU_HIP(hipMemcpy(R, pv, sizeof(double) * n * F, hipMemcpyDefault));
    if (dR) RGPU_HIP(hipMemcpy(dR, pd, sizeof(double) * n * F, hipMemcpyDefault));
  });
}

int rgpu_eval_ylm(rgpu_context* ctx, int64_t n, const double* dirs, double* Y, double* dY) {
  return guarded([&] {
    if (!ctx || !dirs || !Y) throw InvalidArgument("null argument");
    Context& c = ctx->c;
    const int64_t LL = c.rd.LL;
    DeviceBuffer<double> dd, dy, ddy;
    double* pd = dd.reserve(size_t(std::max<int64_t>(3 * n, 1)));
    RGPU_HIP(hipMemcpy(pd, dirs, sizeof(double) * 3 * n, hipMemcpyDefault));
    double* py = dy.reserve(size_t(std::max<int64_t>(n * LL, 1)));
    double* pdy = dY ? ddy.reserve(size_t(std::max<int64_t>(3 * n * LL, 1))) : nullptr;
    if (c.precision == RGPU_FP64)
      launch_ylm_eval<double>(c.rd, pd, n, py, pdy, c.stream);
    else
      launch_ylm_eval<float>(c.rd, pd, n, py, pdy, c.stream);
    RGPU_HIP(hipStreamSynchronize(c.stream));
    RGPU_HIP(hipMemcpy(Y, py, sizeof(double) * n * LL, hipMemcpyDefault));
    if (dY) RGPU_HIP(hipMemcpy(dY, pdy, sizeof(double) * 3 * n * LL, hipMemcpyDefault));
  });
}

}  // extern "C"
