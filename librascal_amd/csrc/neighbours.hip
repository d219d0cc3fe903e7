// Linked-cell neighbour list kernels.  Built with -ffp-contract=off: every
// geometric quantity (ghost image positions, box indices, squared distances)
// is rounded exactly as the reference's separate multiply/add sequence so the
// inclusive cutoff test takes the same decisions.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "neighbours.hh"
#include "scan.hh"

namespace rgpu {

namespace {

void inverse3(const double* C, double* Ci) {
  double a = C[0], b = C[1], c = C[2], d = C[3], e = C[4], f = C[5], g = C[6], h = C[7],
         i = C[8];
  double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
  double id = 1.0 / (a * A + b * B + c * Cc);
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

__device__ __forceinline__ void box_of(const NLGeom& g, const double* x, double rc, int* b) {
  for (int d = 0; d < 3; ++d) {
    double v = x[d] - g.mesh_min[d];
    int k = int(floor(v / rc));
    b[d] = k < 1 ? 1 : k;
  }
}

__device__ __forceinline__ int64_t box_lin(const NLGeom& g, const int* b) {
  return g.box0 + (int64_t(b[0]) * g.nboxes[1] + b[1]) * g.nboxes[2] + b[2];
}

// ghost images of atom a that fall inside the ghost bounds
// (adaptor_neighbour_list.hh:1044-1072); count or write.  One wavefront per
// atom, its lanes over the images in the order of the linear index
// ((p0 - pmin0) rep1 + p1 - pmin1) rep2 + p2 - pmin2 (p2 fastest), 64 per
// step; the ballot keeps that order.  The write pass also puts the atom
// itself in the first n_atoms slots; the count pass also clears the error
// flag and the box counters of the binning that follows.
template <bool WRITE>
__global__ __launch_bounds__(256) void ghosts_kernel(const NLGeom* __restrict__ geom,
                                                     const int* __restrict__ atom_struct,
                                                     const double* __restrict__ pos, int64_t n_atoms,
                                                     int* __restrict__ count, const int64_t* __restrict__ off,
                                                     double* __restrict__ ext_pos, int* __restrict__ ext_orig,
                                                     int64_t cap, int* __restrict__ err,
                                                     int* __restrict__ box_count, int64_t n_boxes) {
  const int lane = threadIdx.x & 63;
  const int64_t a = int64_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (a >= n_atoms) return;
  if (!WRITE && lane == 0) {
    if (a == 0) *err = 0;
    for (int64_t b = a; b < n_boxes; b += n_atoms) box_count[b] = 0;
  }
  const NLGeom& g = geom[atom_struct[a]];
  const double xs[3] = {pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]};
  const int r12 = g.rep[1] * g.rep[2], n_img = g.rep[0] * r12;
  int c = 0;
  const int64_t w = WRITE ? off[a] + n_atoms : 0;
  if (WRITE && lane == 0) {
    ext_pos[3 * a] = xs[0];
    ext_pos[3 * a + 1] = xs[1];
    ext_pos[3 * a + 2] = xs[2];
    ext_orig[a] = int(a);
  }
  for (int k0 = 0; k0 < n_img; k0 += 64) {
    const int k = k0 + lane;
    bool inside = false;
    double xg[3] = {0, 0, 0};
    if (k < n_img) {
      const int q0 = k / r12, rem = k - q0 * r12, q1 = rem / g.rep[2], q2 = rem - q1 * g.rep[2];
      const int p0 = g.pmin[0] + q0, p1 = g.pmin[1] + q1, p2 = g.pmin[2] + q2;
      if (!(p0 == 0 && p1 == 0 && p2 == 0)) {
        inside = true;
        for (int d = 0; d < 3; ++d) {
          double t = g.cell[d * 3 + 0] * double(p0);
          t = t + g.cell[d * 3 + 1] * double(p1);
          t = t + g.cell[d * 3 + 2] * double(p2);
          xg[d] = xs[d] + t;
          if (!((xg[d] - g.ghost_min[d]) > g.tol && (xg[d] - g.ghost_max[d]) < -g.tol)) inside = false;
        }
      }
    }
    const unsigned long long mask = __ballot(inside);
    if (WRITE && inside) {
      const int64_t o = w + c + __popcll(mask & ((1ull << lane) - 1ull));
      if (o < cap) {
        ext_pos[3 * o] = xg[0];
        ext_pos[3 * o + 1] = xg[1];
        ext_pos[3 * o + 2] = xg[2];
        ext_orig[o] = int(a);
      }
    }
    c += __popcll(mask);
  }
  if (!WRITE && lane == 0) count[a] = c;
}


// box of every extended slot (adaptor_neighbour_list.hh:1076-1083,
// :395-422 range check)
__global__ void bin_count_kernel(const NLGeom* __restrict__ geom, const int* __restrict__ atom_struct,
                                 const double* __restrict__ ext_pos, const int* __restrict__ ext_orig,
                                 int64_t n_atoms, const int64_t* __restrict__ n_ghosts, int64_t cap,
                                 double rc, int64_t* __restrict__ ext_box,
                                 int* __restrict__ box_count, int* __restrict__ err) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n_ext = n_atoms + *n_ghosts;
  if (t >= n_ext || t >= cap) return;
  const NLGeom& g = geom[atom_struct[ext_orig[t]]];
  int b[3];
  box_of(g, ext_pos + 3 * t, rc, b);
  for (int d = 0; d < 3; ++d)
    if (b[d] >= g.nboxes[d] - 1 || b[d] <= 0) {
      atomicOr(err, 1);
      ext_box[t] = -1;
      return;
    }
  int64_t bl = box_lin(g, b);
  ext_box[t] = bl;
  atomicAdd(box_count + bl, 1);
}

// slots of a box taken from the top down (box_count returns to 0); the
// order inside a box is set by bin_sort_kernel
__global__ void bin_fill_kernel(const int64_t* __restrict__ ext_box, int64_t n_atoms,
                                const int64_t* __restrict__ n_ghosts, int64_t cap,
                                const int64_t* __restrict__ box_start, int* __restrict__ box_count,
                                int* __restrict__ box_items) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n_ext = n_atoms + *n_ghosts;
  if (t >= n_ext || t >= cap) return;
  int64_t bl = ext_box[t];
  if (bl < 0) return;
  int k = atomicSub(box_count + bl, 1) - 1;
  box_items[box_start[bl] + k] = int(t);
}

// ascending slot order within each box = the reference's tag order; one
// wavefront per box: boxes of up to 64 items are ranked in registers (an
// item's rank is the number of smaller items; slots are distinct), larger
// ones are sorted in place by the wavefront's first lane
__global__ __launch_bounds__(256) void bin_sort_kernel(const int64_t* __restrict__ box_start, int64_t n_boxes,
                                                       int* __restrict__ box_items) {
  const int lane = threadIdx.x & 63;
  const int64_t bx = int64_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (bx >= n_boxes) return;
  const int64_t s = box_start[bx], e = box_start[bx + 1];
  const int m = int(e - s);
  if (e - s <= 64) {
    const int v = lane < m ? box_items[s + lane] : 0;
    int r = 0;
    for (int q = 0; q < m; ++q) r += __shfl(v, q, 64) < v ? 1 : 0;
    if (lane < m) box_items[s + r] = v;  // every lane read its item above
    return;
  }
  if (lane != 0) return;
  for (int64_t i = s + 1; i < e; ++i) {
    int v = box_items[i];
    int64_t j = i - 1;
    while (j >= s && box_items[j] > v) {
      box_items[j + 1] = box_items[j];
      --j;
    }
    box_items[j + 1] = v;
  }
}

// One wavefront per centre: 27-box stencil (last dimension fastest,
// adaptor_neighbour_list.hh:92-158), strict inclusive cutoff on the squared
// distance (adaptor_strict.hh:414-427).  Pass 1 counts, pass 2 writes in
// stencil/box order, preceded by the self pair.  The candidates of the 27
// boxes are taken as one run (box s's items at [first[s], first[s] + size
// of s)), 64 per step, so that the lanes stay busy although a box holds only
// a few atoms; candidate k is item k - first[s] of the last box s whose
// first[s] <= k, and the ballot keeps the stencil/box order.
template <bool WRITE>
__global__ __launch_bounds__(256) void pairs_kernel(
    const NLGeom* __restrict__ geom, const int* __restrict__ atom_struct,
    const double* __restrict__ ext_pos, const int* __restrict__ ext_orig,
    const int64_t* __restrict__ box_start, const int* __restrict__ box_items, int64_t n_atoms,
    double rc, int* __restrict__ pair_count, const int64_t* __restrict__ centre_off,
    int* __restrict__ pair_j, double* __restrict__ pair_vec, double* __restrict__ pair_dist,
    int64_t cap) {
  __shared__ int64_t first[4][32], base[4][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i = int64_t(blockIdx.x) * (blockDim.x / 64) + wave;
  if (i >= n_atoms) return;  // whole wavefronts; no workgroup barrier follows
  const NLGeom& g = geom[atom_struct[i]];
  const double xi0 = ext_pos[3 * i], xi1 = ext_pos[3 * i + 1], xi2 = ext_pos[3 * i + 2];
  const double xi[3] = {xi0, xi1, xi2};
  int b[3];
  box_of(g, xi, rc, b);
  const double rc2 = rc * rc;
  int64_t w = WRITE ? centre_off[i] : 0;
  if (WRITE && lane == 0 && w < cap) {
    pair_j[w] = int(i);
    pair_vec[3 * w] = 0.0;
    pair_vec[3 * w + 1] = 0.0;
    pair_vec[3 * w + 2] = 0.0;
    pair_dist[w] = 0.0;
  }
  // lane s < 27: box s of the stencil; inclusive scan of the box sizes
  int64_t size = 0, start = 0;
  if (lane < 27) {
    int c[3];
    int f = 1;
    for (int d = 2; d >= 0; --d) {
      c[d] = lane / f % 3 + b[d] - 1;
      f *= 3;
    }
    const int64_t bl = box_lin(g, c);
    start = box_start[bl];
    size = box_start[bl + 1] - start;
  }
  int64_t incl = size;
  for (int off = 1; off < 32; off <<= 1) {
    const int64_t o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  const int64_t total = __shfl(incl, 26, 64);
  if (lane < 27) {
    first[wave][lane] = incl - size;
    base[wave][lane] = start;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int64_t n = 1;
  for (int64_t k0 = 0; k0 < total; k0 += 64) {
    const int64_t k = k0 + lane;
    bool keep = false;
    int t = -1;
    double v0 = 0, v1 = 0, v2 = 0, d2 = 0;
    if (k < total) {
      // last box whose run starts at or before k (empty boxes share their
      // start with the next box and are passed over)
      int lo = 0, hi = 26;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[wave][mid] <= k)
          lo = mid;
        else
          hi = mid - 1;
      }
      t = box_items[base[wave][lo] + (k - first[wave][lo])];
      if (t != int(i)) {
        v0 = ext_pos[3 * t] - xi0;
        v1 = ext_pos[3 * t + 1] - xi1;
        v2 = ext_pos[3 * t + 2] - xi2;
        double a0 = v0 * v0;
        double a1 = v1 * v1;
        double a2 = v2 * v2;
        d2 = (a0 + a1) + a2;
        keep = d2 <= rc2;
      }
    }
    unsigned long long mask = __ballot(keep);
    if (WRITE && keep) {
      unsigned long long below = mask & ((1ull << lane) - 1ull);
      int64_t o = w + n + __popcll(below);
      if (o < cap) {
        pair_j[o] = ext_orig[t];
        pair_vec[3 * o] = v0;
        pair_vec[3 * o + 1] = v1;
        pair_vec[3 * o + 2] = v2;
        pair_dist[o] = sqrt(d2);
      }
    }
    n += __popcll(mask);
  }
  if (!WRITE && lane == 0) pair_count[i] = int(n);
}

__global__ void copy_f64_kernel(const double* __restrict__ src, int64_t n, double* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Sizes of the build; when a capacity was short, every centre's pair range
// is emptied (centre_off = 0) so that consumers launched before the host
// sees the sizes read nothing beyond the buffers.
__global__ void nl_totals_kernel(const int64_t* __restrict__ n_ghosts, int64_t* __restrict__ centre_off,
                                 int64_t n_atoms, int64_t cap_ext, int64_t cap_pairs,
                                 const int* __restrict__ err, int64_t* __restrict__ out) {
  const int64_t ng = *n_ghosts, np = centre_off[n_atoms];
  __syncthreads();  // every thread has read the pair total before it may be cleared
  if (threadIdx.x == 0) {
    out[0] = ng;
    out[1] = np;
    out[2] = *err;
  }
  if (n_atoms + ng > cap_ext || np > cap_pairs)
    for (int64_t a = threadIdx.x; a <= n_atoms; a += blockDim.x) centre_off[a] = 0;
}

}  // namespace

NLGeom make_geometry(const double* cell, const int* pbc, double rc) {
  NLGeom g{};
  for (int k = 0; k < 9; ++k) g.cell[k] = cell[k];
  for (int i = 0; i < 3; ++i) {
    double mn = 0.0, mx = 0.0;
    for (int col = 0; col < 3; ++col) {
      double v = cell[i * 3 + col];
      mn += v * static_cast<double>(v < 0.);
      mx += v * static_cast<double>(v > 0.);
    }
    g.mesh_min[i] = mn - 2. * rc;
    double lmesh = mx - mn + 4. * rc;
    g.nboxes[i] = static_cast<int>(std::ceil(lmesh / rc));
    g.ghost_min[i] = mn - rc;
    double lghost = mx - mn + 2 * rc;
    g.ghost_max[i] = std::ceil(lghost / rc) * rc + g.ghost_min[i];
  }
  double maxlen = std::max({g.ghost_max[0] - g.ghost_min[0], g.ghost_max[1] - g.ghost_min[1],
                            g.ghost_max[2] - g.ghost_min[2]});
  g.tol = maxlen * 1e-8;
  double lo[3], hi[3], Ci[9];
  for (int i = 0; i < 3; ++i) {
    lo[i] = g.ghost_min[i] - g.tol;
    hi[i] = g.ghost_max[i] + g.tol;
  }
  inverse3(cell, Ci);
  double xmin[3] = {INFINITY, INFINITY, INFINITY}, xmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = ((c >> i) & 1) ? hi[i] : lo[i];
    for (int d = 0; d < 3; ++d) {
      double m = Ci[d * 3 + 0] * x[0] + Ci[d * 3 + 1] * x[1] + Ci[d * 3 + 2] * x[2];
      xmin[d] = std::min(xmin[d], m);
      xmax[d] = std::max(xmax[d], m);
    }
  }
  g.n_images = 1;
  for (int i = 0; i < 3; ++i) {
    int mmin = static_cast<int>(std::floor(xmin[i]));
    int mmax = static_cast<int>(std::ceil(xmax[i])) - 1;
    if (pbc[i]) {
      g.pmin[i] = mmin;
      g.rep[i] = mmax - mmin + 1;
    } else {
      g.pmin[i] = 0;
      g.rep[i] = 1;
    }
    g.n_images *= g.rep[i];
  }
  return g;
}

void check_structure(int n_atoms, const double* pos, const double* cell) {
  bool all_zero = true;
  for (int k = 0; k < 9; ++k)
    if (std::fabs(cell[k]) >= 1e-10) all_zero = false;
  if (all_zero)
    throw RuntimeError(
        "The unit cell should not be filled with zeros but it should contain all the atoms "
        "even if the structure is not periodic.");
  double Ci[9];
  inverse3(cell, Ci);
  // branch-free over the atoms (vectorisable), one test at the end
  bool outside = false;
  for (int a = 0; a < n_atoms; ++a) {
    const double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
    for (int d = 0; d < 3; ++d) {
      const double s = Ci[d * 3] * x + Ci[d * 3 + 1] * y + Ci[d * 3 + 2] * z;
      outside |= (s < -1e-10) | (s > 1. + 1e-10);
    }
  }
  if (outside)
    throw RuntimeError(
        "Some of the atoms in the structure are not inside the unit cell. Please wrap them "
        "inside it with at least tolerance of 1e-10.");
}

namespace {

// true when the sizes read back fit the capacities (then recorded); else
// the capacities are grown for the next pass
bool settle(PairList& pl, int64_t cap_ext, int64_t cap_pairs) {
  const int64_t* tot = pl.tot_stage.get();
  const int64_t n_ext = pl.n_atoms + tot[0], n_pairs = tot[1];
  if (n_ext <= cap_ext && n_pairs <= cap_pairs) {
    if (tot[2])
      throw RuntimeError(
          "Error: this atom does not fall in one of the linked cell bin. It might have not been "
          "wrapped properly.");
    pl.n_ext = n_ext;
    pl.n_pairs = n_pairs;
    return true;
  }
  pl.cap_ext = std::max(cap_ext, n_ext + n_ext / 8 + 64);
  pl.cap_pairs = std::max(cap_pairs, n_pairs + n_pairs / 8 + 64);
  return false;
}

// every buffer of one device pass at the current capacities (reserved
// before a capture, so that nothing is allocated inside it)
void reserve_pass(PairList& pl) {
  const int64_t n_atoms = pl.n_atoms, boxes = pl.n_boxes;
  pl.pos.reserve(3 * n_atoms);
  pl.err.reserve(1);
  pl.ghost_count.reserve(n_atoms);
  pl.ghost_off.reserve(n_atoms + 1);
  pl.ext_pos.reserve(3 * pl.cap_ext);
  pl.ext_orig.reserve(pl.cap_ext);
  pl.box_count.reserve(boxes);
  pl.box_start.reserve(boxes + 1);
  pl.box_items.reserve(pl.cap_ext);
  pl.ext_box.reserve(pl.cap_ext);
  pl.pair_count.reserve(n_atoms);
  pl.centre_off.reserve(n_atoms + 1);
  pl.pair_j.reserve(pl.cap_pairs);
  pl.pair_vec.reserve(3 * pl.cap_pairs);
  pl.pair_dist.reserve(pl.cap_pairs);
  pl.tot_stage.reserve(3);
  exclusive_scan_reserve(n_atoms, pl.scratch);
  exclusive_scan_reserve(boxes, pl.scratch2);
  exclusive_scan_reserve(n_atoms, pl.scratch3);
}

// One device pass (kernels only): positions from the page-locked staging,
// ghosts, binning, pairs, and the sizes written to page-locked memory.
void launch_pass(PairList& pl, double rc, hipStream_t stream) {
  const int64_t n_atoms = pl.n_atoms, boxes = pl.n_boxes;
  const int64_t cap_ext = pl.cap_ext, cap_pairs = pl.cap_pairs;
  const NLGeom* dgeom = pl.geom.get();
  const int* dstruct = pl.atom_struct.get();
  double* dpos = pl.pos.get();
  const int TB = 256;
  auto blocks = [](int64_t n, int tb) { return dim3(unsigned((n + tb - 1) / tb)); };
  hipLaunchKernelGGL(copy_f64_kernel, blocks(3 * n_atoms, TB), dim3(TB), 0, stream, pl.pos_stage.device(),
                     3 * n_atoms, dpos);
  RGPU_LAUNCH_CHECK();
  int* err = pl.err.get();
  int* bcount = pl.box_count.get();
  // ghosts (one wavefront per atom); the count pass clears err and bcount
  int* gcount = pl.ghost_count.get();
  int64_t* goff = pl.ghost_off.get();
  const dim3 ggrid(unsigned((n_atoms + 3) / 4));
  hipLaunchKernelGGL(ghosts_kernel<false>, ggrid, dim3(256), 0, stream, dgeom, dstruct,
                     dpos, n_atoms, gcount, nullptr, nullptr, nullptr, int64_t(0), err, bcount, boxes);
  RGPU_LAUNCH_CHECK();
  exclusive_scan_device<int>(gcount, goff, n_atoms, stream, pl.scratch);
  double* epos = pl.ext_pos.get();
  int* eorig = pl.ext_orig.get();
  hipLaunchKernelGGL(ghosts_kernel<true>, ggrid, dim3(256), 0, stream, dgeom, dstruct,
                     dpos, n_atoms, nullptr, goff, epos, eorig, cap_ext, nullptr, nullptr, int64_t(0));
  RGPU_LAUNCH_CHECK();
  // binning
  int64_t* bstart = pl.box_start.get();
  int* bitems = pl.box_items.get();
  int64_t* ebox = pl.ext_box.get();
  hipLaunchKernelGGL(bin_count_kernel, blocks(cap_ext, TB), dim3(TB), 0, stream, dgeom, dstruct, epos,
                     eorig, n_atoms, goff + n_atoms, cap_ext, rc, ebox, bcount, err);
  RGPU_LAUNCH_CHECK();
  exclusive_scan_device<int>(bcount, bstart, boxes, stream, pl.scratch2);
  hipLaunchKernelGGL(bin_fill_kernel, blocks(cap_ext, TB), dim3(TB), 0, stream, ebox, n_atoms,
                     goff + n_atoms, cap_ext, bstart, bcount, bitems);
  RGPU_LAUNCH_CHECK();
  hipLaunchKernelGGL(bin_sort_kernel, dim3(unsigned((boxes + 3) / 4)), dim3(256), 0, stream, bstart, boxes,
                     bitems);
  RGPU_LAUNCH_CHECK();
  // pairs
  int* pcount = pl.pair_count.get();
  int64_t* coff = pl.centre_off.get();
  const int WPB = TB / 64;
  dim3 pgrid(unsigned((n_atoms + WPB - 1) / WPB));
  hipLaunchKernelGGL(pairs_kernel<false>, pgrid, dim3(TB), 0, stream, dgeom, dstruct, epos, eorig,
                     bstart, bitems, n_atoms, rc, pcount, nullptr, nullptr, nullptr, nullptr,
                     int64_t(0));
  RGPU_LAUNCH_CHECK();
  exclusive_scan_device<int>(pcount, coff, n_atoms, stream, pl.scratch3);
  hipLaunchKernelGGL(pairs_kernel<true>, pgrid, dim3(TB), 0, stream, dgeom, dstruct, epos, eorig,
                     bstart, bitems, n_atoms, rc, nullptr, coff, pl.pair_j.get(), pl.pair_vec.get(),
                     pl.pair_dist.get(), cap_pairs);
  RGPU_LAUNCH_CHECK();
  hipLaunchKernelGGL(nl_totals_kernel, dim3(1), dim3(256), 0, stream, goff + n_atoms, coff, n_atoms,
                     cap_ext, cap_pairs, err, pl.tot_stage.device());
  RGPU_LAUNCH_CHECK();
}

// one pass replayed from its graph (captured again when anything it depends
// on changed)
void launch_pass_graph(PairList& pl, double rc, hipStream_t stream) {
  const std::vector<uintptr_t> sig = {
      uintptr_t(stream), uintptr_t(pl.n_atoms), uintptr_t(pl.n_boxes), uintptr_t(pl.cap_ext),
      uintptr_t(pl.cap_pairs), uintptr_t(std::llround(rc * 1e12)), uintptr_t(pl.geom.get()),
      uintptr_t(pl.atom_struct.get()), uintptr_t(pl.pos.get()), uintptr_t(pl.pos_stage.device()),
      uintptr_t(pl.err.get()), uintptr_t(pl.ghost_count.get()), uintptr_t(pl.ghost_off.get()),
      uintptr_t(pl.ext_pos.get()), uintptr_t(pl.ext_orig.get()), uintptr_t(pl.box_count.get()),
      uintptr_t(pl.box_start.get()), uintptr_t(pl.box_items.get()),
      uintptr_t(pl.ext_box.get()), uintptr_t(pl.pair_count.get()), uintptr_t(pl.centre_off.get()),
      uintptr_t(pl.pair_j.get()), uintptr_t(pl.pair_vec.get()), uintptr_t(pl.pair_dist.get()),
      uintptr_t(pl.tot_stage.device()), uintptr_t(pl.scratch.get()), uintptr_t(pl.scratch2.get()),
      uintptr_t(pl.scratch3.get())};
  if (!pl.graph || sig != pl.graph_sig) {
    if (pl.graph) (void)hipGraphExecDestroy(pl.graph);
    pl.graph = nullptr;
    hipGraph_t g = nullptr;
    RGPU_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      launch_pass(pl, rc, stream);
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    RGPU_HIP(hipStreamEndCapture(stream, &g));
    const hipError_t e = hipGraphInstantiate(&pl.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    RGPU_HIP(e);
    pl.graph_sig = sig;
    ++pl.graph_captures;
  }
  RGPU_HIP(hipGraphLaunch(pl.graph, stream));
  ++pl.graph_launches;
}

// the device passes over the uploaded positions and geometry
void run_passes(PairList& pl, double rc, hipStream_t stream, bool defer, bool graph) {
  // a short ghost capacity also hides pairs, so up to three passes
  for (int pass = 0; pass < 4; ++pass) {
    const int64_t cap_ext = pl.cap_ext, cap_pairs = pl.cap_pairs;
    reserve_pass(pl);
    if (defer && graph)
      launch_pass_graph(pl, rc, stream);
    else
      launch_pass(pl, rc, stream);
    if (defer) {
      pl.pending = true;
      pl.pending_caps[0] = cap_ext;
      pl.pending_caps[1] = cap_pairs;
      return;
    }
    RGPU_HIP(hipStreamSynchronize(stream));
    if (settle(pl, cap_ext, cap_pairs)) return;
  }
  throw RuntimeError("neighbour list buffers could not be sized");
}

}  // namespace

bool finish_pair_list(PairList& pl, hipStream_t stream) {
  if (!pl.pending) return true;
  RGPU_HIP(hipStreamSynchronize(stream));
  pl.pending = false;
  return settle(pl, pl.pending_caps[0], pl.pending_caps[1]);
}

void rebuild_pair_list(PairList& pl, double rc, hipStream_t stream) {
  if (pl.n_atoms == 0) return;
  run_passes(pl, rc, stream, false, false);
}

void build_pair_list(PairList& pl, int n_struct, const int64_t* atom_offsets,
                     const double* positions, const double* cells, const int* pbc, double rc,
                     hipStream_t stream, bool defer, bool graph) {
  // an unchecked earlier build still owns the staging buffers
  if (pl.pending) (void)finish_pair_list(pl, stream);
  const int64_t n_atoms = atom_offsets[n_struct];
  pl.n_struct = n_struct;
  pl.n_atoms = n_atoms;
  pl.geom_host.resize(n_struct);
  int64_t boxes = 0;
  for (int s = 0; s < n_struct; ++s) {
    int64_t a0 = atom_offsets[s], a1 = atom_offsets[s + 1];
    if (a1 < a0) throw InvalidArgument("atom_offsets must be non-decreasing");
    check_structure(int(a1 - a0), positions + 3 * a0, cells + 9 * s);
    NLGeom g = make_geometry(cells + 9 * s, pbc + 3 * s, rc);
    g.atom0 = int(a0);
    g.n_atoms = int(a1 - a0);
    g.box0 = boxes;
    boxes += int64_t(g.nboxes[0]) * g.nboxes[1] * g.nboxes[2];
    pl.geom_host[s] = g;
  }
  pl.n_boxes = boxes;
  if (n_atoms == 0) {
    pl.n_pairs = 0;
    pl.n_ext = 0;
    pl.centre_off.reserve(1);
    RGPU_HIP(hipMemsetAsync(pl.centre_off.get(), 0, sizeof(int64_t), stream));
    return;
  }
  // geometry and structure map: uploaded again only when they change (the
  // buffers keep their addresses while they do not grow)
  const NLGeom* geom_before = pl.geom.get();
  NLGeom* dgeom = pl.geom.reserve(n_struct);
  if (dgeom != geom_before || pl.geom_uploaded.size() != pl.geom_host.size() ||
      std::memcmp(pl.geom_uploaded.data(), pl.geom_host.data(), sizeof(NLGeom) * n_struct) != 0) {
    RGPU_HIP(hipMemcpyAsync(dgeom, pl.geom_host.data(), sizeof(NLGeom) * n_struct,
                            hipMemcpyHostToDevice, stream));
    pl.geom_uploaded = pl.geom_host;
  }
  // the structure of each atom, from the atom offsets
  const int* struct_before = pl.atom_struct.get();
  int* dstruct = pl.atom_struct.reserve(n_atoms);
  if (dstruct != struct_before || pl.offsets_uploaded.size() != size_t(n_struct + 1) ||
      !std::equal(atom_offsets, atom_offsets + n_struct + 1, pl.offsets_uploaded.begin())) {
    pl.astruct_host.resize(size_t(n_atoms));
    for (int s = 0; s < n_struct; ++s)
      std::fill(pl.astruct_host.begin() + atom_offsets[s], pl.astruct_host.begin() + atom_offsets[s + 1], s);
    RGPU_HIP(hipMemcpyAsync(dstruct, pl.astruct_host.data(), sizeof(int) * n_atoms,
                            hipMemcpyHostToDevice, stream));
    pl.offsets_uploaded.assign(atom_offsets, atom_offsets + n_struct + 1);
  }
  // positions through page-locked staging, read by each device pass with a
  // kernel (free: every earlier build has been waited for, the last one
  // just above if it was deferred)
  double* ps = pl.pos_stage.reserve(size_t(3 * n_atoms));
  std::memcpy(ps, positions, sizeof(double) * 3 * n_atoms);

  // Sizes of the extended (ghost) set and of the pair list are only known on
  // the device; the buffers are sized by capacities kept from earlier builds,
  // every write beyond a capacity is dropped, and the true sizes are read
  // back: at once, or (defer) by finish_pair_list after the consumers of the
  // list have been enqueued.  When a capacity was short, the build is run
  // again with enough room (the first build of a batch, or a large change).
  if (pl.cap_ext < n_atoms) pl.cap_ext = 2 * n_atoms + 1024;
  if (pl.cap_pairs < n_atoms) pl.cap_pairs = 48 * n_atoms + 1024;
  // the first build of a batch sizes the buffers synchronously
  run_passes(pl, rc, stream, defer && pl.n_pairs > 0, graph && stream != nullptr);
}

}  // namespace rgpu
