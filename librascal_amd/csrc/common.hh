// Shared definitions for the MI355X SOAP + sparse-GAP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace rgpu {

// Error classes mirror the reference's exception kinds; the C ABI maps them
// to return codes (see include/rascal_mi355x.h).
struct InvalidArgument : std::logic_error {
  using std::logic_error::logic_error;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess)
    throw HipError(std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file +
                   ":" + std::to_string(line) + ")");
}
#define RGPU_HIP(call) ::rgpu::hip_check((call), #call, __FILE__, __LINE__)
#define RGPU_LAUNCH_CHECK() ::rgpu::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

constexpr double kPi = 3.14159265358979323846;
constexpr double kDblFtol = 100.0 * 2.220446049250313080847e-16;
constexpr int kMaxSpecies = 64;        // key masks are 64-bit
constexpr int kMaxChemElements = 130;  // reference utils.hh:399
constexpr int kHypTerms = 200;         // series terms (expansion.hh:697)

// Per-structure neighbour-list geometry (computed on the host in FP64 from
// the cell, following adaptor_neighbour_list.hh:935-1011).
struct NLGeom {
  double cell[9];       // cell[d*3+k]: component d of lattice vector k
  double mesh_min[3];
  double ghost_min[3];
  double ghost_max[3];
  double tol;
  int nboxes[3];
  int pmin[3];
  int rep[3];
  int n_images;         // rep[0]*rep[1]*rep[2]
  int atom0;            // first atom of the structure
  int n_atoms;
  int64_t box0;         // first linked-cell box of the structure
  int64_t ext0;         // first extended (atom or ghost) slot
};

}  // namespace rgpu
