// Linked-cell full neighbour list on the GPU with the self pair first and a
// strict cutoff, for a ragged batch of structures.
//
// Reproduces the pair sets and pair order of the reference stack
// AdaptorNeighbourList -> AdaptorCenterContribution -> AdaptorStrict
// (structure_managers/adaptor_neighbour_list.hh:904-1118,
//  adaptor_center_contribution.hh:334-403, adaptor_strict.hh:373-455).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hh"
#include "device_buffer.hh"

namespace rgpu {

struct PairList {
  int64_t n_atoms = 0;
  int64_t n_pairs = 0;
  int64_t n_ext = 0;      // atoms + ghosts
  int64_t n_boxes = 0;
  int n_struct = 0;
  std::vector<NLGeom> geom_host;
  DeviceBuffer<NLGeom> geom;
  DeviceBuffer<int> atom_struct;     // [n_atoms]
  DeviceBuffer<double> pos;          // [n_atoms][3]
  DeviceBuffer<int> ghost_count;     // [n_atoms]
  DeviceBuffer<int64_t> ghost_off;   // [n_atoms+1]
  DeviceBuffer<double> ext_pos;      // [n_ext][3]
  DeviceBuffer<int> ext_orig;        // [n_ext] original atom (global index)
  DeviceBuffer<int64_t> ext_box;     // [n_ext] global box index
  DeviceBuffer<int> box_count;       // [n_boxes]
  DeviceBuffer<int64_t> box_start;   // [n_boxes+1]
  DeviceBuffer<int> box_items;       // [n_ext]
  DeviceBuffer<int> pair_count;      // [n_atoms]
  DeviceBuffer<int64_t> centre_off;  // [n_atoms+1]
  DeviceBuffer<int> pair_j;          // [n_pairs] original atom, global index
  DeviceBuffer<double> pair_vec;     // [n_pairs][3]  r_j(image) - r_i
  DeviceBuffer<double> pair_dist;    // [n_pairs]
  DeviceBuffer<int> err;             // [1]
  DeviceBuffer<int64_t> scratch, scratch2, scratch3;  // scan tile sums (one per scan)
  int64_t cap_ext = 0, cap_pairs = 0;  // allocated sizes of the extended set / pair list
  // host staging: positions and totals [ghosts, pairs, error flag]
  // (page-locked, read and written by kernels), and the last uploaded
  // structure map and geometry (not uploaded again while unchanged)
  PinnedBuffer<double> pos_stage;
  PinnedBuffer<int64_t> tot_stage;
  std::vector<int> astruct_host;           // source of the last structure-map upload
  std::vector<int64_t> offsets_uploaded;   // the atom offsets it was made from
  std::vector<NLGeom> geom_uploaded;
  // a deferred build whose sizes have not been read yet, and its capacities
  bool pending = false;
  int64_t pending_caps[2] = {0, 0};
  // HIP graph of one pass of the device build, replayed by deferred builds
  // while every size, capacity and buffer it was captured with is unchanged
  hipGraphExec_t graph = nullptr;
  std::vector<uintptr_t> graph_sig;
  int64_t graph_captures = 0, graph_launches = 0;
  PairList() = default;
  PairList(const PairList&) = delete;
  PairList& operator=(const PairList&) = delete;
  ~PairList() {
    if (graph) (void)hipGraphExecDestroy(graph);
  }
};

// Host inputs: positions [n_atoms][3], cells [n_struct][9] (lattice vectors
// as columns), pbc [n_struct][3], atom_offsets [n_struct+1].  Validates the
// structures (cell not zero, atoms inside the cell within 1e-10) and builds
// the list on `stream`.  Without defer it synchronises for the sizes (n_ext,
// n_pairs).  With defer (and capacities kept from an earlier build of the
// batch) it returns at once: device consumers may use the list straight
// away, and finish_pair_list reads the sizes later; if a capacity was short,
// the list then holds no pairs (every centre's range empty) until
// rebuild_pair_list.
// With graph (a deferred build on a stream other than the null stream), the
// device pass is replayed from a captured HIP graph.
void build_pair_list(PairList& pl, int n_struct, const int64_t* atom_offsets,
                     const double* positions, const double* cells, const int* pbc,
                     double cutoff, hipStream_t stream, bool defer = false, bool graph = false);
// Reads the sizes of a deferred build (waits for `stream`); false when a
// capacity was short (the capacities are grown for rebuild_pair_list).
// Throws if an atom fell outside the cell bins.  True when nothing is pending.
bool finish_pair_list(PairList& pl, hipStream_t stream);
// The device passes again over the positions and geometry of the last
// build, with the current capacities; synchronous.
void rebuild_pair_list(PairList& pl, double cutoff, hipStream_t stream);

// Per-structure linked-cell geometry (host, FP64).
NLGeom make_geometry(const double* cell, const int* pbc, double cutoff);

// Structure checks of structure_manager_centers.cc:67-99 and
// atomic_structure.hh:249-255 (throws RuntimeError).
void check_structure(int n_atoms, const double* pos, const double* cell);

}  // namespace rgpu
