// SOAP calculator hypers, parsed from the reference's JSON format
// (representations/calculator_spherical_expansion.hh:1596-1742,
//  calculator_spherical_invariants.hh:257-468,
//  bindings/rascal/representations/spherical_invariants.py:209-305).
#pragma once

#include <string>
#include <vector>

#include "json_lite.hh"

namespace rgpu {

struct SoapHypers {
  enum Kind { kPowerSpectrum = 0, kExpansion = 1 };
  Kind kind = kPowerSpectrum;
  int max_radial = 0;
  int max_angular = 0;
  double cutoff = 0, smooth_width = 0, gaussian_sigma = 0;
  int cutoff_type = 0;  // 0 ShiftedCosine, 1 RadialScaling
  double rs_rate = 0, rs_scale = 1, rs_exponent = 0;
  int radial_basis = 0;  // 0 GTO, 1 DVR
  int optimization = 0;  // 0 direct evaluation, 1 Spline
  // RadialDimReduction (with Spline): per species (atomic number) the
  // projections of the radial functions, [L+1][N components][N], flattened
  std::vector<int> rdr_species;
  std::vector<double> rdr_proj;
  double spline_accuracy = 0;
  bool normalize = true;
  bool compute_gradients = false;
  std::string expansion_by_species = "environment wise";
  std::vector<int> global_species;
  // coefficient_subselection (sparsified power spectrum): feature k is
  // p^{sel_a[k] sel_b[k]}[sel_n1[k], sel_n2[k], sel_l[k]] (atomic numbers a <= b)
  bool sparsified = false;
  std::vector<int> sel_a, sel_b, sel_n1, sel_n2, sel_l;
  std::string json;  // the hypers as given (calculator name, serialisation)
};

SoapHypers parse_soap_hypers(const std::string& json_text);

}  // namespace rgpu
