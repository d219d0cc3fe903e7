#include "hypers.hh"

#include <algorithm>
#include <set>
#include <string>

#include "common.hh"

namespace rgpu {

namespace {
double unit_value(const json::Value& v) {
  // {"value": x, "unit": "AA"} or a bare number
  if (v.is_object()) return v.at("value").as_number();
  return v.as_number();
}

// RadialDimReduction projection matrices: {"Z": [L+1][N][N]}
// (calculator_spherical_expansion.hh:1188-1323)
void parse_projections(const json::Value& rdr, SoapHypers& o) {
  if (!rdr.has("projection_matrices")) throw InvalidArgument("No projection matrices were given.");
  const json::Value& pm = rdr.at("projection_matrices");
  if (!pm.is_object() || pm.obj.empty())
    throw InvalidArgument(
        "Projection matrices should be of type std::map<std::string, "
        "std::vector<std::vector<std::vector<double>>>>.");
  const int L1 = o.max_angular + 1, N = o.max_radial;
  const std::string shape = "projection_matrices should have the shape (n_species, max_angular+1, "
                            "max_radial, max_radial)";
  for (const auto& kv : pm.obj) {
    int z = 0;
    try {
      z = std::stoi(kv.first);
    } catch (const std::exception&) {
      throw InvalidArgument("projection_matrices keys must be atomic numbers, got '" + kv.first + "'");
    }
    const json::Value& per_l = kv.second;
    if (per_l.kind != json::Value::Array || int(per_l.arr.size()) != L1)
      throw InvalidArgument("Projection matrices at species=" + kv.first +
                            " should have max_angular+1 entries; " + shape);
    o.rdr_species.push_back(z);
    for (const json::Value& m : per_l.arr) {
      if (m.kind != json::Value::Array || int(m.arr.size()) != N)
        throw InvalidArgument("Projection matrices at species=" + kv.first +
                              " should have max_radial rows; " + shape);
      for (const json::Value& row : m.arr) {
        if (row.kind != json::Value::Array || int(row.arr.size()) != N)
          throw InvalidArgument("Projection matrices at species=" + kv.first +
                                " should have max_radial columns; " + shape);
        for (const json::Value& x : row.arr) o.rdr_proj.push_back(x.as_number());
      }
    }
  }
}
std::vector<int> int_list(const json::Value& v) {
  if (v.kind != json::Value::Array) throw InvalidArgument("json: value is not a list");
  std::vector<int> out;
  for (const json::Value& x : v.arr) out.push_back(static_cast<int>(x.as_number()));
  return out;
}

// coefficient_subselection {"a", "b", "n1", "n2", "l"} with the checks of
// calculator_spherical_invariants.hh:315-393 (the same messages)
void parse_subselection(const json::Value& cs, SoapHypers& o) {
  o.sparsified = true;
  o.sel_a = int_list(cs.at("a"));
  o.sel_b = int_list(cs.at("b"));
  o.sel_n1 = int_list(cs.at("n1"));
  o.sel_n2 = int_list(cs.at("n2"));
  o.sel_l = int_list(cs.at("l"));
  const size_t n = o.sel_a.size();
  const char* pre = "coefficient_subselection should have elements of the same size but: ";
  if (o.sel_b.size() != n) throw InvalidArgument(std::string(pre) + "a.size() != b.size()");
  if (o.sel_n1.size() != n) throw InvalidArgument(std::string(pre) + "a.size() != n1.size()");
  if (o.sel_n2.size() != n) throw InvalidArgument(std::string(pre) + "a.size() != n2.size()");
  if (o.sel_l.size() != n) throw InvalidArgument(std::string(pre) + "a.size() != l.size()");
  if (n == 0) throw InvalidArgument("coefficient_subselection selects no coefficient");
  auto max_of = [](const std::vector<int>& v) { return *std::max_element(v.begin(), v.end()); };
  auto min_of = [](const std::vector<int>& v) { return *std::min_element(v.begin(), v.end()); };
  if (min_of(o.sel_n1) < 0 || min_of(o.sel_n2) < 0 || min_of(o.sel_l) < 0)
    throw InvalidArgument("coefficient_subselection indices must be non-negative");
  const int lmax = max_of(o.sel_l), n1max = max_of(o.sel_n1), n2max = max_of(o.sel_n2);
  if (lmax >= o.max_angular + 1)
    throw RuntimeError("'max(l)' in coefficient_subselection is larger than 'max_angular+1'" +
                       std::to_string(lmax) + " >= " + std::to_string(o.max_angular + 1));
  if (n1max >= o.max_radial)
    throw RuntimeError("'max(n1)' in coefficient_subselection is larger than 'max_radial'" +
                       std::to_string(n1max) + " >= " + std::to_string(o.max_radial));
  if (n2max >= o.max_radial)
    throw RuntimeError("'max(n2)' in coefficient_subselection is larger than 'max_radial'" +
                       std::to_string(n2max) + " >= " + std::to_string(o.max_radial));
  for (size_t k = 0; k < n; ++k)
    if (o.sel_a[k] > o.sel_b[k])
      throw InvalidArgument(
          "coefficient_subselection should have only lexicographically sorted pairs of atomic "
          "species in the selection but: a > b");
}
}  // namespace

SoapHypers parse_soap_hypers(const std::string& text) {
  json::Value h;
  try {
    h = json::parse(text);
  } catch (const std::invalid_argument& e) {
    throw InvalidArgument(e.what());
  }
  SoapHypers o;
  o.json = text;
  try {
    o.max_radial = static_cast<int>(h.at("max_radial").as_number());
    o.max_angular = static_cast<int>(h.at("max_angular").as_number());
    if (h.has("soap_type")) {
      const std::string& t = h.at("soap_type").as_string();
      if (t == "PowerSpectrum") {
        o.kind = SoapHypers::kPowerSpectrum;
      } else if (t == "RadialSpectrum" || t == "BiSpectrum") {
        throw InvalidArgument("soap_type '" + t +
                              "' is not provided by this library (PowerSpectrum only)");
      } else {
        throw InvalidArgument("Requested SphericalInvariants type '" + t +
                              "' has not been implemented.  Must be one of: "
                              "'PowerSpectrum', 'RadialSpectrum', or 'BiSpectrum'.");
      }
      o.normalize = h.at("normalize").as_bool();
      if (h.has("coefficient_subselection")) parse_subselection(h.at("coefficient_subselection"), o);
    } else {
      o.kind = SoapHypers::kExpansion;
    }
    if (h.has("compute_gradients")) o.compute_gradients = h.at("compute_gradients").as_bool();
    if (h.has("expansion_by_species_method")) {
      const std::string& m = h.at("expansion_by_species_method").as_string();
      if (m != "environment wise" && m != "user defined" && m != "structure wise")
        throw InvalidArgument("expansion_by_species_method provided:'" + m +
                              "' is not part of the implemented methods");
      o.expansion_by_species = m;
    }
    if (h.has("global_species")) {
      for (auto& v : h.at("global_species").arr)
        o.global_species.push_back(static_cast<int>(v.as_number()));
    } else if (o.expansion_by_species == "user defined") {
      throw InvalidArgument(
          "expansion_by_species_method is 'user defined' but global_species is not defined.");
    }
    const json::Value& gd = h.at("gaussian_density");
    const std::string& gt = gd.at("type").as_string();
    if (gt != "Constant")
      throw InvalidArgument("Requested Smearing type '" + gt +
                            "' has not been implemented.  Must be one of: 'Constant'.");
    o.gaussian_sigma = unit_value(gd.at("gaussian_sigma"));
    const json::Value& rcn = h.at("radial_contribution");
    const std::string& rt = rcn.at("type").as_string();
    // calculator_spherical_expansion.hh:1682-1699
    if (rt == "GTO")
      o.radial_basis = 0;
    else if (rt == "DVR")
      o.radial_basis = 1;
    else
      throw InvalidArgument("Requested Radial contribution type '" + rt +
                            "' has not been implemented.  Must be one of: 'GTO' or 'DVR'. ");
    o.optimization = 0;
    if (rcn.has("optimization")) {
      const json::Value& op = rcn.at("optimization");
      bool spline = op.has("Spline"), rdr = op.has("RadialDimReduction");
      if (rdr && !spline) throw InvalidArgument("RadialDimReduction requires' Spline to be set.");
      if (rdr) parse_projections(op.at("RadialDimReduction"), o);
      if (spline) {
        o.optimization = 1;
        const json::Value& sp = op.at("Spline");
        if (!sp.has("accuracy")) throw InvalidArgument("No Spline accuracy was given.");
        o.spline_accuracy = sp.at("accuracy").as_number();
      }
    }
    const json::Value& fc = h.at("cutoff_function");
    const std::string& ft = fc.at("type").as_string();
    o.cutoff = unit_value(fc.at("cutoff"));
    o.smooth_width = unit_value(fc.at("smooth_width"));
    if (ft == "ShiftedCosine") {
      o.cutoff_type = 0;
    } else if (ft == "RadialScaling") {
      o.cutoff_type = 1;
      o.rs_rate = unit_value(fc.at("rate"));
      if (o.rs_rate < 0) throw RuntimeError("RadialScaling's rate should be positive");
      o.rs_exponent = unit_value(fc.at("exponent"));
      o.rs_scale = unit_value(fc.at("scale"));
    } else {
      throw InvalidArgument("Requested cutoff function type '" + ft +
                            "' has not been implemented.  Must be one of: 'ShiftedCosine' or "
                            "'RadialScaling'.");
    }
  } catch (const std::invalid_argument& e) {
    throw InvalidArgument(e.what());
  }
  if (o.max_radial < 1 || o.max_radial > 24) throw InvalidArgument("max_radial must be in [1, 24]");
  if (o.max_angular < 0 || o.max_angular > 16)
    throw InvalidArgument("max_angular must be in [0, 16]");
  if (!(o.cutoff > 0)) throw InvalidArgument("cutoff must be positive");
  if (!(o.gaussian_sigma > 0)) throw InvalidArgument("gaussian_sigma must be positive");
  if (o.optimization == 1 && !(o.spline_accuracy > 0))
    throw InvalidArgument("Error bound must be > 0");
  // math/gauss_legendre.cc:33-37 (the DVR basis integrates on max_radial
  // Gauss-Legendre points)
  if (o.radial_basis == 1 && o.max_radial < 2)
    throw InvalidArgument("there should be at least 2 integration points but: order_n < 2 == true");
  return o;
}

}  // namespace rgpu
