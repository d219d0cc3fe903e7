"""SphericalInvariants / SphericalExpansion with the reference's Python API.

Constructor arguments and the JSON hypers they produce follow
bindings/rascal/representations/spherical_invariants.py:209-342,
spherical_expansion.py and representations/base.py:28-101, so hypers dicts
are interchangeable with the reference.  ``transform(frames)`` computes the
representation on the MI355X through the C ABI and returns an AtomsList.
"""
import json
from copy import deepcopy

import numpy as np

from . import _capi
from .neighbourlist import AtomsList

_SUPPORTED_OPTIMIZATION = ("Spline", "RadialDimReduction")


def cutoff_function_dict_switch(cutoff_function_type, **kwargs):
    """Cutoff-function part of the hypers (representations/base.py:40-64)."""
    if cutoff_function_type == "ShiftedCosine":
        return dict(type=cutoff_function_type,
                    cutoff=dict(value=kwargs["interaction_cutoff"], unit="AA"),
                    smooth_width=dict(value=kwargs["cutoff_smooth_width"], unit="AA"))
    if cutoff_function_type == "RadialScaling":
        return dict(type=cutoff_function_type,
                    cutoff=dict(value=kwargs["interaction_cutoff"], unit="AA"),
                    smooth_width=dict(value=kwargs["cutoff_smooth_width"], unit="AA"),
                    rate=dict(value=kwargs["rate"], unit="AA"),
                    scale=dict(value=kwargs["scale"], unit="AA"),
                    exponent=dict(value=kwargs["exponent"], unit="AA"))
    raise NotImplementedError(
        "cutoff_function: " + cutoff_function_type + " has not been implemented.")


def check_optimization_for_spherical_representations(optimization, optimization_args=None):
    """Default optimisation is Spline(accuracy=1e-8) (representations/base.py:67-101)."""
    if optimization_args is not None:
        optimization = dict(Spline=dict(accuracy=optimization_args["accuracy"]))
    if optimization is None:
        optimization = dict(Spline=dict(accuracy=1e-8))
    optimization = deepcopy(optimization)
    if "RadialDimReduction" in optimization and "Spline" not in optimization:
        raise ValueError("Optimization key `RadialDimReduction` can only be used with `Spline`")
    return optimization


class _SphericalBase:
    name = None

    def _finish(self, hypers):
        self.hypers = hypers
        rc = hypers["cutoff_function"]["cutoff"]["value"]
        self.nl_options = [
            dict(name="centers", args=[]),
            dict(name="neighbourlist", args=dict(cutoff=rc)),
            dict(name="centercontribution", args=dict()),
            dict(name="strict", args=dict(cutoff=rc)),
        ]
        # validates the hypers on the device side (raises like the reference)
        self._context = _capi.Context(json.dumps(self.hypers))
        self.precision = _capi.FP64
        # the batch whose structures are loaded in the context (an AtomsList,
        # or a species tuple for an empty batch that only fixes the layout)
        self._loaded = None
        self._sparse_token = None

    @property
    def cutoff(self):
        return self.hypers["cutoff_function"]["cutoff"]["value"]

    def set_precision(self, precision):
        """'fp64' (default, parity path) or 'mixed' (FP32 expansion and
        invariants, BF16 hi+lo kernel GEMM)."""
        self.precision = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[precision]
        self._context.set_precision(self.precision)
        if self._loaded is not None and not isinstance(self._loaded, tuple):
            self._context.compute()
        return self

    def _load(self, atoms_list):
        """Make atoms_list the batch held by the context (recomputing its
        representation if another batch was loaded since)."""
        if self._loaded is not atoms_list:
            atoms_list._upload(self)
            self._context.compute()
            self._loaded = atoms_list
            self._sparse_token = None

    def _load_species(self, species):
        """Empty batch with the given species list (feature layout only)."""
        sp = tuple(sorted(set(int(z) for z in species)))
        if self._loaded != sp:
            self._context.set_structures(np.zeros(1, np.int64), np.zeros((0, 3)),
                                         np.zeros(0, np.int32), np.zeros((0, 9)),
                                         np.zeros((0, 3), np.int32), list(sp))
            self._loaded = sp
            self._sparse_token = None

    def get_name(self):
        return json.dumps(self.hypers, sort_keys=True)

    def transform(self, frames, species=None):
        """Compute the representation of a list of structures (dicts with
        positions/numbers/cell/pbc, ase.Atoms-like objects, or an AtomsList)."""
        if not isinstance(frames, AtomsList):
            frames = AtomsList(frames, self.nl_options)
        frames.compute(self, species=species)
        return frames

    def to_dict(self):
        return dict(class_name=type(self).__name__, init_params=deepcopy(self._init_params))


class SphericalInvariants(_SphericalBase):
    """SOAP power spectrum (soap_type 'PowerSpectrum'); arguments as the
    reference (spherical_invariants.py:209-232)."""

    name = "sphericalinvariants"

    def __init__(self, interaction_cutoff, cutoff_smooth_width, max_radial, max_angular,
                 gaussian_sigma_type, gaussian_sigma_constant=0.3,
                 cutoff_function_type="ShiftedCosine", soap_type="PowerSpectrum",
                 inversion_symmetry=True, radial_basis="GTO", normalize=True,
                 optimization=None, optimization_args=None,
                 expansion_by_species_method="environment wise", global_species=None,
                 compute_gradients=False, cutoff_function_parameters=dict(),
                 coefficient_subselection=None):
        self._init_params = dict(
            interaction_cutoff=interaction_cutoff, cutoff_smooth_width=cutoff_smooth_width,
            max_radial=max_radial, max_angular=max_angular,
            gaussian_sigma_type=gaussian_sigma_type,
            gaussian_sigma_constant=gaussian_sigma_constant,
            cutoff_function_type=cutoff_function_type, soap_type=soap_type,
            inversion_symmetry=inversion_symmetry, radial_basis=radial_basis,
            normalize=normalize, optimization=optimization,
            expansion_by_species_method=expansion_by_species_method,
            global_species=global_species, compute_gradients=compute_gradients,
            cutoff_function_parameters=deepcopy(cutoff_function_parameters),
            coefficient_subselection=coefficient_subselection)
        if global_species is None:
            global_species = []
        elif not isinstance(global_species, list):
            global_species = list(global_species)
        if soap_type != "PowerSpectrum":
            raise ValueError("soap_type '%s' is not provided by this library" % soap_type)
        hypers = dict(max_radial=max_radial, max_angular=max_angular, soap_type=soap_type,
                      normalize=normalize, inversion_symmetry=inversion_symmetry,
                      expansion_by_species_method=expansion_by_species_method,
                      global_species=global_species, compute_gradients=compute_gradients)
        if coefficient_subselection is not None:
            hypers["coefficient_subselection"] = coefficient_subselection
        cfp = deepcopy(cutoff_function_parameters)
        cfp.update(interaction_cutoff=interaction_cutoff, cutoff_smooth_width=cutoff_smooth_width)
        hypers["cutoff_function"] = cutoff_function_dict_switch(cutoff_function_type, **cfp)
        hypers["gaussian_density"] = dict(type=gaussian_sigma_type,
                                          gaussian_sigma=dict(value=gaussian_sigma_constant,
                                                              unit="AA"))
        opt = check_optimization_for_spherical_representations(optimization, optimization_args)
        hypers["radial_contribution"] = dict(type=radial_basis, optimization=opt)
        self._finish(hypers)

    def get_num_coefficients(self, n_species=1):
        N = self.hypers["max_radial"]
        L1 = self.hypers["max_angular"] + 1
        return n_species * (n_species + 1) // 2 * N * N * L1

    def get_keys(self, species):
        sp = sorted(species)
        return [(a, b) for i, a in enumerate(sp) for b in sp[i:]]


class SphericalExpansion(_SphericalBase):
    """Density expansion coefficients c_nlm (spherical_expansion.py)."""

    name = "sphericalexpansion"

    def __init__(self, interaction_cutoff, cutoff_smooth_width, max_radial, max_angular,
                 gaussian_sigma_type, gaussian_sigma_constant=0.3,
                 cutoff_function_type="ShiftedCosine", radial_basis="GTO", optimization=None,
                 optimization_args=None, expansion_by_species_method="environment wise",
                 global_species=None, compute_gradients=False,
                 cutoff_function_parameters=dict()):
        self._init_params = dict(
            interaction_cutoff=interaction_cutoff, cutoff_smooth_width=cutoff_smooth_width,
            max_radial=max_radial, max_angular=max_angular,
            gaussian_sigma_type=gaussian_sigma_type,
            gaussian_sigma_constant=gaussian_sigma_constant,
            cutoff_function_type=cutoff_function_type, radial_basis=radial_basis,
            optimization=optimization, expansion_by_species_method=expansion_by_species_method,
            global_species=global_species, compute_gradients=compute_gradients,
            cutoff_function_parameters=deepcopy(cutoff_function_parameters))
        if global_species is None:
            global_species = []
        hypers = dict(max_radial=max_radial, max_angular=max_angular,
                      expansion_by_species_method=expansion_by_species_method,
                      global_species=list(global_species), compute_gradients=compute_gradients)
        cfp = deepcopy(cutoff_function_parameters)
        cfp.update(interaction_cutoff=interaction_cutoff, cutoff_smooth_width=cutoff_smooth_width)
        hypers["cutoff_function"] = cutoff_function_dict_switch(cutoff_function_type, **cfp)
        hypers["gaussian_density"] = dict(type=gaussian_sigma_type,
                                          gaussian_sigma=dict(value=gaussian_sigma_constant,
                                                              unit="AA"))
        opt = check_optimization_for_spherical_representations(optimization, optimization_args)
        hypers["radial_contribution"] = dict(type=radial_basis, optimization=opt)
        self._finish(hypers)

    def get_num_coefficients(self, n_species=1):
        N = self.hypers["max_radial"]
        L1 = self.hypers["max_angular"] + 1
        return n_species * N * L1 * L1

    def get_keys(self, species):
        return [(a,) for a in sorted(species)]


def from_hypers(hypers):
    """Calculator object for a raw reference hypers dict."""
    obj = object.__new__(SphericalInvariants if "soap_type" in hypers else SphericalExpansion)
    obj._init_params = dict(hypers=deepcopy(hypers))
    obj._finish(deepcopy(hypers))
    return obj


__all__ = ["SphericalInvariants", "SphericalExpansion", "from_hypers", "np"]
