"""Host-side checks of the C-ABI library and the Python layer (CPU only).

No compute call is made: the library is loaded, every function declared in
include/rascal_mi355x.h must be exported, and hypers are validated before any
device is touched (bad hypers are argument errors on any host).
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT, soap_hypers

HEADER = os.path.join(ROOT, "include", "rascal_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(rgpu_\w+)\s*\(", text, re.M)))


def load():
    from librascal_amd import _capi
    return _capi.lib()


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("rgpu_context_create", "rgpu_set_structures", "rgpu_compute", "rgpu_get_features",
                 "rgpu_features_backward", "rgpu_set_sparse_points", "rgpu_gap_knm",
                 "rgpu_gap_predict", "rgpu_get_neighbours", "rgpu_set_stream"):
        assert must in names
    # every declaration cites the reference interface it replaces (file:line)
    text = open(HEADER).read()
    assert len(re.findall(r"\w+\.(?:hh|cc|py):\d+", text)) >= len(names) // 2


def test_library_exports_every_declared_symbol():
    lib = load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_has_gfx950_code_object():
    from librascal_amd import build
    blob = open(build.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def _create(hp):
    lib = load()
    lib.rgpu_context_create.restype = ctypes.c_int
    lib.rgpu_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    rc = lib.rgpu_context_create(json.dumps(hp).encode() if not isinstance(hp, bytes) else hp,
                                 ctypes.byref(out))
    return rc, lib.rgpu_last_error().decode(), out


@pytest.mark.parametrize("mutate,needle", [
    (lambda h: h.update(soap_type="BiSpectrum"), "BiSpectrum"),
    (lambda h: h["radial_contribution"].update(type="Chebyshev"), "Chebyshev"),
    (lambda h: h.update(max_radial=1) or h["radial_contribution"].update(type="DVR"),
     "integration points"),
    (lambda h: h["radial_contribution"].update(
        optimization=dict(Spline=dict(accuracy=1e-8),
                          RadialDimReduction=dict(projection_matrices={}))), "Projection matrices"),
    (lambda h: h["cutoff_function"].update(type="Step"), "Step"),
    (lambda h: h["gaussian_density"].update(type="PieceWise"), "PieceWise"),
    (lambda h: h.update(expansion_by_species_method="magic"), "magic"),
    (lambda h: h["radial_contribution"].update(optimization=dict(Spline=dict())), "accuracy"),
    (lambda h: h.update(max_radial=0), "max_radial"),
])
def test_bad_hypers_are_argument_errors(mutate, needle):
    hp = soap_hypers()
    mutate(hp)
    rc, msg, out = _create(hp)
    assert rc == 1, (rc, msg)  # RGPU_ERR_INVALID, the reference's std::logic_error
    assert needle in msg
    assert not out.value


def test_malformed_json_is_an_argument_error():
    rc, msg, _ = _create(b"{not json")
    assert rc == 1 and msg


def test_valid_hypers_without_device_fail_loudly():
    """Valid hypers pass validation; with no HIP device the call fails with a
    HIP error rather than falling back to any CPU path."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    rc, msg, out = _create(soap_hypers())
    assert rc == 3 and "HIP device" in msg and not out.value


class _RecordingContext:
    """Stands in for the device context: records the hypers JSON it gets."""
    seen = []

    def __init__(self, hypers_json):
        self.seen.append(json.loads(hypers_json))

    def set_precision(self, p):
        pass


def test_python_hypers_mirror_reference_defaults(monkeypatch):
    """spherical_invariants.py:209-300: keyword arguments -> hypers JSON,
    with the Python default optimisation (Spline, accuracy 1e-8)."""
    from librascal_amd import representations as R
    monkeypatch.setattr(R._capi, "Context", _RecordingContext)
    _RecordingContext.seen.clear()
    rep = R.SphericalInvariants(interaction_cutoff=5.0, cutoff_smooth_width=0.5, max_radial=8,
                                max_angular=6, gaussian_sigma_type="Constant",
                                gaussian_sigma_constant=0.5)
    hp = _RecordingContext.seen[-1]
    assert hp == rep.hypers
    assert hp["soap_type"] == "PowerSpectrum" and hp["normalize"] is True
    assert hp["cutoff_function"]["type"] == "ShiftedCosine"
    assert hp["cutoff_function"]["cutoff"]["value"] == 5.0
    assert hp["cutoff_function"]["smooth_width"]["value"] == 0.5
    assert hp["gaussian_density"] == dict(type="Constant",
                                          gaussian_sigma=dict(value=0.5, unit="AA"))
    assert hp["radial_contribution"]["type"] == "GTO"
    assert hp["radial_contribution"]["optimization"]["Spline"]["accuracy"] == 1e-8
    assert hp["expansion_by_species_method"] == "environment wise"
    assert rep.get_num_coefficients(2) == 3 * 8 * 8 * 7
    rs = R.SphericalInvariants(interaction_cutoff=4.0, cutoff_smooth_width=0.5, max_radial=4,
                               max_angular=3, gaussian_sigma_type="Constant",
                               cutoff_function_type="RadialScaling",
                               cutoff_function_parameters=dict(rate=1.0, scale=2.0, exponent=3),
                               optimization=dict(Spline=dict(accuracy=1e-6)))
    fc = rs.hypers["cutoff_function"]
    assert fc["type"] == "RadialScaling" and fc["rate"]["value"] == 1.0
    assert fc["scale"]["value"] == 2.0 and fc["exponent"]["value"] == 3
    with pytest.raises(ValueError):
        R.SphericalInvariants(interaction_cutoff=4.0, cutoff_smooth_width=0.5, max_radial=4,
                              max_angular=3, gaussian_sigma_type="Constant",
                              soap_type="BiSpectrum")


def test_pack_and_synthetic_boxes():
    from librascal_amd.synthetic import pack, periodic_box, molecules, diamond_cells
    sts = periodic_box(500, 0.05, (1, 8), d_min=1.0, seed=3)
    off, pos, num, cells, pbc = pack(sts)
    assert off.tolist() == [0, 500] and pos.shape == (500, 3)
    side = cells[0].reshape(3, 3)[0, 0]
    assert abs(500 / side ** 3 - 0.05) < 1e-12
    # minimum-image separation respects d_min
    d = pos[:, None, :] - pos[None, :, :]
    d -= side * np.round(d / side)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(500) * 1e9
    assert r.min() >= 1.0
    assert sorted(set(num.tolist())) == [1, 8] and abs((num == 1).sum() - 250) <= 1
    mols = molecules(n_mol=6, seed=5)
    for m in mols:
        assert 8 <= len(m["numbers"]) <= 64 and not any(m["pbc"])
    dc = diamond_cells(n_cells=2, seed=1)
    assert len(dc) == 2 and len(dc[0]["numbers"]) == 64
