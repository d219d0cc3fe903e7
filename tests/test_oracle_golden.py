"""The FP64 oracle against librascal's own golden data (CPU).

Each test names the reference test and tolerance it mirrors.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, golden_structures, rel_err, wrap_cubic_periodic

DBL_FTOL = 100 * np.finfo(float).eps


def test_hyp1f1_vs_mpmath():
    """tests/test_math_hyp1f1.cc:42-88: 4800 mpmath (200 digits) values and
    derivatives, relative error <= 15 DBL_FTOL."""
    d = np.load(os.path.join(GOLDEN, "hyp1f1_reference.npz"))["data"]
    worst_v = worst_d = 0.0
    for a, b, z, val, der in d:
        v = O.hyp1f1(a, b, z, False, 200, 1e-15)
        dv = O.hyp1f1(a, b, z, True, 200, 1e-15)
        worst_v = max(worst_v, abs((val - v) / val))
        worst_d = max(worst_d, abs((der - dv) / der))
    assert worst_v <= 15 * DBL_FTOL
    assert worst_d <= 15 * DBL_FTOL


def test_spherical_harmonics_reference():
    """tests/test_math_spherical_harmonics.cc:50-140: harmonics (L2 error
    <= 2 DBL_FTOL) and associated Legendre functions (<= 10 DBL_FTOL), l <= 30."""
    d = np.load(os.path.join(GOLDEN, "spherical_harmonics_reference.npz"))
    for L, u, h, alps in zip(d["max_angular"], d["directions"], d["harmonics"], d["alps"]):
        y, _, alp = O.ylm(int(L), u)
        assert np.linalg.norm(y - h) <= 2 * DBL_FTOL
        assert np.linalg.norm(alp[:, :L + 1] - np.asarray(alps).reshape(L + 1, L + 1)) <= 10 * DBL_FTOL


@pytest.mark.parametrize("L", [1, 4, 9])
def test_ylm_gradient_two_term_identity(L):
    """The adjoint's Y_lm gradients: r^l Y_lm is a harmonic polynomial, so its
    gradient is one of degree l-1 and, on the unit sphere, dY_c[lm] =
    sum_m' a^c_{lm,m'} Y_{l-1,m'} - l r_c Y_lm with at most two nonzero
    a^c_{lm,m'} per (lm, c).  The coefficients are projected with Gauss-Legendre
    (L + 2 points in cos theta) times 2L + 3 uniform points in phi, exact for
    these products, from the oracle's harmonics and gradients; the identity
    must then reproduce the gradients at random directions."""
    nt, nph = L + 2, 2 * L + 3
    ct, wt = np.polynomial.legendre.leggauss(nt)
    ph = 2 * np.pi * (np.arange(nph) + 0.5) / nph
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([np.outer(st, np.cos(ph)).ravel(), np.outer(st, np.sin(ph)).ravel(),
                     np.repeat(ct, nph)], axis=1)
    wq = np.repeat(wt, nph) * 2 * np.pi / nph
    Y, dY = zip(*[O.ylm(L, d, derivatives=True)[:2] for d in dirs])
    Y, dY = np.array(Y), np.array(dY)
    test_dirs = np.random.default_rng(0).normal(size=(50, 3))
    test_dirs /= np.linalg.norm(test_dirs, axis=1)[:, None]
    Yt, dYt = zip(*[O.ylm(L, d, derivatives=True)[:2] for d in test_dirs])
    Yt, dYt = np.array(Yt), np.array(dYt)
    for l in range(1, L + 1):
        lo = (l - 1) ** 2
        for lm in range(l * l, (l + 1) ** 2):
            for c in range(3):
                f = dY[:, c, lm] + l * dirs[:, c] * Y[:, lm]
                a = (wq * f) @ Y[:, lo:l * l] / (wq @ Y[:, lo:l * l] ** 2)
                a[np.abs(a) < 1e-9] = 0.0
                assert np.count_nonzero(a) <= 2
                v = Yt[:, lo:l * l] @ a - l * test_dirs[:, c] * Yt[:, lm]
                assert np.max(np.abs(v - dYt[:, c, lm])) < 1e-12 * max(1.0, np.abs(dYt[:, c, lm]).max())


def _expansion_rows():
    rows = json.load(open(os.path.join(GOLDEN, "spherical_expansion_reference.json")))
    return [r for r in rows]


_STRUCT_OF = {"CaCrP2O7_mvc-11955_symmetrized.json": "CaCrP2O7",
              "small_molecule.json": "small_molecule"}


@pytest.mark.parametrize("row", _expansion_rows(), ids=lambda r: r["name"])
def test_spherical_expansion_reference(row):
    """tests/test_calculator.cc:460-493: dense spherical-expansion features of
    CaCrP2O7 and a small molecule for GTO hypers (direct Hyp1f1 and Spline
    1e-8; ShiftedCosine and RadialScaling).  The reference bar is 3e-6 with
    relative_error(eps 1e-14); the oracle is held to 1e-8."""
    feats = np.load(os.path.join(GOLDEN, "spherical_expansion_reference.npz"))[row["name"]]
    st = golden_structures()[_STRUCT_OF[row["structure"]]]
    calc = O.Calculator(row["hypers"])
    b = O.BatchData(st)
    c, _ = calc.expansion(b)
    X = c.reshape(c.shape[0], -1)
    assert X.shape == feats.shape
    assert rel_err(feats, X).max() < 1e-8


def test_power_spectrum_via_kernel_reference():
    """tests/test_kernels.cc:62-101: cosine kernels (X X^T)^zeta of the
    normalised power spectrum (n 6, l 6, rc 3.5, direct Hyp1f1) of the first
    5 structures of small_molecules-20, Structure (mean) and Atom targets,
    relative error <= 1e-10."""
    meta = json.load(open(os.path.join(GOLDEN, "kernel_reference.json")))
    mats = np.load(os.path.join(GOLDEN, "kernel_reference.npz"))
    sts = golden_structures()["small_molecules_20"][meta["start"]:meta["start"] + meta["length"]]
    b = O.BatchData(sts)
    done = 0
    for row in meta["rows"]:
        hp = dict(row["hypers"])
        hp["radial_contribution"] = dict(type="GTO", optimization={})  # "None" in the fixture
        calc = O.Calculator(hp)
        X, _ = calc.power_spectrum(b)
        z = row["kernel"]["zeta"]
        K = (X @ X.T) ** z
        if row["kernel"]["target_type"] == "Structure":
            off = b.atom_offsets
            K = np.array([[K[off[i]:off[i + 1], off[j]:off[j + 1]].mean() for j in range(b.n_struct)]
                          for i in range(b.n_struct)])
        ref = mats[row["name"]]
        assert np.max(np.abs(K - ref) / np.abs(ref)) < 1e-10
        done += 1
    assert done == 6


def test_sparse_points_of_gap_model_fixture():
    """reference_data/tests_only/simple_gap_model.json holds 30 sparse points
    (power spectra with RadialScaling rc 7.5, n 4, l 2, direct Hyp1f1) taken
    from environments of the first 100 methane dimers; every one must match
    an oracle environment of the same species."""
    gm = json.load(open(os.path.join(GOLDEN, "simple_gap_model.json")))
    xt = gm["init_params"]["X_train"]
    ip = xt["init_params"]["representation"]["init_params"]
    sp = xt["data"]["sparse_points"]
    hp = dict(max_radial=ip["max_radial"], max_angular=ip["max_angular"], soap_type="PowerSpectrum",
              normalize=ip["normalize"],
              cutoff_function=dict(type="RadialScaling",
                                   cutoff=dict(value=ip["interaction_cutoff"]),
                                   smooth_width=dict(value=ip["cutoff_smooth_width"]),
                                   rate=dict(value=ip["cutoff_function_parameters"]["rate"]),
                                   scale=dict(value=ip["cutoff_function_parameters"]["scale"]),
                                   exponent=dict(value=ip["cutoff_function_parameters"]["exponent"])),
              gaussian_density=dict(type="Constant",
                                    gaussian_sigma=dict(value=ip["gaussian_sigma_constant"])),
              radial_contribution=dict(type="GTO", optimization=ip["optimization"]))
    sts = [wrap_cubic_periodic(s) for s in golden_structures()["methane_dimers"]]
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=[1, 6])
    X, _ = calc.power_spectrum(b)
    inner = sp["inner_size"]
    keys = [tuple(k) for k in sp["keys"]]
    assert keys == [(1, 1), (1, 6), (6, 6)] and X.shape[1] == 3 * inner
    n_checked = 0
    for sp_type, by_key in sp["values"]:
        counts = dict((c[0], c[1]) for c in sp["counters"])[sp_type]
        T = np.zeros((counts, 3 * inner))
        for key, vals in by_key:
            k = keys.index(tuple(key))
            T[:, k * inner:(k + 1) * inner] = np.asarray(vals).reshape(counts, inner)
        cand = X[b.types == sp_type]
        for t in T:
            d = np.min(np.linalg.norm(cand - t[None, :], axis=1))
            assert d < 1e-8
            n_checked += 1
    assert n_checked == 30
