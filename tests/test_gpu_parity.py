"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Bars: neighbour lists bit-identical (same pairs, same order, same vectors);
FP64 GPU path within 1e-10 relative, measured per block: for the expansion
per (centre, species, l) block, for the power spectrum per (centre, key, l)
block, for radial functions per (n, l) over the sampled distances (max |diff|
/ max |ref| of the block).  Element-wise relative errors are not used for
entries that are many orders of magnitude below their block (they come from
cancellation and carry only rounding noise of the large entries).  The mixed
path is held to the tolerances stated per test.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import (expansion_rel, golden_structures, oracle_backward, ps_rel, radial_rel,
                      random_sparse, rel_err, soap_hypers, wrap_cubic_periodic)
from librascal_amd import _capi
from librascal_amd.synthetic import diamond_cells, molecules, pack, periodic_box, triclinic_cells

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def make_ctx(hp, structures, precision=_capi.FP64, species=None):
    ctx = _capi.Context(json.dumps(hp))
    ctx.set_precision(precision)
    off, pos, num, cells, pbc = pack(structures)
    sp = sorted(set(num.tolist()) | set(species or []))
    ctx.set_structures(off, pos, num, cells, pbc, sp)
    return ctx, sp


def structure_sets():
    g = golden_structures()
    sets = {
        "CaCrP2O7_rc3": (g["CaCrP2O7"], 3.0),
        "small_molecule_rc3": (g["small_molecule"], 3.0),
        "diamond_small_cell_rc5": (g["diamond_cubic_distorted"], 5.0),
        "methane_dimers_rc4": ([wrap_cubic_periodic(s) for s in g["methane_dimers"][:6]], 4.0),
        "SiC_rc4": (g["SiC_moissanite_supercell"], 4.0),
        "diamond_jitter_rc5": (diamond_cells(n_cells=2), 5.0),
        "molecules_rc4": (molecules(n_mol=6), 4.0),
        "triclinic_rc6": (triclinic_cells(n_cells=3, n_max=80), 6.0),
    }
    return sets


@pytest.mark.parametrize("name", list(structure_sets().keys()))
def test_neighbour_list_bit_exact(name):
    sts, rc = structure_sets()[name]
    ctx, _ = make_ctx(soap_hypers(rc=rc, N=2, L=1), sts)
    off, pj, vec = ctx.neighbours()
    a0 = 0
    for st in sts:
        n = len(st["numbers"])
        ro, rj, rv, _ = O.neighbours(st["positions"], st["cell"], st["pbc"], rc)
        go = off[a0:a0 + n + 1] - off[a0]
        assert np.array_equal(go, ro), name
        seg = slice(off[a0], off[a0 + n])
        assert np.array_equal(pj[seg] - a0, rj), name
        assert np.array_equal(vec[seg], rv), name
        a0 += n


@pytest.mark.parametrize("m", [24, 42])
def test_neighbour_list_bit_exact_large_box(m):
    """Jittered simple cubic lattices, periodic: 13824 atoms (the ghost and
    pair offsets scanned by one workgroup in two chunks of 8192) and 74088
    atoms (more atoms than one workgroup scans: the multi-block scan; its
    ~12000 boxes take the two-chunk single-workgroup scan)."""
    a = 1.5
    rng = np.random.default_rng(11)
    grid = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (grid + 0.5 + rng.uniform(-0.2, 0.2, size=grid.shape)) * a
    st = dict(positions=pos, numbers=np.where(rng.random(len(pos)) < 0.5, 1, 8).astype(np.int32),
              cell=np.eye(3) * m * a, pbc=np.ones(3, np.int32))
    ctx, _ = make_ctx(soap_hypers(rc=3.0, N=2, L=1), [st])
    off, pj, vec = ctx.neighbours()
    ro, rj, rv, _ = O.neighbours(st["positions"], st["cell"], st["pbc"], 3.0)
    assert len(pos) > (65536 if m == 42 else 8192)
    assert np.array_equal(off, ro)
    assert np.array_equal(pj, rj)
    assert np.array_equal(vec, rv)


def test_ylm_parity_fp64_and_fp32():
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    dirs[:4] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0.0, 0.05, 0.99874921777]]
    for L in (3, 6, 9):
        ctx, _ = make_ctx(soap_hypers(rc=3.0, N=2, L=L), golden_structures()["small_molecule"])
        Y, dY = ctx.eval_ylm(dirs, derivative=True)
        for k, d in enumerate(dirs):
            y, dy, _ = O.ylm(L, d, derivatives=True)
            assert np.max(np.abs(Y[k] - y)) < 1e-13
            assert np.max(np.abs(dY[k] - dy)) < 1e-12
        ctx.set_precision(_capi.MIXED)
        Y32, dY32 = ctx.eval_ylm(dirs, derivative=True)
        # FP32 recursion: a few tens of ulp on values of order one
        assert np.max(np.abs(Y32 - Y)) < 1e-5
        assert np.max(np.abs(dY32 - dY)) < 1e-4


@pytest.mark.parametrize("spline", [None, 1e-8])
@pytest.mark.parametrize("N,L,sigma,rc", [(4, 3, 0.5, 3.0), (8, 6, 0.5, 5.0), (12, 9, 0.5, 5.0),
                                          (6, 2, 0.3, 4.0)])
def test_radial_parity(spline, N, L, sigma, rc):
    hp = soap_hypers(rc=rc, N=N, L=L, sigma=sigma, spline=spline)
    ctx, _ = make_ctx(hp, golden_structures()["small_molecule"])
    calc = O.Calculator(hp)
    if spline is not None:
        assert ctx.info("spline_grid") == calc.radial_info()["grid"]
    r = np.linspace(0.05, rc, 137)
    R, dR = ctx.eval_radial(r, derivative=True)
    ref = [calc.radial(x, derivative=True) for x in r]
    assert radial_rel(np.array([a for a, _ in ref]), R) < FP64_TOL
    assert radial_rel(np.array([b for _, b in ref]), dR) < FP64_TOL


EXP_CASES = [
    ("CaCrP2O7", 3.0, 0.5, 4, 3, None),
    ("CaCrP2O7", 3.0, 0.2, 4, 3, 1e-8),
    ("small_molecule", 3.0, 0.3, 6, 4, None),
    ("diamond_cubic_distorted", 5.0, 0.5, 8, 6, 1e-8),
    ("SiC_moissanite_supercell", 4.0, 0.4, 5, 5, None),
]


@pytest.mark.parametrize("case", EXP_CASES)
def test_expansion_and_power_spectrum_fp64(case):
    name, rc, sigma, N, L, spline = case
    sts = golden_structures()[name]
    hp = soap_hypers(rc=rc, sigma=sigma, N=N, L=L, spline=spline)
    ctx, sp = make_ctx(hp, sts)
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=sp)
    c_ref, mask_ref = calc.expansion(b)
    c, mask = ctx.expansion()
    assert np.array_equal(mask, mask_ref)
    assert expansion_rel(c_ref, c, L) < FP64_TOL
    p_ref, inv_ref = calc.power_spectrum(b)
    X, inv = ctx.features()
    assert ps_rel(p_ref, X, len(sp), N, L) < FP64_TOL
    assert rel_err(inv_ref, inv).max() < FP64_TOL


@pytest.mark.parametrize("case", [("CaCrP2O7", 3.0, 0.5, 4, 3, None),
                                  ("diamond_cubic_distorted", 5.0, 0.5, 4, 4, 1e-8),
                                  ("small_molecule", 3.5, 0.4, 5, 3, 1e-8)])
def test_features_backward_fp64(case):
    name, rc, sigma, N, L, spline = case
    sts = golden_structures()[name]
    hp = soap_hypers(rc=rc, sigma=sigma, N=N, L=L, spline=spline)
    ctx, sp = make_ctx(hp, sts)
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=sp)
    rng = np.random.default_rng(7)
    G = rng.normal(size=(b.n_atoms, calc.D(len(sp))))
    ref_g, ref_s = oracle_backward(calc, b, G)
    g, s = ctx.features_backward(G)
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() < FP64_TOL * scale
    assert np.abs(s - ref_s).max() < FP64_TOL * np.abs(ref_s).max()


@pytest.mark.parametrize("zeta", [1, 2, 4])
def test_gap_predict_fp64(zeta):
    sts = golden_structures()["CaCrP2O7"] + golden_structures()["SiC_moissanite_supercell"]
    hp = soap_hypers(rc=3.5, sigma=0.5, N=4, L=3, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=sp)
    rng = np.random.default_rng(11 + zeta)
    spec, vals, w = random_sparse(calc, b, 5, rng, zeta)
    ctx.set_sparse_points(zeta, spec, vals, w)
    ss = O.SparseSet(spec, vals, zeta)
    e_ref, g_ref, s_ref = O.gap_predict(calc, b, ss, w)
    e, g, s = ctx.gap_predict()
    assert rel_err(e_ref, e).max() < FP64_TOL
    assert np.abs(g - g_ref).max() < FP64_TOL * np.abs(g_ref).max()
    assert np.abs(s - s_ref).max() < FP64_TOL * np.abs(s_ref).max()
    for target in ("Structure", "Atom"):
        k_ref = O.gap_knm(calc, b, ss, target)
        # the oracle orders columns species-sorted; map back to input order
        k_ref_in = np.zeros_like(k_ref)
        k_ref_in[:, ss.order] = k_ref
        k = ctx.gap_knm(target)
        assert rel_err(k_ref_in, k).max() < FP64_TOL


def test_mixed_precision_power_spectrum_and_forces():
    sts = periodic_box(256, 0.06, (1, 6, 7, 8), seed=5)
    hp = soap_hypers(rc=5.0, sigma=0.5, N=8, L=6, spline=1e-8)
    calc = O.Calculator(hp)
    ctx, sp = make_ctx(hp, sts)
    b = O.BatchData(sts, species=sp)
    rng = np.random.default_rng(5)
    spec, vals, w = random_sparse(calc, b, 16, rng, 4)
    ctx.set_sparse_points(4, spec, vals, w)
    e64, g64, s64 = ctx.gap_predict()
    X64, _ = ctx.features()
    ctx.set_precision(_capi.MIXED)
    X32, _ = ctx.features()
    e32, g32, s32 = ctx.gap_predict()
    assert np.abs(X32 - X64).max() < 1e-5
    assert abs(e32[0] - e64[0]) / abs(e64[0]) < 1e-4
    rms = np.sqrt(np.mean((g32 - g64) ** 2)) / np.sqrt(np.mean(g64 ** 2))
    assert rms < 1e-4


_GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
_STRUCT_OF = {"CaCrP2O7_mvc-11955_symmetrized.json": "CaCrP2O7",
              "small_molecule.json": "small_molecule"}


def _dvr_rows():
    return json.load(open(os.path.join(_GOLDEN_DIR, "spherical_expansion_dvr_reference.json")))


def _gto_rows():
    return json.load(open(os.path.join(_GOLDEN_DIR, "spherical_expansion_reference.json")))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("row", _gto_rows(), ids=lambda r: r["name"])
def test_gto_expansion_reference_on_gpu(row, precision):
    """tests/test_calculator.cc:460-493, the 32 GTO rows of the reference's
    spherical_expansion_reference.ubjson (CaCrP2O7 and a small molecule, rc
    2 and 3, sigma 0.2 and 0.5, ShiftedCosine and RadialScaling, direct and
    Spline 1e-8) straight against the GPU expansion: FP64 at 1e-8
    (relative_error as the reference's, which holds itself to 3e-6), FP32
    at 1e-5 of each row's largest coefficient."""
    feats = np.load(os.path.join(_GOLDEN_DIR, "spherical_expansion_reference.npz"))[row["name"]]
    sts = golden_structures()[_STRUCT_OF[row["structure"]]]
    ctx, _ = make_ctx(row["hypers"], sts, precision=_capi.FP64 if precision == "fp64" else _capi.MIXED)
    c, _ = ctx.expansion()
    X = c.reshape(c.shape[0], -1)
    assert X.shape == feats.shape
    if precision == "fp64":
        assert rel_err(feats, X).max() < 1e-8
    else:
        row_max = np.abs(feats).max(axis=1, keepdims=True)
        assert np.all(np.abs(X - feats) <= 1e-5 * row_max)


@pytest.mark.parametrize("row", _dvr_rows(), ids=lambda r: r["name"])
def test_dvr_expansion_reference(row):
    """tests/test_calculator.cc:460-493, the DVR rows of the reference's
    spherical_expansion_reference.ubjson (CaCrP2O7 and a small molecule, rc
    2 and 3, sigma 0.2 and 0.5, ShiftedCosine and RadialScaling, direct and
    Spline 1e-8): the FP64 GPU expansion against the reference's own
    features (the oracle has no DVR basis; the reference's bar is 3e-6 with
    relative_error, held here to 1e-8)."""
    feats = np.load(os.path.join(_GOLDEN_DIR, "spherical_expansion_dvr_reference.npz"))[row["name"]]
    sts = golden_structures()[_STRUCT_OF[row["structure"]]]
    ctx, _ = make_ctx(row["hypers"], sts)
    c, _ = ctx.expansion()
    X = c.reshape(c.shape[0], -1)
    assert X.shape == feats.shape
    assert rel_err(feats, X).max() < 1e-8


@pytest.mark.parametrize("spline,N", [(None, 6), (1e-8, 6), (None, 5)])
def test_dvr_gradients_vs_finite_differences(spline, N):
    """The adjoint with the DVR radial derivatives (Bessel-function gradient
    recursion): dL/dr for L = sum G . X of the normalised power spectrum
    against central differences (h = 1e-5 A) of the GPU features.  N = 5 has
    a quadrature node at the middle of the interval, whose weight the
    reference's routine forms with the previous root's derivative."""
    hp = soap_hypers(rc=3.5, sw=0.5, sigma=0.4, N=N, L=4, spline=spline)
    hp["radial_contribution"]["type"] = "DVR"
    sts = golden_structures()["small_molecule"]
    ctx, _ = make_ctx(hp, sts)
    X, _ = ctx.features()
    assert np.isfinite(X).all()
    G = np.random.default_rng(3).normal(size=X.shape)
    g, _ = ctx.features_backward(G)
    pos = pack(sts)[1]
    h = 1e-5
    for atom in range(min(4, len(pos))):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                p = pos.copy()
                p[atom, d] += sgn * h
                ctx.update_positions(p)
                vals.append(np.sum(G * ctx.features()[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[atom, d]) < 1e-6 * np.abs(g).max() + 1e-6 * abs(fd)
    ctx.update_positions(pos)


def _reference_gauss_legendre(r0, r1, n):
    """The reference's Gauss-Legendre loop (math/gauss_legendre.cc:30-73)
    as plain Python: Newton's iteration per root from z1 = 0 while
    |z - z1| > DBL_FTOL, the Legendre derivative pp carried over from the
    previous root when the loop does not run."""
    x, w = np.zeros(n), np.zeros(n)
    mid, half = 0.5 * (r0 + r1), 0.5 * (r1 - r0)
    pp = 0.0
    for i in range((n + 1) // 2):
        z, z1 = np.cos(np.pi * (i + 0.75) / (n + 0.5)), 0.0
        while abs(z - z1) > 100 * 2.220446049250313e-16:
            p1, p2 = 1.0, 0.0
            for j in range(n):
                p1, p2 = ((2 * j + 1) * z * p1 - j * p2) / (j + 1), p1
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1, z = z, z - p1 / pp
        x[i], x[n - 1 - i] = mid - half * z, mid + half * z
        w[i] = w[n - 1 - i] = 2.0 * half / ((1.0 - z * z) * pp * pp)
    return x, w


@pytest.mark.parametrize("N", [4, 5, 7])
def test_dvr_quadrature_follows_the_reference(N):
    """DVR centre term of an isolated atom: sqrt(w_n) x_n exp(-x_n^2 / 2
    sigma^2) times one common factor (calculator_spherical_expansion.hh
    :809-848), with the points and weights of the reference's Gauss-Legendre
    loop on [0, rc + 3 sigma].  For odd N that loop skips Newton's iteration
    at the middle root and keeps the previous root's derivative; the middle
    weight then differs from the textbook quadrature's, and the GPU basis
    follows the reference there too."""
    rc, sigma = 3.5, 0.4
    hp = soap_hypers(rc=rc, sw=0.5, sigma=sigma, N=N, L=2, spline=None, normalize=False)
    hp["radial_contribution"]["type"] = "DVR"
    st = dict(positions=np.array([[10.0, 10.0, 10.0]]), numbers=np.array([1]),
              cell=np.eye(3) * 20.0, pbc=np.zeros(3, int))
    ctx, _ = make_ctx(hp, [st])
    c, _ = ctx.expansion()
    c0 = c.reshape(N, -1)[:, 0]
    x, w = _reference_gauss_legendre(0.0, rc + 3 * sigma, N)
    ratio = c0 / (np.sqrt(w) * x * np.exp(-0.5 * x ** 2 / sigma ** 2))
    assert np.abs(ratio / ratio[0] - 1.0).max() < 1e-12
    zt, wt = np.polynomial.legendre.leggauss(N)
    wt = wt * 0.5 * (rc + 3 * sigma)
    if N % 2:
        assert abs(wt[N // 2] / w[N // 2] - 1.0) > 1e-2
    else:
        assert np.abs(wt / w - 1.0).max() < 1e-12


def _rdr_rows():
    d = json.load(open(os.path.join(_GOLDEN_DIR, "spherical_expansion_rdr_reference.json")))
    for r in d["rows"]:
        r["hypers"]["radial_contribution"]["optimization"]["RadialDimReduction"][
            "projection_matrices"] = d["projection_matrices"]
    return d["rows"]


@pytest.mark.parametrize("row", _rdr_rows(), ids=lambda r: r["name"])
def test_radial_dim_reduction_expansion_reference(row):
    """tests/test_calculator.cc:460-493, the RadialDimReduction rows of the
    reference's spherical_expansion_reference.ubjson (GTO and DVR, Spline
    1e-8, per-species projection matrices of every l): the FP64 GPU
    expansion against the reference's own features, relative_error < 1e-8
    (the reference's bar is 3e-6)."""
    feats = np.load(os.path.join(_GOLDEN_DIR, "spherical_expansion_rdr_reference.npz"))[row["name"]]
    sts = golden_structures()[_STRUCT_OF[row["structure"]]]
    ctx, _ = make_ctx(row["hypers"], sts)
    c, _ = ctx.expansion()
    X = c.reshape(c.shape[0], -1)
    assert X.shape == feats.shape
    assert rel_err(feats, X).max() < 1e-8


def test_radial_dim_reduction_missing_species_and_gradients():
    """A batch species without projection matrices is an error (the
    reference's runtime_error); with projections for every species, the
    adjoint (spline derivatives of each species' table) matches central
    differences of the GPU features."""
    hp = soap_hypers(rc=3.5, sw=0.5, sigma=0.4, N=4, L=3, spline=1e-8)
    rng = np.random.default_rng(7)
    sts = golden_structures()["small_molecule"]
    zs = sorted(set(pack(sts)[2].tolist()))
    proj = {str(z): rng.uniform(size=(4, 4, 4)).tolist() for z in zs}
    hp["radial_contribution"]["optimization"]["RadialDimReduction"] = dict(
        projection_matrices={k: v for k, v in proj.items() if k != str(zs[0])})
    with pytest.raises(_capi.RascalGPUError, match="missing projection matrices"):
        make_ctx(hp, sts)
    hp["radial_contribution"]["optimization"]["RadialDimReduction"]["projection_matrices"] = proj
    ctx, _ = make_ctx(hp, sts)
    X, _ = ctx.features()
    G = rng.normal(size=X.shape)
    g, _ = ctx.features_backward(G)
    pos = pack(sts)[1]
    h = 1e-5
    for atom in range(min(4, len(pos))):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                p = pos.copy()
                p[atom, d] += sgn * h
                ctx.update_positions(p)
                vals.append(np.sum(G * ctx.features()[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[atom, d]) < 1e-6 * np.abs(g).max() + 1e-6 * abs(fd)
    ctx.update_positions(pos)


@pytest.mark.parametrize("kind", ["DVR", "DVR+RDR", "GTO+RDR"])
def test_radial_options_mixed_precision(kind):
    """The mixed path (FP32 expansion and invariants from the same FP32
    spline tables) with the DVR basis and RadialDimReduction: normalised
    invariants within 1e-5 of the FP64 path (|X| = 1)."""
    hp = soap_hypers(rc=3.5, sw=0.5, sigma=0.4, N=6, L=4, spline=1e-8)
    sts = golden_structures()["CaCrP2O7"]
    if kind.startswith("DVR"):
        hp["radial_contribution"]["type"] = "DVR"
    if kind.endswith("RDR"):
        rng = np.random.default_rng(2)
        zs = sorted(set(pack(sts)[2].tolist()))
        hp["radial_contribution"]["optimization"]["RadialDimReduction"] = dict(
            projection_matrices={str(z): rng.uniform(size=(5, 6, 6)).tolist() for z in zs})
    ctx, _ = make_ctx(hp, sts)
    X64, _ = ctx.features()
    ctx.set_precision(_capi.MIXED)
    X32, _ = ctx.features()
    assert np.abs(X32 - X64).max() < 1e-5
