"""The benchmark configurations at their own sizes.

- The headline workload (periodic box of 8192 atoms, 4 species in equal
  fractions, density 0.06 / A^3, rc 5 A, n_max 12, l_max 9, 1024 sparse
  points per species, zeta 4): the mixed path against the FP64 GPU path
  (forces 1e-4 relative RMS, energy and negative stress 1e-4 relative), and
  the FP64 GPU path against the oracle on the first 200 centres (features,
  their energy, forces and stress; 1e-10 relative).
- The ragged molecule batch (4096 molecules of 8-64 atoms, H/C/N/O,
  rc 4 A, n_max 8, l_max 6) in one launch replayed from a HIP graph: FP32
  against FP64 GPU features (1e-5 of each row's largest entry), graph
  replays equal to eager launches, and FP64 against the oracle on the first
  24 molecules.
- Large triclinic cells (16 cells of 512 atoms, 3 species, rc 6 A,
  n_max 12, l_max 9, 8192 sparse points, zeta 2): mixed against FP64
  (forces 1e-4 relative RMS, energies 1e-4).
"""
import json

import numpy as np
import pytest

import oracle as O
from conftest import ps_rel, soap_hypers
from librascal_amd import _capi
from librascal_amd.synthetic import molecules, pack, periodic_box, triclinic_cells

pytestmark = pytest.mark.gpu


def _ctx(hp, sts, precision):
    ctx = _capi.Context(json.dumps(hp))
    ctx.set_precision(precision)
    off, pos, num, cells, pbc = pack(sts)
    ctx.set_structures(off, pos, num, cells, pbc, sorted(set(num.tolist())))
    return ctx


def _sparse(X, num, per_species, rng, replace=False):
    spec, vals = [], []
    for t in sorted(set(num.tolist())):
        idx = np.where(num == t)[0]
        pick = rng.choice(idx, size=per_species, replace=replace or per_species > len(idx))
        spec += [t] * per_species
        vals.append(X[pick])
    return np.asarray(spec, np.int32), np.concatenate(vals)


def _rms_rel(ref, test):
    return float(np.sqrt(np.mean((test - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


def test_headline_box_mixed_vs_fp64_and_oracle():
    sts = periodic_box(8192, 0.06, (1, 6, 7, 8), d_min=1.0, seed=20263)
    hp = soap_hypers(rc=5.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8)
    zeta = 4
    c64 = _ctx(hp, sts, _capi.FP64)
    X, _ = c64.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(20263)
    spec, vals = _sparse(X, num, 1024, rng)
    w = rng.normal(size=len(spec))
    c64.set_sparse_points(zeta, spec, vals, w)
    e64, g64, s64 = c64.gap_predict()
    cm = _ctx(hp, sts, _capi.MIXED)
    cm.enable_graphs(True)
    cm.set_sparse_points(zeta, spec, vals, w)
    em, gm, sm = cm.gap_predict()
    assert _rms_rel(g64, gm) < 1e-4
    assert abs(em[0] - e64[0]) < 1e-4 * abs(e64[0])
    assert np.abs(sm - s64).max() < 1e-4 * np.abs(s64).max()
    # a second (graph-replayed) step gives the same result to rounding
    off, pos, _, _, _ = pack(sts)
    cm.update_positions(pos)
    em2, gm2, _ = cm.gap_predict()
    assert _rms_rel(gm, gm2) < 1e-6 and abs(em2[0] - em[0]) < 1e-8 * abs(em[0])

    # FP64 GPU path vs the oracle on the first 200 centres: features, and
    # the energy, forces and stress of those centres' environments (the
    # GPU adjoint applied to their dE/dX)
    n0 = 200
    calc = O.Calculator(hp)
    b = O.BatchData(sts)
    Xr, _ = calc.power_spectrum(b, max_centres=n0)
    assert ps_rel(Xr, X[:n0], 4, 12, 9) < 1e-10
    same = num[:n0, None] == spec[None, :]
    K = X[:n0] @ vals.T
    e_sub = np.sum(np.where(same, K ** zeta, 0.0) @ w)
    dEdX = np.zeros_like(X)
    dEdX[:n0] = (np.where(same, zeta * K ** (zeta - 1), 0.0) * w[None, :]) @ vals
    g_sub, s_sub = c64.features_backward(dEdX)
    ss = O.SparseSet(spec, vals, zeta)
    e_ref, g_ref, s_ref = O.gap_predict(calc, b, ss, w, max_centres=n0)
    terms = np.abs(np.where(same, K ** zeta, 0.0)) @ np.abs(w)
    assert abs(e_sub - e_ref[0]) < 1e-10 * terms.sum()
    assert np.abs(g_sub - g_ref).max() < 1e-10 * np.abs(g_ref).max()
    assert np.abs(s_sub - s_ref).max() < 1e-10 * np.abs(s_ref).max()


def test_ragged_molecule_batch_one_graph_launch():
    sts = molecules(n_mol=4096, seed=20261)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=8, L=6, spline=1e-8)
    c64 = _ctx(hp, sts, _capi.FP64)
    X64, _ = c64.features()
    c32 = _ctx(hp, sts, _capi.MIXED)
    c32.enable_graphs(True)
    c32.compute()
    X32, _ = c32.features()
    assert c32.info("graph_launches") >= 1
    S = len(set(pack(sts)[2].tolist()))
    # FP32 features: error per row against the row's largest entry (blocks
    # far below it carry FP32 rounding of the large ones)
    row = np.abs(X64).max(axis=1)
    assert np.all(np.abs(X32 - X64).max(axis=1) <= 1e-5 * row)
    # replays of the captured representation stage equal the first launch
    for _ in range(2):
        c32.compute()
        assert np.array_equal(c32.features()[0], X32)
    assert c32.info("graph_launches") >= 3 and c32.info("graph_captures") == 1
    # FP64 GPU vs the oracle on the first 24 molecules
    sub = sts[:24]
    n_sub = sum(len(s["numbers"]) for s in sub)
    species = sorted(set(pack(sts)[2].tolist()))
    Xr, _ = O.Calculator(hp).power_spectrum(O.BatchData(sub, species=species))
    assert ps_rel(Xr, X64[:n_sub], S, 8, 6) < 1e-10


def test_large_triclinic_cells_mixed_vs_fp64():
    sts = triclinic_cells(n_cells=16, n_min=512, n_max=512, seed=20264)
    hp = soap_hypers(rc=6.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8)
    zeta = 2
    c64 = _ctx(hp, sts, _capi.FP64)
    X, _ = c64.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(20264)
    spec, vals = _sparse(X, num, 2731, rng, replace=True)
    vals = vals + 1e-3 * rng.normal(size=vals.shape)  # distinct points
    w = rng.normal(size=len(spec))
    c64.set_sparse_points(zeta, spec, vals, w)
    e64, g64, s64 = c64.gap_predict()
    cm = _ctx(hp, sts, _capi.MIXED)
    cm.set_sparse_points(zeta, spec, vals, w)
    cm.autotune()
    em, gm, sm = cm.gap_predict()
    assert _rms_rel(g64, gm) < 1e-4
    assert np.abs(em - e64).max() < 1e-4 * np.abs(e64).max()
    assert np.abs(sm - s64).max() < 1e-4 * np.abs(s64).max()
