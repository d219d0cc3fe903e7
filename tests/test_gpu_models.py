"""The reference's Python model API (SparsePoints, Kernel, KRR, module
gradient functions) and the autograd bindings, on the GPU, against the FP64
oracle (tolerance 1e-10 relative, FP64 path)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, golden_structures, oracle_backward, rel_err
from librascal_amd.models import (KRR, Kernel, SparsePoints, compute_sparse_kernel_gradients,
                                  compute_sparse_kernel_neg_stress, train_gap_model)
from librascal_amd.representations import SphericalInvariants, from_hypers
from librascal_amd.synthetic import wrap_positions

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rep():
    return SphericalInvariants(interaction_cutoff=3.5, cutoff_smooth_width=0.5, max_radial=4,
                               max_angular=3, gaussian_sigma_type="Constant",
                               gaussian_sigma_constant=0.5)


def _frames():
    g = golden_structures()
    return g["CaCrP2O7"] + g["SiC_moissanite_supercell"]


def _model(rep, zeta, target="Structure"):
    frames = _frames()
    al = rep.transform(frames)
    sp = SparsePoints(rep)
    sel = [list(range(0, al.atom_offsets[s + 1] - al.atom_offsets[s], 3)) for s in range(len(al))]
    sp.extend(al, sel)
    kernel = Kernel(rep, name="GAP", zeta=zeta, target_type=target, kernel_type="Sparse")
    rng = np.random.default_rng(5 + zeta)
    w = rng.normal(size=sp.size())
    return frames, al, sp, kernel, w


@pytest.mark.parametrize("zeta", [1, 2, 3])
def test_kernel_kmm_and_knm(zeta):
    rep = _rep()
    frames, al, sp, kernel, _ = _model(rep, zeta)
    calc = O.Calculator(rep.hypers)
    b = O.BatchData(frames, species=al.species)
    ss = O.SparseSet(sp.numbers, sp.values_for(al.species), zeta)
    # K_MM (rows/columns in the order the points were added)
    kmm = kernel(sp)
    ref = np.zeros_like(kmm)
    ref[np.ix_(ss.order, ss.order)] = O.gap_kmm(ss)
    assert np.abs(kmm - ref).max() < TOL * np.abs(ref).max()
    # K_NM for both targets
    for target in ("Structure", "Atom"):
        kernel.target_type = target
        k = kernel(al, sp)
        kr = np.zeros_like(k)
        kr[:, ss.order] = O.gap_knm(calc, b, ss, target)
        assert rel_err(kr, k).max() < TOL


@pytest.mark.parametrize("zeta", [1, 2, 3])
def test_kernel_gradient_rows(zeta):
    """K_NM gradient rows with the negative-stress rows
    (SparseKernel::compute_derivative) vs the oracle, columns in the order
    the sparse points were added; contracted with the weights they give the
    model's energy gradient and negative stress."""
    rep = _rep()
    frames, al, sp, kernel, w = _model(rep, zeta)
    calc = O.Calculator(rep.hypers)
    b = O.BatchData(frames, species=al.species)
    ss = O.SparseSet(sp.numbers, sp.values_for(al.species), zeta)
    kd = kernel(al, sp, grad=(True, False), compute_neg_stress=True)
    assert kd.shape == (3 * b.n_atoms + 6 * b.n_struct, len(sp.numbers))
    ref = np.zeros_like(kd)
    ref[:, ss.order] = O.gap_knm_derivative(calc, b, ss, neg_stress=True)
    assert np.abs(kd - ref).max() < TOL * np.abs(ref).max()
    g = compute_sparse_kernel_gradients(rep, kernel, al, sp, w)
    ns = compute_sparse_kernel_neg_stress(rep, kernel, al, sp, w)
    assert np.abs(kd[:3 * b.n_atoms] @ w - g.ravel()).max() < TOL * np.abs(g).max()
    assert np.abs(kd[3 * b.n_atoms:] @ w - ns.ravel()).max() < TOL * np.abs(ns).max()


def _reference_relative_error(ref, test, epsilon):
    """math::relative_error (math/utils.hh:152-194): |ref - test| / |ref|,
    0 where both are below epsilon, the plain difference where only ref is."""
    diff = np.abs(ref - test)
    zr, zt = np.abs(ref) < epsilon, np.abs(test) < epsilon
    out = np.where(zr, diff, diff / np.where(zr, 1.0, np.abs(ref)))
    return np.where(zr & zt, 0.0, out)


_SKI_STRUCTURES = {"diamond_cubic_distorted.json": "diamond_cubic_distorted",
                   "CaCrP2O7_mvc-11955_symmetrized.json": "CaCrP2O7",
                   "small_molecules-20.json": "small_molecules_20"}


@pytest.mark.parametrize("case", range(6))
def test_sparse_kernel_reference_cases(case):
    """The reference's own sparse-kernel gradient cases
    (sparse_kernel_inputs.json, test_sparse_kernels.cc:126-330): its hypers
    (GTO, direct radial integrals, n 3, l 3, structure- or environment-wise
    keys, normalised or not), structures and selected sparse points, GAP
    zeta 2 on Structure targets.  K_NM derivative rows vs central finite
    differences of K_NM with the case's h (atoms displaced and wrapped as
    compute_displaced_kernel does; cells strained as displace_strain_tensor
    does, divided by -volume), with the reference's thresholds: relative
    error 1e-3 for the position rows (zero below 1e-14) and 6e-3 for the
    negative-stress rows (zero below 1e-11); and the prediction routines
    equal to the rows contracted with unit weights."""
    c = json.load(open(os.path.join(GOLDEN, "sparse_kernel_inputs.json")))[case]
    frames = golden_structures()[_SKI_STRUCTURES[os.path.basename(c["filename"])]]
    frames = [dict(f) for f in frames[:c["n_structures"]]]
    rep = from_hypers(c["calculator"])
    al = rep.transform(frames)
    sp = SparsePoints(rep)
    sp.extend(al, c["selected_ids"])
    kernel = Kernel(rep, name="GAP", zeta=c["kernel"]["zeta"], target_type="Structure",
                    kernel_type="Sparse")
    kd = kernel(al, sp, grad=(True, False), compute_neg_stress=True)
    n_at = sum(len(f["numbers"]) for f in frames)
    assert kd.shape == (3 * n_at + 6 * len(frames), sp.size())
    h = c["h"]
    copies, rows = [], []
    row = 0
    for f in frames:
        for a in range(len(f["numbers"])):
            for d in range(3):
                for sgn in (1.0, -1.0):
                    pos = f["positions"].copy()
                    pos[a, d] += sgn * h
                    copies.append(dict(f, positions=wrap_positions(pos, f["cell"], f["pbc"])))
                rows.append((row, 1.0 / (2 * h)))
                row += 1
    for f in frames:
        vol = abs(np.linalg.det(f["cell"]))
        for (x, y) in ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)):
            for sgn in (1.0, -1.0):
                shift = np.eye(3)
                shift[x, y] += sgn * h
                cell = shift.T @ f["cell"]
                pos = (cell @ np.linalg.inv(f["cell"]) @ f["positions"].T).T
                copies.append(dict(f, positions=pos, cell=cell))
            rows.append((row, -1.0 / (2 * h * vol)))
            row += 1
    k = kernel(rep.transform(copies), sp)
    fd = np.stack([(k[2 * q] - k[2 * q + 1]) * sc for q, (_, sc) in enumerate(rows)])
    npos = 3 * n_at
    ef = _reference_relative_error(kd[:npos], fd[:npos], 1e-14)
    es = _reference_relative_error(kd[npos:], fd[npos:], 1e-11)
    assert ef.max() < 1e-3, (np.unravel_index(ef.argmax(), ef.shape), ef.max())
    assert es.max() < 6e-3, (np.unravel_index(es.argmax(), es.shape), es.max())
    w = np.ones(sp.size())
    g = compute_sparse_kernel_gradients(rep, kernel, al, sp, w).ravel()
    ns = compute_sparse_kernel_neg_stress(rep, kernel, al, sp, w).ravel()
    assert _reference_relative_error(kd[:npos] @ w, g, 1e-14).max() < 1e-3
    assert _reference_relative_error(kd[npos:] @ w, ns, 1e-11).max() < 6e-3


@pytest.mark.parametrize("shape", [(300, 70), (1, 5), (129, 64), (1000, 257), (0, 4)])
def test_gram_f64_matches_numpy(shape):
    """rgpu_gram_f64 (A^T A on the FP64 matrix cores, the normal matrix of
    train_gap_model, krr.py:468) against numpy, for sizes off the 64 x 64
    tile and an empty A."""
    from librascal_amd import _capi
    A = np.random.default_rng(shape[0]).normal(size=shape)
    G = _capi.gram_f64(A)
    ref = A.T @ A
    assert G.shape == ref.shape
    assert np.abs(G - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0)


def test_krr_predictions_and_gradient_functions():
    rep = _rep()
    frames, al, sp, kernel, w = _model(rep, 2)
    self_contrib = {int(z): 0.1 * int(z) for z in set(al.numbers.tolist())}
    model = KRR(w, kernel, sp, self_contrib)
    calc = O.Calculator(rep.hypers)
    b = O.BatchData(frames, species=al.species)
    ss = O.SparseSet(sp.numbers, sp.values_for(al.species), 2)
    e_ref, g_ref, s_ref = O.gap_predict(calc, b, ss, w)
    y0 = np.array([sum(self_contrib[int(z)] for z in f["numbers"]) for f in frames])
    e = model.predict(al)
    assert rel_err(e_ref + y0, e).max() < TOL
    f = model.predict_forces(al)
    assert np.abs(f + g_ref).max() < TOL * np.abs(g_ref).max()
    s = model.predict_stress(al)
    assert np.abs(s + s_ref).max() < TOL * np.abs(s_ref).max()
    g = compute_sparse_kernel_gradients(rep, kernel, al, sp, w)
    ns = compute_sparse_kernel_neg_stress(rep, kernel, al, sp, w)
    assert np.abs(g - g_ref).max() < TOL * np.abs(g_ref).max()
    assert np.abs(ns - s_ref).max() < TOL * np.abs(s_ref).max()
    # a precomputed K_NM gives the same energies (krr.py:110-126)
    knm = kernel(al, sp)
    assert rel_err(e, model.predict(al, KNM=knm)).max() < TOL


def test_train_gap_model_matches_regularised_least_squares():
    """train_gap_model (krr.py:350-477) on noisy energies and energy
    gradients of a known model: its weights minimise
    |Lambda^-1 (K_NM a - y)|^2 + a^T (K_MM + jitter) a with Lambda^-1 =
    lambdas[0] / sigma_y * sqrt(N_atoms) on the energy rows and lambdas[1] /
    sigma_y on the gradient rows (sigma_y of the energies less the self
    contributions).  The same minimiser is found here by another route: one
    least-squares solve over the stacked matrix [Lambda^-1 K_NM; R], R^T R =
    K_MM + jitter (Cholesky), instead of the normal equations.  The noise and
    the moderate regularisation make the solution depend on every weight."""
    rep = _rep()
    frames, al, sp, kernel, w_true = _model(rep, 2)
    self_contrib = {int(z): -0.5 * int(z) for z in set(al.numbers.tolist())}
    true = KRR(w_true, kernel, sp, self_contrib)
    rng = np.random.default_rng(11)
    y = true.predict(al)
    grad = -true.predict_forces(al)
    y0 = np.array([sum(self_contrib[int(z)] for z in f["numbers"]) for f in frames])
    y = y + 0.2 * np.std(y - y0) * rng.normal(size=y.shape)
    grad = grad + 0.2 * np.std(grad) * rng.normal(size=grad.shape)
    knm = kernel(al, sp)
    kd = kernel(al, sp, grad=(True, False))
    lambdas, jitter = (0.05, 0.2), 1e-8
    model = train_gap_model(kernel, frames, np.vstack([knm, kd]), sp, y, self_contrib,
                            grad_train=grad, lambdas=lambdas, jitter=jitter)
    nat = np.diff(al.atom_offsets).astype(np.float64)
    yc = y - y0
    delta = np.std(yc)
    s_e = lambdas[0] / delta * np.sqrt(nat)
    s_g = lambdas[1] / delta
    A = np.vstack([knm / s_e[:, None], kd / s_g])
    b = np.concatenate([yc / s_e, grad.ravel() / s_g])
    kmm = kernel(sp) + jitter * np.eye(sp.size())
    R = np.linalg.cholesky(kmm).T
    alpha = np.linalg.lstsq(np.vstack([A, R]), np.concatenate([b, np.zeros(sp.size())]),
                            rcond=None)[0]
    w = np.asarray(model.get_weights()).ravel()
    # the predictions of the two solutions agree (the weights themselves
    # only as far as K_MM's conditioning allows)
    for K in (knm, kd):
        p_ref = K @ alpha
        assert np.abs(K @ w - p_ref).max() < 1e-6 * np.abs(p_ref).max()


def test_context_reloads_its_batch():
    """Two AtomsLists computed with one calculator: each one's features stay
    its own (the context is reloaded on demand)."""
    rep = _rep()
    g = golden_structures()
    a1 = rep.transform(g["CaCrP2O7"])
    x1 = a1.get_dense_features(rep).copy()
    a2 = rep.transform(g["small_molecule"])
    x2 = a2.get_dense_features(rep)
    assert x1.shape != x2.shape or np.abs(x1 - x2).max() > 0
    assert np.array_equal(a1.get_dense_features(rep), x1)


def test_autograd_features_backward_and_finite_differences():
    import torch
    from librascal_amd.autograd import soap_features
    rep = _rep()
    frames = golden_structures()["CaCrP2O7"]
    al = rep.transform(frames)
    pos = torch.tensor(al.positions, dtype=torch.float64, device="cuda", requires_grad=True)
    X = soap_features(pos, rep, al)
    rng = np.random.default_rng(3)
    R = torch.tensor(rng.normal(size=tuple(X.shape)), dtype=torch.float64, device="cuda")
    (X * R).sum().backward()
    calc = O.Calculator(rep.hypers)
    b = O.BatchData(frames, species=al.species)
    ref, _ = oracle_backward(calc, b, R.cpu().numpy())
    grad = pos.grad.cpu().numpy()
    assert np.abs(grad - ref).max() < TOL * np.abs(ref).max()
    # central differences of the GPU forward itself (h = 1e-4 A)
    h = 1e-4
    base = al.positions.copy()
    for atom, c in ((0, 0), (3, 2), (7, 1)):
        vals = []
        for sgn in (1, -1):
            p = base.copy()
            p[atom, c] += sgn * h
            Xs = soap_features(torch.tensor(p, device="cuda"), rep, al)
            vals.append(float((Xs * R).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grad[atom, c]) < 1e-6 * np.abs(grad).max()


def test_autograd_gap_energy_matches_forces():
    import torch
    from librascal_amd.autograd import gap_energy
    rep = _rep()
    frames, al, sp, kernel, w = _model(rep, 2)
    model = KRR(w, kernel, sp, {int(z): 0.0 for z in set(al.numbers.tolist())})
    forces = model.predict_forces(al)
    pos = torch.tensor(al.positions, dtype=torch.float64, device="cuda", requires_grad=True)
    E = gap_energy(pos, model, al)
    assert rel_err(model.predict(al), E.detach().cpu().numpy()).max() < TOL
    (2.0 * E.sum()).backward()
    assert np.abs(pos.grad.cpu().numpy() + 2.0 * forces).max() < TOL * np.abs(forces).max()


def test_empty_batch_fixes_the_layout():
    rep = _rep()
    rep._load_species([1, 6, 8])
    X, inv = rep._context.features()
    assert X.shape == (0, rep.get_num_coefficients(3)) and inv.shape == (0,)


def _model_file():
    import os
    from conftest import GOLDEN
    return os.path.join(GOLDEN, "simple_gap_model.json")


def test_reference_model_file_predicts_like_oracle():
    """The reference's saved KRR model (simple_gap_model.json: methane
    dimers, RadialScaling rc 7.5, n 4, l 2, direct Hyp1f1, 30 sparse points,
    zeta 2) loaded with io.load_obj: energies, forces and stress of methane
    dimers vs the oracle with the same hypers, sparse points and weights."""
    from conftest import wrap_cubic_periodic
    from librascal_amd.io import load_obj
    model = load_obj(_model_file())
    rep = model.get_representation_calculator()
    sts = [wrap_cubic_periodic(s) for s in golden_structures()["methane_dimers"][:8]]
    al = rep.transform(sts)
    e = model.predict(al)
    f = model.predict_forces(al)
    s = model.predict_stress(al)
    calc = O.Calculator(rep.hypers)
    b = O.BatchData(sts, species=al.species)
    ss = O.SparseSet(model.X_train.numbers, model.X_train.values_for(al.species), model.kernel.zeta)
    e_ref, g_ref, s_ref = O.gap_predict(calc, b, ss, model.weights)
    # the energies are small sums of large terms (|w| up to ~100): the error
    # is measured against the size of the terms, sum_m |K_Am w_m|
    K = np.zeros((b.n_struct, len(model.weights)))
    K[:, ss.order] = O.gap_knm(calc, b, ss, "Structure")
    assert np.all(np.abs(e - e_ref) < TOL * (np.abs(K) @ np.abs(model.weights)))
    assert np.abs(f + g_ref).max() < TOL * np.abs(g_ref).max()
    assert np.abs(s + s_ref).max() < TOL * np.abs(s_ref).max()
    # one GPU pass for all three
    e1, f1, s1 = model.predict_energy_forces_stress(al)
    assert np.abs(e1 - e).max() < 1e-12 * np.abs(e).max()
    assert np.abs(f1 - f).max() < 1e-12 * np.abs(f).max()
    assert np.abs(s1 - s).max() < 1e-12 * np.abs(s).max()


def test_generic_md_calculator():
    """models/genericmd.py behaviour (tests/python/md_calculator_test.py of
    the reference, on a methane dimer in its periodic box): repeated calls
    agree, uniform translations leave energy, forces and stress unchanged, a
    different cell changes them, the stress matrix is symmetric, and the
    argument checks raise the reference's errors."""
    from conftest import wrap_cubic_periodic
    from librascal_amd.genericmd import GenericMDCalculator
    st = wrap_cubic_periodic(golden_structures()["methane_dimers"][0])
    calc = GenericMDCalculator(_model_file(), True, structure_template=st)
    pos = np.asarray(st["positions"])
    cell_rows = np.asarray(st["cell"]).T
    e, f, s = calc.calculate(pos, cell_rows)
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(s)) and np.isfinite(e).all()
    assert f.shape == pos.shape and s.shape == (3, 3) and np.allclose(s, s.T)
    e2, f2, s2 = calc.calculate(pos, cell_rows)
    assert np.allclose(e, e2) and np.allclose(f, f2) and np.allclose(s, s2)
    e3, f3, s3 = calc.calculate(pos + 5.0, cell_rows)
    assert np.allclose(e, e3) and np.allclose(f, f3) and np.allclose(s, s3)
    e4, f4, _ = calc.calculate(pos, np.eye(3) * 1.2 * np.abs(cell_rows).max())
    assert not np.allclose(e, e4)
    other = GenericMDCalculator(_model_file(), True, atomic_numbers=st["numbers"])
    e5, f5, s5 = other.calculate(pos, cell_rows)
    assert np.allclose(e, e5) and np.allclose(f, f5) and np.allclose(s, s5)
    with pytest.raises(ValueError, match="Must specify one of 'structure_template' or 'atomic_numbers'"):
        GenericMDCalculator(_model_file(), True)
    with pytest.raises(ValueError, match="Structure template PBC flags.*incompatible"):
        GenericMDCalculator(_model_file(), False, structure_template=st)
    with pytest.raises(ValueError, match="Improper shape of positions"):
        calc.calculate(pos[:5], cell_rows)
    with pytest.raises(ValueError, match=r"Improper shape of cell info \(expected 3x3 matrix\)"):
        calc.calculate(pos, cell_rows[:2])


def test_autograd_backward_after_a_later_forward():
    """X1 = f(p1), X2 = f(p2) on the same AtomsList, then a backward through
    X1: the adjoint must use the expansion at p1 (the context holds p2)."""
    import torch
    from librascal_amd.autograd import soap_features
    rep = _rep()
    frames = golden_structures()["CaCrP2O7"]
    al = rep.transform(frames)
    rng = np.random.default_rng(8)
    p1 = torch.tensor(al.positions, dtype=torch.float64, device="cuda", requires_grad=True)
    p2 = torch.tensor(al.positions + 0.05 * rng.normal(size=al.positions.shape),
                      dtype=torch.float64, device="cuda", requires_grad=True)
    X1 = soap_features(p1, rep, al)
    X2 = soap_features(p2, rep, al)
    assert X1.is_cuda and X2.is_cuda
    R = torch.tensor(rng.normal(size=tuple(X1.shape)), dtype=torch.float64, device="cuda")
    (X1 * R).sum().backward()
    calc = O.Calculator(rep.hypers)
    st1 = [dict(f, positions=p1.detach().cpu().numpy()[al.atom_offsets[k]:al.atom_offsets[k + 1]])
           for k, f in enumerate(al.structures)]
    b1 = O.BatchData(st1, species=al.species)
    X1_ref, _ = calc.power_spectrum(b1)
    assert np.abs(X1.detach().cpu().numpy() - X1_ref).max() < TOL * np.abs(X1_ref).max()
    ref, _ = oracle_backward(calc, b1, R.cpu().numpy())
    grad = p1.grad.cpu().numpy()
    assert np.abs(grad - ref).max() < TOL * np.abs(ref).max()
    # and X2's backward still sees p2
    (X2 * R).sum().backward()
    st2 = [dict(f, positions=al.positions[al.atom_offsets[k]:al.atom_offsets[k + 1]])
           for k, f in enumerate(al.structures)]
    ref2, _ = oracle_backward(calc, O.BatchData(st2, species=al.species), R.cpu().numpy())
    assert np.abs(p2.grad.cpu().numpy() - ref2).max() < TOL * np.abs(ref2).max()


def test_sparse_points_rebuilt_with_a_reused_identity():
    """Sparse points made in a helper and freed, then others of the same
    size and species: the context must upload the new points (Python may
    give the new object the freed object's id)."""
    rep = _rep()
    frames, al, sp0, kernel, w = _model(rep, 2)

    def points(step):
        sp = SparsePoints(rep)
        sel = [list(range(step, al.atom_offsets[s + 1] - al.atom_offsets[s], 3))[:4]
               for s in range(len(al))]
        sp.extend(al, sel)
        return sp

    sp_a = points(0)
    n = sp_a.size()
    k_a = kernel(al, sp_a).copy()
    del sp_a
    sp_b = points(1)
    assert sp_b.size() == n
    k_b = kernel(al, sp_b)
    ref = O.gap_knm(O.Calculator(rep.hypers), O.BatchData(frames, species=al.species),
                    O.SparseSet(sp_b.numbers, sp_b.values_for(al.species), 2), "Structure")
    ss = O.SparseSet(sp_b.numbers, sp_b.values_for(al.species), 2)
    kr = np.zeros_like(ref)
    kr[:, ss.order] = ref
    assert rel_err(kr, k_b).max() < TOL
    assert np.abs(k_a - k_b).max() > 1e-6
    # extending a set bumps its version: the next kernel call sees the new rows
    sp_b.extend(al, [[0]] + [[] for _ in range(len(al) - 1)])
    assert kernel(al, sp_b).shape[1] == n + 1


def test_capi_sparse_points_dropped_with_another_species_list():
    """Sparse points set for species {1, 6}, then a batch with {1, 8}: the
    C ABI refuses GAP calls until sparse points are set again."""
    from librascal_amd import _capi
    import json
    from conftest import soap_hypers
    ctx = _capi.Context(json.dumps(soap_hypers()))
    cell = (np.eye(3) * 8.0).ravel()
    pos = np.array([[1.0, 1.0, 1.0], [2.0, 1.5, 1.0], [1.5, 2.5, 1.2]])
    ctx.set_structures([0, 3], pos, [1, 6, 6], cell, [1, 1, 1], [1, 6])
    X, _ = ctx.features()
    ctx.set_sparse_points(2, np.array([1, 6], np.int32), X[:2], np.ones(2))
    e, _, _ = ctx.gap_predict()
    ctx.set_structures([0, 3], pos, [1, 8, 8], cell, [1, 1, 1], [1, 8])
    for call in (lambda: ctx.gap_predict(), lambda: ctx.gap_knm(), lambda: ctx.gap_kmm()):
        with pytest.raises(_capi.RascalGPUError, match="no sparse points set"):
            call()
    # the same species list keeps them
    ctx.set_structures([0, 3], pos, [1, 6, 6], cell, [1, 1, 1], [1, 6])
    ctx.set_sparse_points(2, np.array([1, 6], np.int32), X[:2], np.ones(2))
    moved = pos.copy()
    moved[1, 0] += 0.2
    ctx.set_structures([0, 3], moved, [1, 6, 6], cell, [1, 1, 1], [1, 6])
    e2, _, _ = ctx.gap_predict()
    assert np.isfinite(e2).all() and abs(e2[0] - e[0]) > 0


def test_generic_md_step_equals_update_and_predict():
    """A GenericMDCalculator step after the first moves the atoms of its
    batch (set_positions) and runs one prediction: energy, forces and the
    Voigt stress equal a fresh transform + predict to 1e-12; the 3x3 matrix
    is the reference's (genericmd.py:123-129): Voigt entries in the upper
    triangle plus the transposed upper triangle, so the diagonal is doubled."""
    from conftest import wrap_cubic_periodic
    from librascal_amd.genericmd import GenericMDCalculator
    from librascal_amd.io import load_obj
    from librascal_amd.synthetic import wrap_positions
    st = wrap_cubic_periodic(golden_structures()["methane_dimers"][0])
    calc = GenericMDCalculator(_model_file(), True, structure_template=st)
    pos = np.asarray(st["positions"])
    cell_rows = np.asarray(st["cell"]).T
    calc.calculate(pos, cell_rows)
    al = calc.manager
    rng = np.random.default_rng(4)
    for step in range(3):
        p = pos + 0.02 * rng.normal(size=pos.shape)
        e, f, s = calc.calculate(p, cell_rows)
        assert calc.manager is al  # the batch is reused
        model = load_obj(_model_file())
        rep = model.get_representation_calculator()
        pw = wrap_positions(p, np.asarray(st["cell"]), np.asarray(st["pbc"]))
        ref = rep.transform([dict(st, positions=pw)])
        e_r, f_r, s_r = model.predict(ref), model.predict_forces(ref), model.predict_stress(ref)
        assert abs(e[0] - e_r[0]) < 1e-12 * max(abs(e_r[0]), 1.0)
        assert np.abs(f - f_r).max() < 1e-12 * np.abs(f_r).max()
        v = s_r.ravel()
        m = np.array([[2 * v[0], v[5], v[4]], [v[5], 2 * v[1], v[3]], [v[4], v[3], 2 * v[2]]])
        assert np.abs(s - m).max() < 1e-12 * np.abs(m).max()
    ctx = calc.representation._context
    assert ctx.info("nl_graph_launches") >= 1
