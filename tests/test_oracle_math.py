"""Mathematical pins of the oracle (CPU): closed forms, integrals,
independent library routines, finite differences, brute force and
invariances.  Each is chosen so that a plausible slip (a wrong sign or
index, a dropped factor, a transposed operand) fails it."""
import numpy as np
import pytest
from scipy import integrate, special

import oracle as O
from conftest import golden_structures, soap_hypers, wrap_cubic_periodic
from librascal_amd.synthetic import diamond_cells, molecules, periodic_box, triclinic_cells

DBL_FTOL = 100 * np.finfo(float).eps


# --- Hyp1f1 -----------------------------------------------------------------
def test_hyp1f1_recursion_matches_direct():
    """tests/test_math_hyp1f1.cc:91-127 with its fixture (tests/test_math.hh
    :476-515: l_max, n_max in {4,5,9,15,16,20}, fac_a 0.4, rc in {2,3,5,7,8},
    tolerance 1e-14): downward recursion vs direct evaluation, mean relative
    difference <= 3 DBL_FTOL for values and derivatives."""
    maxes = (4, 5, 9, 15, 16, 20)
    for rc in (2.0, 3.0, 5.0, 7.0, 8.0):
        for L in maxes:
            for N in maxes:
                sig = np.array([(rc - 0.5) * max(np.sqrt(n), 1.0) / N for n in range(N)])
                fac_b = 0.5 * sig ** 2  # as written in the reference fixture
                for r in (1., 2., 3., 4., 5.5, 6.5, 7.5, 7.9):
                    if r >= rc:
                        continue
                    v1, d1 = O.hyp1f1_sph(N, L, r, 0.4, fac_b, recursion=False, tol=1e-14)
                    v2, d2 = O.hyp1f1_sph(N, L, r, 0.4, fac_b, recursion=True, tol=1e-14)
                    assert np.mean(np.abs(v1 - v2) / np.abs(v1)) <= 3 * DBL_FTOL
                    assert np.mean(np.abs(d1 - d2) / np.abs(d1)) <= 3 * DBL_FTOL


def test_hyp1f1_recursion_benchmark_regime():
    """Same comparison in the regime of the benchmark hypers (sigma 0.3-0.5,
    r up to 5): values to 3 DBL_FTOL on average; derivatives, which cancel
    near their zeros, as mean |difference| / mean |value| <= 1e-11."""
    for N, L in ((8, 6), (12, 9), (6, 3)):
        rc = 5.0
        sig = np.array([max(np.sqrt(n), 1.0) * rc / N for n in range(N)])
        fac_b = 0.5 / sig ** 2
        for r in (0.5, 1.7, 3.3, 4.9):
            for alpha in (0.5 / 0.5 ** 2, 0.5 / 0.3 ** 2):
                v1, d1 = O.hyp1f1_sph(N, L, r, alpha, fac_b, recursion=False)
                v2, d2 = O.hyp1f1_sph(N, L, r, alpha, fac_b, recursion=True)
                assert np.mean(np.abs(v1 - v2) / np.abs(v1)) <= 3 * DBL_FTOL
                assert np.mean(np.abs(d1 - d2)) / np.mean(np.abs(d1)) <= 1e-11


def test_hyp1f1_kummer_transformation():
    """1F1(a,b,z) = e^z 1F1(b-a,b,-z): ties the series, the asymptotic branch
    and the switching point to an identity (mpmath-free)."""
    for (a, b) in ((1.5, 1.5), (2.0, 1.5), (4.5, 2.5), (7.0, 4.5), (10.5, 9.5)):
        for z in (0.1, 1.0, 5.0, 20.0, 60.0):
            lhs = O.hyp1f1(a, b, z, False, 200, 1e-15)
            rhs = np.exp(z) * special.hyp1f1(b - a, b, -z)
            assert abs(lhs - rhs) / abs(rhs) < 1e-12


# --- spherical harmonics ------------------------------------------------------
def _real_ylm_scipy(L, d):
    x, y, z = d
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    out = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            c = special.sph_harm_y(l, abs(m), theta, phi)
            if m > 0:
                v = np.sqrt(2) * c.real
            elif m < 0:
                v = np.sqrt(2) * c.imag
            else:
                v = c.real
            out[l * l + l + m] = v
    return out


def test_ylm_closed_form_scipy():
    """Real harmonics with Condon-Shortley phase, m<0 = sine parts
    (spherical_harmonics.hh:41-75) equal sqrt(2) Re / Im of scipy's complex
    Y_l^|m| (and Re Y_l^0)."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        y, _, _ = O.ylm(8, d)
        assert np.max(np.abs(y - _real_ylm_scipy(8, d))) < 1e-13


def test_ylm_gradients_finite_difference():
    """dY/dr (on the sphere, times r) vs central differences of Y(x/|x|);
    also near the poles where the other branch of the gradient is used."""
    rng = np.random.default_rng(2)
    dirs = [rng.normal(size=3) for _ in range(10)] + [np.array([0.01, 0.02, 1.0]),
                                                     np.array([0.03, -0.02, -1.0])]
    L = 7
    h = 1e-6
    for d in dirs:
        d = d / np.linalg.norm(d)
        _, dy, _ = O.ylm(L, d, derivatives=True)
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            yp, _, _ = O.ylm(L, (d + e) / np.linalg.norm(d + e))
            ym, _, _ = O.ylm(L, (d - e) / np.linalg.norm(d - e))
            fd = (yp - ym) / (2 * h)
            assert np.max(np.abs(fd - dy[c])) < 1e-7


# --- cutoff and radial ---------------------------------------------------------
@pytest.mark.parametrize("ctype,rs", [("ShiftedCosine", None), ("RadialScaling", (1.0, 2.5, 3.0))])
def test_cutoff_derivative(ctype, rs):
    hp = soap_hypers(rc=4.0, sw=0.7, cutoff_type=ctype, rs=rs)
    for r in np.linspace(0.3, 3.95, 40):
        fd = (O.cutoff_fc(hp, r + 1e-6) - O.cutoff_fc(hp, r - 1e-6)) / 2e-6
        assert abs(fd - O.cutoff_dfc(hp, r)) < 1e-7
    assert O.cutoff_fc(hp, 4.0 + 1e-12) == 0.0


def test_gto_radial_integral_closed_form():
    """The raw GTO radial integral must equal, up to one constant common to
    all (n, l, r), e^{-a r^2} int_0^inf x^{n+2} e^{-(a+b_n) x^2} i_l(2 a r x) dx
    (Gaussian density expanded in spherical Bessel functions)."""
    hp = soap_hypers(rc=5.0, sigma=0.5, N=8, L=6)
    calc = O.Calculator(hp)
    info = calc.radial_info()
    a = 0.5 / 0.5 ** 2
    ratios = []
    for r in (0.4, 1.3, 2.9, 4.6):
        R = calc.radial(r)
        for n in range(8):
            for l in range(7):
                def f(x, n=n, l=l):
                    z = max(2 * a * r * x, 1e-300)
                    # i_l(z) e^{-z} = sqrt(pi / 2z) I_{l+1/2}(z) e^{-z}
                    il_scaled = np.sqrt(np.pi / (2 * z)) * special.ive(l + 0.5, z)
                    return x ** (n + 2) * np.exp(-(a + info["fac_b"][n]) * x * x + 2 * a * r * x) \
                        * il_scaled
                val, _ = integrate.quad(f, 0, np.inf, epsabs=0, epsrel=1e-13, limit=400)
                ratios.append(R[n, l] / (np.exp(-a * r * r) * val))
    ratios = np.array(ratios)
    assert np.max(np.abs(ratios / ratios[0] - 1)) < 1e-9


@pytest.mark.parametrize("spline", [None, 1e-8])
def test_radial_derivative_finite_difference(spline):
    calc = O.Calculator(soap_hypers(rc=5.0, N=6, L=4, spline=spline))
    for r in (0.7, 2.2, 3.9, 4.7):
        R, dR = calc.radial(r, derivative=True)
        fd = (calc.radial(r + 1e-6) - calc.radial(r - 1e-6)) / 2e-6
        assert np.max(np.abs(fd - dR)) < 1e-6 * max(1.0, np.abs(dR).max())


def test_spline_tracks_direct_evaluation():
    """The Spline(1e-8) radial function stays within the requested mean error
    of the direct one (interpolator error bound, test_math_interpolator.cc)."""
    direct = O.Calculator(soap_hypers(rc=5.0, N=8, L=6))
    spl = O.Calculator(soap_hypers(rc=5.0, N=8, L=6, spline=1e-8))
    O_mat = direct.radial_info()["ortho_norm"]
    errs = []
    for r in np.linspace(0.1, 4.99, 300):
        errs.append(np.abs(O_mat.T @ direct.radial(r) - spl.radial(r)).mean())
    assert np.mean(errs) < 1e-8


# --- neighbour list ------------------------------------------------------------
def brute_force_pairs(pos, cell, pbc, rc):
    inv = np.linalg.inv(cell)
    span = [int(np.ceil(rc * np.linalg.norm(inv[d]))) + 1 if pbc[d] else 0 for d in range(3)]
    images = np.array([[i, j, k] for i in range(-span[0], span[0] + 1)
                       for j in range(-span[1], span[1] + 1) for k in range(-span[2], span[2] + 1)])
    out = set()
    for i in range(len(pos)):
        for j in range(len(pos)):
            for n in images:
                v = pos[j] + cell @ n - pos[i]
                if v @ v <= rc * rc and not (i == j and not n.any()):
                    out.add((i, j, tuple(n)))
    return out


@pytest.mark.parametrize("name", ["diamond", "triclinic", "molecule", "small_cell"])
def test_neighbour_list_brute_force(name):
    """tests/test_adaptor_strict.cc:160-300: strict pair sets equal a brute
    force search over periodic images."""
    if name == "diamond":
        st, rc = diamond_cells(n_cells=1)[0], 5.0
    elif name == "triclinic":
        st, rc = triclinic_cells(n_cells=1, n_max=40)[0], 6.0
    elif name == "molecule":
        st, rc = molecules(n_mol=1)[0], 4.0
    else:
        st, rc = golden_structures()["diamond_cubic_distorted"][0], 5.0
    pos, cell, pbc = st["positions"], st["cell"], st["pbc"]
    off, j, vec, dist = O.neighbours(pos, cell, pbc, rc)
    inv = np.linalg.inv(cell)
    got = set()
    for i in range(len(pos)):
        assert j[off[i]] == i and np.all(vec[off[i]] == 0)  # self pair first
        for p in range(off[i] + 1, off[i + 1]):
            n = np.rint(inv @ (vec[p] - (pos[j[p]] - pos[i]))).astype(int)
            got.add((i, int(j[p]), tuple(n)))
    assert got == brute_force_pairs(pos, cell, pbc, rc)
    assert np.allclose(dist, np.linalg.norm(vec, axis=1), rtol=0, atol=1e-14)


def test_neighbour_list_invariant_under_cell_skew():
    """test_adaptor_neighbour_list.cc:433-555: counts unchanged when the
    same lattice is described by a skewed cell (a' = a + b)."""
    st = diamond_cells(n_cells=1)[0]
    cell = st["cell"]
    skew = cell.copy()
    skew[:, 0] = cell[:, 0] + cell[:, 1]
    inv = np.linalg.inv(skew)
    pos = np.mod(st["positions"] @ inv.T, 1.0) @ skew.T
    c1 = np.diff(O.neighbours(st["positions"], cell, [1, 1, 1], 4.0)[0])
    c2 = np.diff(O.neighbours(pos, skew, [1, 1, 1], 4.0)[0])
    assert np.array_equal(c1, c2)


# --- representation gradients, invariances ------------------------------------
def _fd_features(calc, st, species, h=1e-5, expansion=False):
    b = O.BatchData([st], species=species)
    base = calc.expansion(b)[0].reshape(b.n_atoms, -1) if expansion else calc.power_spectrum(b)[0]
    n = b.n_atoms
    out = np.zeros((n, 3, n, base.shape[1]))  # d X_i / d r_k
    for k in range(n):
        for c in range(3):
            res = []
            for sgn in (1, -1):
                s2 = dict(st)
                p = np.array(st["positions"], float)
                p[k, c] += sgn * h
                s2["positions"] = p
                bb = O.BatchData([s2], species=species)
                res.append(calc.expansion(bb)[0].reshape(n, -1) if expansion
                           else calc.power_spectrum(bb)[0])
            out[k, c] = (res[0] - res[1]) / (2 * h)
    return out


def _analytic_from_slots(calc, st, species, expansion=False):
    b = O.BatchData([st], species=species)
    if expansion:
        g = calc.expansion_gradients(b)
        g = g.reshape(g.shape[0], 3, -1)
    else:
        g, _ = calc.power_spectrum_gradients(b)
    off, j, _, _ = O.neighbours(st["positions"], st["cell"], st["pbc"], calc._h.cutoff)
    n = b.n_atoms
    out = np.zeros((n, 3, n, g.shape[2]))
    for i in range(n):
        for p in range(off[i], off[i + 1]):
            out[j[p], :, i] += g[p]
    return out


@pytest.mark.parametrize("spline", [None, 1e-8])
@pytest.mark.parametrize("expansion", [True, False])
def test_representation_gradients_finite_difference(spline, expansion):
    """test_calculator.cc:571-626 (RepresentationCalculatorGradientProvider,
    tolerance 3e-6): summed pair-slot gradients equal central differences,
    for the expansion and the normalised power spectrum."""
    st = periodic_box(14, 0.05, (1, 8), d_min=1.2, seed=3)[0]  # cell 6.5 A > 2 rc
    hp = soap_hypers(rc=3.0, sw=0.5, sigma=0.4, N=4, L=3, spline=spline)
    calc = O.Calculator(hp)
    ana = _analytic_from_slots(calc, st, [1, 8], expansion)
    fd = _fd_features(calc, st, [1, 8], expansion=expansion)
    scale = np.abs(ana).max()
    assert np.max(np.abs(ana - fd)) < 3e-6 * scale


def test_power_spectrum_invariances():
    """Rotation and permutation invariance, unit norm."""
    st = molecules(n_mol=1, n_min=12, n_max=12)[0]
    hp = soap_hypers(rc=4.0, N=6, L=5, spline=1e-8)
    calc = O.Calculator(hp)
    sp = sorted(set(st["numbers"].tolist()))
    X, _ = calc.power_spectrum(O.BatchData([st], species=sp))
    assert np.allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-13)
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    c = st["positions"].mean(axis=0)
    side = st["cell"][0, 0]
    rot = (st["positions"] - c) @ q.T
    # a larger non-periodic box holding the rotated molecule
    big = dict(st, positions=rot - rot.min(axis=0) + 1.0, cell=np.eye(3) * (side * 2 + 2))
    X2, _ = calc.power_spectrum(O.BatchData([big], species=sp))
    assert np.max(np.abs(X - X2)) < 1e-9
    perm = rng.permutation(len(st["numbers"]))
    st3 = dict(st, positions=st["positions"][perm], numbers=st["numbers"][perm])
    X3, _ = calc.power_spectrum(O.BatchData([st3], species=sp))
    assert np.max(np.abs(X[perm] - X3)) < 1e-12


# --- sparse GAP ----------------------------------------------------------------
def _gap_setup(zeta, seed=4):
    sts = periodic_box(12, 0.05, (1, 8), d_min=1.2, seed=seed)
    hp = soap_hypers(rc=3.0, sw=0.5, sigma=0.4, N=4, L=3, spline=1e-8)
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=[1, 8])
    X, _ = calc.power_spectrum(b)
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(X), 6, replace=False)
    ss = O.SparseSet(b.types[pick], X[pick] + 0.05 * rng.normal(size=(6, X.shape[1])), zeta)
    w = rng.normal(size=6)
    return sts, calc, ss, w


@pytest.mark.parametrize("zeta", [1, 2, 3, 4])
def test_gap_energy_closed_form(zeta):
    """K_NM = sum over centres of delta(species) (X_i . T_m)^zeta."""
    sts, calc, ss, w = _gap_setup(zeta)
    b = O.BatchData(sts, species=[1, 8])
    X, _ = calc.power_spectrum(b)
    K = O.gap_knm(calc, b, ss, "Atom")
    ref = np.where(b.types[:, None] == ss.species[None, :], (X @ ss.values.T) ** zeta, 0.0)
    assert np.max(np.abs(K - ref)) < 1e-13
    e, _, _ = O.gap_predict(calc, b, ss, w, True, False, False)
    assert abs(e[0] - ref.sum(axis=0) @ w[ss.order]) < 1e-12 * max(1, abs(e[0]))


@pytest.mark.parametrize("zeta", [1, 2, 4])
def test_gap_forces_and_stress_finite_difference(zeta):
    """test_sparse_kernels.cc grad_test / grad_stress_test: analytic forces
    and negative stress vs central differences of the energy (positions
    displaced, cell strained), and compute_sparse_kernel_gradients equal to
    K_NM_der . w."""
    sts, calc, ss, w = _gap_setup(zeta)
    b = O.BatchData(sts, species=[1, 8])
    e0, g, s = O.gap_predict(calc, b, ss, w)
    h = 1e-5
    st = sts[0]
    fd = np.zeros_like(g)
    for k in range(len(st["numbers"])):
        for c in range(3):
            es = []
            for sgn in (1, -1):
                p = st["positions"].copy()
                p[k, c] += sgn * h
                bb = O.BatchData([dict(st, positions=p)], species=[1, 8])
                es.append(O.gap_predict(calc, bb, ss, w, True, False, False)[0][0])
            fd[k, c] = (es[0] - es[1]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-6 * np.abs(g).max()
    # neg stress: d E / d eps_ab / V, Voigt xx yy zz yz xz xy (sym. strain)
    voigt = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    vol = abs(np.linalg.det(st["cell"]))
    for v, (a, c) in enumerate(voigt):
        es = []
        for sgn in (1, -1):
            eps = np.eye(3)
            eps[a, c] += sgn * h
            if a != c:
                eps[c, a] += sgn * h
            bb = O.BatchData([dict(st, positions=st["positions"] @ eps.T, cell=eps @ st["cell"])],
                             species=[1, 8])
            es.append(O.gap_predict(calc, bb, ss, w, True, False, False)[0][0])
        d = (es[0] - es[1]) / (2 * h) / vol
        if a != c:
            d /= 2
        # negative stress = -(1/V) dE/d(strain) (numerical_kernel_gradients.hh:160-162)
        assert abs(-d - s[0, v]) < 1e-6 * np.abs(s).max()
    kd = O.gap_knm_derivative(calc, b, ss, neg_stress=True)
    assert np.max(np.abs(kd[:3 * b.n_atoms] @ w[ss.order] - g.ravel())) < 1e-12 * np.abs(g).max()
    assert np.max(np.abs(kd[3 * b.n_atoms:] @ w[ss.order] - s.ravel())) < 1e-12 * np.abs(s).max()


@pytest.mark.parametrize("zeta", [1, 2, 3])
def test_gap_kmm_closed_form(zeta):
    """sparse_kernels.hh:136-150: K_MM is block diagonal by species with
    blocks (T_a T_a^T)^zeta, rows and columns in species-sorted order."""
    _, _, ss, _ = _gap_setup(zeta)
    K = O.gap_kmm(ss)
    T, s = ss.values, ss.species  # stored species-sorted
    ref = np.where(s[:, None] == s[None, :], (T @ T.T) ** zeta, 0.0)
    assert np.max(np.abs(K - ref)) < 1e-13 * max(1.0, np.abs(ref).max())
    assert np.allclose(K, K.T)


def test_gap_knm_derivative_columns_finite_difference():
    """Each column of the K_NM derivative (sparse_kernels.hh:258-499) against
    central differences of the Structure kernel column, for two columns."""
    sts, calc, ss, _ = _gap_setup(2)
    b = O.BatchData(sts, species=[1, 8])
    kd = O.gap_knm_derivative(calc, b, ss, neg_stress=False)
    st = sts[0]
    h = 1e-5
    for m in (0, 3):
        fd = np.zeros(3 * len(st["numbers"]))
        for k in range(len(st["numbers"])):
            for c in range(3):
                vals = []
                for sgn in (1, -1):
                    p = st["positions"].copy()
                    p[k, c] += sgn * h
                    bb = O.BatchData([dict(st, positions=p)], species=[1, 8])
                    vals.append(O.gap_knm(calc, bb, ss, "Structure")[0, m])
                fd[3 * k + c] = (vals[0] - vals[1]) / (2 * h)
        assert np.max(np.abs(fd - kd[:, m])) < 1e-6 * max(1e-12, np.abs(kd[:, m]).max())


def test_gap_predict_centre_ranges_add_up():
    """The oracle's prediction over consecutive centre ranges (the CPU
    baseline's parallel workers) sums to the prediction over all centres."""
    from librascal_amd.synthetic import periodic_box
    sts = periodic_box(40, 0.05, (1, 8), d_min=1.0, seed=3)
    hp = soap_hypers(rc=3.5, N=4, L=3, sigma=0.5)
    calc = O.Calculator(hp)
    b = O.BatchData(sts)
    X, _ = calc.power_spectrum(b)
    rng = np.random.default_rng(2)
    sp = O.SparseSet(b.types[:6], X[:6] + 0.01 * rng.normal(size=X[:6].shape), 2)
    w = rng.normal(size=6)
    e, g, s = O.gap_predict(calc, b, sp, w)
    parts = [O.gap_predict(calc, b, sp, w, first_centre=a, max_centres=15) for a in (0, 15, 30)]
    assert np.allclose(sum(p[0] for p in parts), e, rtol=1e-12, atol=0)
    assert np.allclose(sum(p[1] for p in parts), g, rtol=1e-12, atol=1e-14)
    assert np.allclose(sum(p[2] for p in parts), s, rtol=1e-12, atol=1e-14)
