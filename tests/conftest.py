import json
import os
import sys

import numpy as np
import pytest

try:  # load PyTorch's HIP runtime before the library's (one runtime per process)
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def golden_structures():
    with open(os.path.join(GOLDEN, "structures.json")) as fh:
        d = json.load(fh)
    out = {}
    for k, v in d.items():
        out[k] = [dict(positions=np.asarray(s["positions"], float), numbers=np.asarray(s["numbers"]),
                       cell=np.asarray(s["cell"], float), pbc=np.asarray(s["pbc"], int),
                       energy=s.get("energy")) for s in v]
    return out


def wrap_cubic_periodic(st):
    """Shift + wrap a periodic structure into its cell (environments are
    unchanged)."""
    cell = np.asarray(st["cell"], float)
    inv = np.linalg.inv(cell)
    f = np.mod(st["positions"] @ inv.T, 1.0)
    return dict(st, positions=f @ cell.T)


def soap_hypers(rc=3.0, sw=0.5, sigma=0.4, N=4, L=3, spline=None, normalize=True,
                cutoff_type="ShiftedCosine", rs=None, invariants=True):
    fc = dict(type=cutoff_type, cutoff=dict(value=rc, unit="AA"),
              smooth_width=dict(value=sw, unit="AA"))
    if cutoff_type == "RadialScaling":
        rate, scale, expo = rs
        fc.update(rate=dict(value=rate, unit="AA"), scale=dict(value=scale, unit="AA"),
                  exponent=dict(value=expo, unit="AA"))
    h = dict(max_radial=N, max_angular=L, expansion_by_species_method="environment wise",
             global_species=[], compute_gradients=False, cutoff_function=fc,
             gaussian_density=dict(type="Constant", gaussian_sigma=dict(value=sigma, unit="AA")),
             radial_contribution=dict(type="GTO", optimization={} if spline is None else
                                      dict(Spline=dict(accuracy=spline))))
    if invariants:
        h.update(soap_type="PowerSpectrum", normalize=normalize, inversion_symmetry=True)
    return h


def rel_err(ref, test, eps=1e-14):
    """librascal's relative_error metric (math/utils.hh:152-194)."""
    ref = np.asarray(ref)
    test = np.asarray(test)
    diff = np.abs(ref - test)
    zr = np.abs(ref) < eps
    zt = np.abs(test) < eps
    out = np.where(zr & zt, 0.0, np.where(~zr, diff / np.where(zr, 1.0, np.abs(ref)), diff))
    return out


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def structures():
    return golden_structures()


def _block_rel(ref, test, scale):
    """max |test-ref| / scale with scale broadcast over the block; blocks that
    are identically zero in the reference are compared against the row
    scale."""
    return float(np.max(np.abs(np.asarray(test) - np.asarray(ref)) / scale))


FLOOR = 1e-4


def expansion_rel(ref, test, L):
    """Relative error per (centre, species, l) block of c[n][s][N][(L+1)^2]:
    max |diff| / max |ref| over the block's (n, m) entries.  Blocks smaller
    than FLOOR times the centre's largest coefficient (e.g. zero by symmetry,
    holding only rounding noise of the large entries) are measured against
    that floor."""
    ref = np.asarray(ref)
    test = np.asarray(test)
    worst = 0.0
    row = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    for l in range(L + 1):
        sl = slice(l * l, (l + 1) * (l + 1))
        r = ref[..., sl]
        sc = np.abs(r).max(axis=(2, 3), keepdims=True)
        sc = np.maximum(sc, FLOOR * row[:, None, None, None] + 1e-300)
        worst = max(worst, _block_rel(r, test[..., sl], sc))
    return worst


def ps_rel(ref, test, S, N, L):
    """Relative error per (centre, key, l) block of p[n][K][N*N][L+1]."""
    K = S * (S + 1) // 2
    ref = np.asarray(ref).reshape(-1, K, N * N, L + 1)
    test = np.asarray(test).reshape(ref.shape)
    sc = np.abs(ref).max(axis=2, keepdims=True)
    row = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    sc = np.maximum(sc, FLOOR * row[:, None, None, None] + 1e-300)
    return _block_rel(ref, test, sc)


def radial_rel(ref, test):
    """Relative error per radial function (n, l) over the sampled r:
    ref/test [n_r][N][L+1]."""
    ref = np.asarray(ref)
    sc = np.abs(ref).max(axis=0, keepdims=True)
    sc = np.where(sc > 0, sc, 1e-300)
    return _block_rel(ref, test, sc)


def oracle_backward(calc, batch, G):
    """dL/dr and neg stress of L = sum_i G_i . X_i from the oracle's stored
    per-pair feature gradients (the reference's gradient layout)."""
    import oracle as O
    grads, info = calc.power_spectrum_gradients(batch)
    n = batch.n_atoms
    out = np.zeros((n, 3))
    ns = np.zeros((batch.n_struct, 6))
    off = batch.atom_offsets
    vo = [(4, 2), (5, 0), (3, 1)]
    row = 0
    for s in range(batch.n_struct):
        st = slice(off[s], off[s + 1])
        ro, rj, rv, _ = O.neighbours(batch.positions[st], batch.cells[s], batch.pbc[s], calc._h.cutoff)
        for i in range(off[s + 1] - off[s]):
            for p in range(ro[i], ro[i + 1]):
                f = grads[row] @ G[off[s] + i]
                out[off[s] + rj[p]] += f
                rji = -rv[p]
                for x in range(3):
                    ns[s, x] += rji[x] * f[x]
                    ns[s, vo[x][0]] += rji[vo[x][1]] * f[x]
                row += 1
        ns[s] /= abs(np.linalg.det(batch.cells[s]))
    return out, ns


def random_sparse(calc, batch, n_per_species, rng, zeta):
    X, _ = calc.power_spectrum(batch)
    species, values = [], []
    for t in sorted(set(batch.types.tolist())):
        idx = np.where(batch.types == t)[0]
        pick = rng.choice(idx, size=n_per_species, replace=True)
        species += [t] * n_per_species
        values.append(X[pick] + 0.01 * rng.normal(size=(n_per_species, X.shape[1])))
    species = np.asarray(species)
    values = np.concatenate(values)
    perm = rng.permutation(len(species))
    return species[perm], values[perm], rng.normal(size=len(species))
