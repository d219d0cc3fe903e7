"""Seeded synthetic structures for the benchmark workloads.

Every generator returns a list of structures as dicts with keys
``positions`` [n][3] (Angstrom, inside the cell), ``numbers`` [n],
``cell`` [3][3] (lattice vectors as COLUMNS) and ``pbc`` [3].  Seeds follow
numpy PCG64 with seed 20260 + workload id.
"""
import numpy as np

SEED0 = 20260


def _structure(pos, numbers, cell_cols, pbc):
    return dict(positions=np.ascontiguousarray(pos, np.float64),
                numbers=np.ascontiguousarray(numbers, np.int32),
                cell=np.ascontiguousarray(cell_cols, np.float64),
                pbc=np.asarray(pbc, np.int32))


def wrap_positions(pos, cell_cols, pbc):
    inv = np.linalg.inv(cell_cols)
    frac = pos @ inv.T
    for d in range(3):
        if pbc[d]:
            frac[:, d] = np.mod(frac[:, d], 1.0)
    return frac @ cell_cols.T


def diamond_cells(n_cells=8, reps=2, a=5.43, jitter=0.1, species=14, seed=None):
    """Diamond-lattice supercells (8-atom conventional basis, reps^3 cells)
    with Gaussian jitter, periodic."""
    rng = np.random.Generator(np.random.PCG64(SEED0 + 0 if seed is None else seed))
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    grid = np.array([[i, j, k] for i in range(reps) for j in range(reps) for k in range(reps)])
    frac = (grid[:, None, :] + basis[None, :, :]).reshape(-1, 3) / reps
    cell = np.eye(3) * a * reps
    out = []
    for _ in range(n_cells):
        pos = frac @ cell.T + rng.normal(0.0, jitter, size=frac.shape)
        pos = wrap_positions(pos, cell, [1, 1, 1])
        out.append(_structure(pos, np.full(len(pos), species), cell, [1, 1, 1]))
    return out


def _rsa_box(rng, n, cell_cols, d_min, pbc, max_tries=200000):
    """Random sequential addition inside a cell with a minimum separation
    (minimum image over the fractional wrap for periodic directions)."""
    inv = np.linalg.inv(cell_cols)
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    pbc = np.asarray(pbc, bool)
    while k < n:
        tries += 1
        if tries > max_tries * max(n, 1):
            raise RuntimeError("random sequential addition did not converge")
        f = rng.random(3)
        x = cell_cols @ f
        if k:
            dfrac = (pos[:k] - x) @ inv.T
            dfrac[:, pbc] -= np.round(dfrac[:, pbc])
            d = dfrac @ cell_cols.T
            if np.min(np.einsum("ij,ij->i", d, d)) < d_min * d_min:
                continue
        pos[k] = x
        k += 1
    return pos


def molecules(n_mol=4096, n_min=8, n_max=64, density=0.08, d_min=0.9,
              species=(1, 6, 7, 8), probs=(0.5, 0.3, 0.1, 0.1), seed=None):
    """Non-periodic 'molecules': atom counts uniform in [n_min, n_max],
    uniform positions in a cube sized to the density, minimum separation."""
    rng = np.random.Generator(np.random.PCG64(SEED0 + 1 if seed is None else seed))
    out = []
    for _ in range(n_mol):
        n = int(rng.integers(n_min, n_max + 1))
        side = (n / density) ** (1.0 / 3.0)
        cell = np.eye(3) * side
        pos = _rsa_box(rng, n, cell, d_min, [0, 0, 0])
        num = rng.choice(np.asarray(species), size=n, p=np.asarray(probs))
        out.append(_structure(pos, num, cell, [0, 0, 0]))
    return out


def periodic_box(n_atoms, density, species, d_min=1.0, seed=0):
    """One cubic periodic box with equal (or given) species fractions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    side = (n_atoms / density) ** (1.0 / 3.0)
    cell = np.eye(3) * side
    pos = _rsa_box(rng, n_atoms, cell, d_min, [1, 1, 1])
    sp = np.asarray(species)
    num = sp[np.arange(n_atoms) % len(sp)]
    num = rng.permutation(num)
    return [_structure(pos, num, cell, [1, 1, 1])]


def triclinic_cells(n_cells=2048, n_min=32, n_max=512, species=(1, 6, 8),
                    fractions=(0.5, 0.3, 0.2), rho=(0.04, 0.08), lengths=(8.0, 20.0),
                    angles=(70.0, 110.0), d_min=1.0, seed=None):
    """Random triclinic periodic cells; lengths are rescaled so that the
    volume matches n_atoms / rho."""
    rng = np.random.Generator(np.random.PCG64(SEED0 + 4 if seed is None else seed))
    out = []
    while len(out) < n_cells:
        n = int(rng.integers(n_min, n_max + 1))
        r = rng.uniform(*rho)
        a, b, c = rng.uniform(*lengths, size=3)
        al, be, ga = np.radians(rng.uniform(*angles, size=3))
        cx = np.cos(be)
        cy = (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
        cz2 = 1 - cx * cx - cy * cy
        if cz2 <= 0.05:
            continue
        A = np.array([a, 0, 0])
        B = np.array([b * np.cos(ga), b * np.sin(ga), 0])
        Cv = np.array([c * cx, c * cy, c * np.sqrt(cz2)])
        cell = np.stack([A, B, Cv], axis=1)
        vol = abs(np.linalg.det(cell))
        cell *= ((n / r) / vol) ** (1.0 / 3.0)
        pos = _rsa_box(rng, n, cell, d_min, [1, 1, 1])
        num = rng.choice(np.asarray(species), size=n, p=np.asarray(fractions))
        out.append(_structure(pos, num, cell, [1, 1, 1]))
    return out


def pack(structures):
    """Concatenate structures into the flat batch arrays of the C ABI."""
    off = [0]
    for s in structures:
        off.append(off[-1] + len(s["numbers"]))
    pos = np.concatenate([np.asarray(s["positions"], np.float64).reshape(-1, 3)
                          for s in structures]) if structures else np.zeros((0, 3))
    num = np.concatenate([np.asarray(s["numbers"], np.int32) for s in structures]) \
        if structures else np.zeros(0, np.int32)
    cells = np.stack([np.asarray(s["cell"], np.float64) for s in structures]) \
        if structures else np.zeros((0, 3, 3))
    pbc = np.stack([np.asarray(s["pbc"], np.int32) for s in structures]) \
        if structures else np.zeros((0, 3), np.int32)
    return np.asarray(off, np.int64), pos, num, cells, pbc
