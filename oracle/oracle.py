"""ctypes front-end of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module: it is the checker and the CPU baseline, never the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class Hypers(C.Structure):
    _fields_ = [
        ("cutoff", C.c_double),
        ("smooth_width", C.c_double),
        ("gaussian_sigma", C.c_double),
        ("max_radial", C.c_int),
        ("max_angular", C.c_int),
        ("cutoff_type", C.c_int),
        ("rs_rate", C.c_double),
        ("rs_scale", C.c_double),
        ("rs_exponent", C.c_double),
        ("optimization", C.c_int),
        ("spline_accuracy", C.c_double),
        ("normalize", C.c_int),
        ("compute_gradients", C.c_int),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("n_struct", C.c_int),
        ("atom_offsets", C.POINTER(C.c_int64)),
        ("positions", C.POINTER(C.c_double)),
        ("types", C.POINTER(C.c_int)),
        ("cells", C.POINTER(C.c_double)),
        ("pbc", C.POINTER(C.c_int)),
        ("n_species", C.c_int),
        ("species", C.POINTER(C.c_int)),
    ]


class Sparse(C.Structure):
    _fields_ = [
        ("n_sparse", C.c_int),
        ("sp_species", C.POINTER(C.c_int)),
        ("values", C.POINTER(C.c_double)),
        ("zeta", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        d, i, i64, p = C.c_double, C.c_int, C.c_int64, C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_hyp1f1.restype = d
        L.orc_hyp1f1.argtypes = [d, d, d, i, i, d]
        L.orc_hyp1f1_switch_point.restype = d
        L.orc_hyp1f1_switch_point.argtypes = [d, d]
        L.orc_hyp1f1_sph.argtypes = [i, i, d, d, p, i, d, p, p]
        L.orc_ylm.argtypes = [i, p, p, p, p]
        L.orc_cutoff_fc.restype = d
        L.orc_cutoff_fc.argtypes = [C.POINTER(Hypers), d]
        L.orc_cutoff_dfc.restype = d
        L.orc_cutoff_dfc.argtypes = [C.POINTER(Hypers), d]
        L.orc_calc_create.restype = p
        L.orc_calc_create.argtypes = [C.POINTER(Hypers)]
        L.orc_calc_destroy.argtypes = [p]
        L.orc_calc_radial_info.argtypes = [p, p, p, p, p, p, p]
        L.orc_calc_radial.argtypes = [p, d, p, p]
        L.orc_calc_spline_tables.argtypes = [p, p, p]
        L.orc_nl_count.restype = i64
        L.orc_nl_count.argtypes = [i, p, p, p, d]
        L.orc_nl_fill.argtypes = [i, p, p, p, d, p, p, p, p]
        L.orc_batch_pairs.restype = i64
        L.orc_batch_pairs.argtypes = [p, C.POINTER(Batch)]
        L.orc_expansion.argtypes = [p, C.POINTER(Batch), p, p]
        L.orc_expansion_gradients.argtypes = [p, C.POINTER(Batch), p]
        L.orc_power_spectrum.argtypes = [p, C.POINTER(Batch), p, p, i64]
        L.orc_power_spectrum_gradients.argtypes = [p, C.POINTER(Batch), p, p]
        L.orc_gap_knm.argtypes = [p, C.POINTER(Batch), C.POINTER(Sparse), i, p]
        L.orc_gap_kmm.argtypes = [C.POINTER(Sparse), i, p]
        L.orc_gap_knm_derivative.argtypes = [p, C.POINTER(Batch), C.POINTER(Sparse), i, p]
        L.orc_gap_predict.argtypes = [p, C.POINTER(Batch), C.POINTER(Sparse), p, p, p, p, i64, i64]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def hyp1f1(a, b, z, derivative=False, mmax=200, tol=1e-15):
    return lib().orc_hyp1f1(a, b, z, int(derivative), int(mmax), float(tol))


def hyp1f1_switch_point(a, b):
    return lib().orc_hyp1f1_switch_point(a, b)


def hyp1f1_sph(max_radial, max_angular, r, alpha, fac_b, recursion=True, tol=1e-13):
    fac_b = np.ascontiguousarray(fac_b, dtype=np.float64)
    v = np.zeros((max_radial, max_angular + 1))
    dv = np.zeros_like(v)
    _check(lib().orc_hyp1f1_sph(max_radial, max_angular, r, alpha, _ptr(fac_b),
                                int(recursion), float(tol), _ptr(v), _ptr(dv)))
    return v, dv


def ylm(max_angular, direction, derivatives=False):
    d = np.ascontiguousarray(direction, dtype=np.float64)
    n = (max_angular + 1) ** 2
    y = np.zeros(n)
    dy = np.zeros((3, n)) if derivatives else None
    alp = np.zeros((max_angular + 1, max_angular + 2))
    _check(lib().orc_ylm(max_angular, _ptr(d), _ptr(y), _ptr(dy), _ptr(alp)))
    return y, dy, alp


def make_hypers(hp):
    """Build the oracle hyper struct from a reference-style hypers dict."""
    h = Hypers()
    fc = hp["cutoff_function"]
    h.cutoff = float(fc["cutoff"]["value"])
    h.smooth_width = float(fc["smooth_width"]["value"])
    h.gaussian_sigma = float(hp["gaussian_density"]["gaussian_sigma"]["value"])
    h.max_radial = int(hp["max_radial"])
    h.max_angular = int(hp["max_angular"])
    if fc["type"] == "ShiftedCosine":
        h.cutoff_type = 0
    elif fc["type"] == "RadialScaling":
        h.cutoff_type = 1
        h.rs_rate = float(fc["rate"]["value"])
        h.rs_scale = float(fc["scale"]["value"])
        h.rs_exponent = float(fc["exponent"]["value"])
    else:
        raise ValueError(fc["type"])
    opt = hp.get("radial_contribution", {}).get("optimization", {}) or {}
    if "Spline" in opt:
        h.optimization = 1
        h.spline_accuracy = float(opt["Spline"]["accuracy"])
    else:
        h.optimization = 0
    h.normalize = int(bool(hp.get("normalize", True)))
    h.compute_gradients = int(bool(hp.get("compute_gradients", False)))
    return h


def cutoff_fc(hp, r):
    h = make_hypers(hp)
    return lib().orc_cutoff_fc(C.byref(h), r)


def cutoff_dfc(hp, r):
    h = make_hypers(hp)
    return lib().orc_cutoff_dfc(C.byref(h), r)


def neighbours(positions, cell, pbc, cutoff):
    """Strict pair list of one structure: (offsets, j, vec, dist)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    cl = np.ascontiguousarray(cell, dtype=np.float64)
    pb = np.ascontiguousarray(pbc, dtype=np.int32)
    n = pos.shape[0]
    npairs = lib().orc_nl_count(n, _ptr(pos), _ptr(cl), _ptr(pb), cutoff)
    if npairs < 0:
        _check(-1)
    off = np.zeros(n + 1, np.int64)
    j = np.zeros(npairs, np.int32)
    vec = np.zeros((npairs, 3))
    dist = np.zeros(npairs)
    _check(lib().orc_nl_fill(n, _ptr(pos), _ptr(cl), _ptr(pb), cutoff, _ptr(off),
                             _ptr(j), _ptr(vec), _ptr(dist)))
    return off, j, vec, dist


class BatchData:
    """Packed batch of structures.  cell[s][d][k] = component d of lattice
    vector k (lattice vectors are columns)."""

    def __init__(self, structures, species=None):
        pos, typ, cells, pbc, off = [], [], [], [], [0]
        for st in structures:
            pos.append(np.asarray(st["positions"], np.float64).reshape(-1, 3))
            typ.append(np.asarray(st["numbers"], np.int32).ravel())
            cells.append(np.asarray(st["cell"], np.float64).reshape(3, 3))
            pbc.append(np.asarray(st["pbc"], np.int32).reshape(3))
            off.append(off[-1] + pos[-1].shape[0])
        self.positions = np.ascontiguousarray(np.concatenate(pos))
        self.types = np.ascontiguousarray(np.concatenate(typ))
        self.cells = np.ascontiguousarray(np.stack(cells))
        self.pbc = np.ascontiguousarray(np.stack(pbc))
        self.atom_offsets = np.asarray(off, np.int64)
        if species is None:
            species = sorted(set(self.types.tolist()))
        self.species = np.asarray(sorted(species), np.int32)
        self.n_struct = len(structures)
        self.n_atoms = int(off[-1])
        self._c = Batch()
        self._c.n_struct = self.n_struct
        self._c.atom_offsets = self.atom_offsets.ctypes.data_as(C.POINTER(C.c_int64))
        self._c.positions = self.positions.ctypes.data_as(C.POINTER(C.c_double))
        self._c.types = self.types.ctypes.data_as(C.POINTER(C.c_int))
        self._c.cells = self.cells.ctypes.data_as(C.POINTER(C.c_double))
        self._c.pbc = self.pbc.ctypes.data_as(C.POINTER(C.c_int))
        self._c.n_species = len(self.species)
        self._c.species = self.species.ctypes.data_as(C.POINTER(C.c_int))

    @property
    def n_species(self):
        return len(self.species)

    @property
    def c(self):
        return C.byref(self._c)


class Calculator:
    """Oracle SOAP calculator for one hypers dict."""

    def __init__(self, hypers):
        self.hypers = dict(hypers)
        self._h = make_hypers(hypers)
        self._p = lib().orc_calc_create(C.byref(self._h))
        if not self._p:
            _check(-1)
        self.N = self._h.max_radial
        self.L = self._h.max_angular

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_calc_destroy(self._p)
            self._p = None

    def radial_info(self):
        N = self.N
        fb = np.zeros(N)
        O = np.zeros((N, N))
        cen = np.zeros(N)
        g = C.c_int(0)
        x1 = C.c_double(0)
        x2 = C.c_double(0)
        lib().orc_calc_radial_info(self._p, _ptr(fb), _ptr(O), _ptr(cen), C.byref(g),
                                   C.byref(x1), C.byref(x2))
        return dict(fac_b=fb, ortho_norm=O, center=cen, grid=g.value, x1=x1.value, x2=x2.value)

    def radial(self, r, derivative=False):
        R = np.zeros((self.N, self.L + 1))
        dR = np.zeros_like(R) if derivative else None
        _check(lib().orc_calc_radial(self._p, r, _ptr(R), _ptr(dR)))
        return (R, dR) if derivative else R

    def spline_tables(self):
        info = self.radial_info()
        g = info["grid"]
        F = self.N * (self.L + 1)
        y = np.zeros((g, F))
        d2 = np.zeros((g, F))
        if lib().orc_calc_spline_tables(self._p, _ptr(y), _ptr(d2)) != 0:
            raise RuntimeError("no spline")
        return y, d2

    def D(self, n_species):
        return n_species * (n_species + 1) // 2 * self.N * self.N * (self.L + 1)

    def n_pairs(self, batch):
        n = lib().orc_batch_pairs(self._p, batch.c)
        if n < 0:
            _check(-1)
        return n

    def expansion(self, batch):
        S = batch.n_species
        c = np.zeros((batch.n_atoms, S, self.N, (self.L + 1) ** 2))
        mask = np.zeros(batch.n_atoms, np.uint64)
        _check(lib().orc_expansion(self._p, batch.c, _ptr(c), _ptr(mask)))
        return c, mask

    def expansion_gradients(self, batch):
        S = batch.n_species
        npairs = self.n_pairs(batch)
        g = np.zeros((npairs, 3, S, self.N, (self.L + 1) ** 2))
        _check(lib().orc_expansion_gradients(self._p, batch.c, _ptr(g)))
        return g

    def power_spectrum(self, batch, max_centres=-1):
        """(features [n][D], inverse norms [n]) of the first max_centres
        centres (all when negative)."""
        D = self.D(batch.n_species)
        n = batch.n_atoms if max_centres < 0 else min(max_centres, batch.n_atoms)
        p = np.zeros((n, D))
        inv = np.zeros(n)
        _check(lib().orc_power_spectrum(self._p, batch.c, _ptr(p), _ptr(inv),
                                        C.c_int64(int(max_centres))))
        return p, inv

    def power_spectrum_gradients(self, batch):
        D = self.D(batch.n_species)
        npairs = self.n_pairs(batch)
        g = np.zeros((npairs, 3, D))
        info = np.zeros((npairs, 5), np.int32)
        _check(lib().orc_power_spectrum_gradients(self._p, batch.c, _ptr(g), _ptr(info)))
        return g, info


class SparseSet:
    def __init__(self, species, values, zeta):
        order = np.argsort(np.asarray(species), kind="stable")
        self.species = np.ascontiguousarray(np.asarray(species, np.int32)[order])
        self.values = np.ascontiguousarray(np.asarray(values, np.float64)[order])
        self.order = order
        self.zeta = int(zeta)
        self._c = Sparse()
        self._c.n_sparse = len(self.species)
        self._c.sp_species = self.species.ctypes.data_as(C.POINTER(C.c_int))
        self._c.values = self.values.ctypes.data_as(C.POINTER(C.c_double))
        self._c.zeta = self.zeta

    @property
    def c(self):
        return C.byref(self._c)


def gap_knm(calc, batch, sparse, target="Structure"):
    M = sparse._c.n_sparse
    rows = batch.n_struct if target == "Structure" else batch.n_atoms
    k = np.zeros((rows, M))
    _check(lib().orc_gap_knm(calc._p, batch.c, sparse.c, 0 if target == "Structure" else 1,
                             _ptr(k)))
    return k


def gap_kmm(sparse):
    M = sparse._c.n_sparse
    k = np.zeros((M, M))
    _check(lib().orc_gap_kmm(sparse.c, sparse.values.shape[1], _ptr(k)))
    return k


def gap_knm_derivative(calc, batch, sparse, neg_stress=False):
    M = sparse._c.n_sparse
    rows = 3 * batch.n_atoms + (6 * batch.n_struct if neg_stress else 0)
    out = np.zeros((rows, M))
    _check(lib().orc_gap_knm_derivative(calc._p, batch.c, sparse.c, int(neg_stress),
                                        _ptr(out)))
    return out


def gap_predict(calc, batch, sparse, weights, energy=True, gradient=True,
                neg_stress=True, max_centres=-1, first_centre=0):
    """Energies, dE/dr and the negative stress of the centres [first_centre,
    first_centre + max_centres) of the batch (all centres by default)."""
    w = np.ascontiguousarray(np.asarray(weights, np.float64)[sparse.order])
    e = np.zeros(batch.n_struct) if energy else None
    g = np.zeros((batch.n_atoms, 3)) if gradient else None
    s = np.zeros((batch.n_struct, 6)) if neg_stress else None
    _check(lib().orc_gap_predict(calc._p, batch.c, sparse.c, _ptr(w), _ptr(e), _ptr(g),
                                 _ptr(s), int(max_centres), int(first_centre)))
    return e, g, s
