"""ctypes binding of the C ABI in include/rascal_mi355x.h.

The shared library must have been built (``python -m librascal_amd.build``);
there is no CPU fallback: loading fails loudly when it is missing, and every
call raises when the library reports an error.
"""
import ctypes as C
import os

import numpy as np

from .build import LIB_PATH

RGPU_OK = 0
RGPU_ERR_INVALID = 1
RGPU_ERR_RUNTIME = 2
RGPU_ERR_HIP = 3
RGPU_ERR_STATE = 4

FP64 = 0
MIXED = 1

STAGES = dict(neighbours=0, expansion=1, power_spectrum=2, gap_kernel=3, gap_force=4, backward=5)

INFO = dict(max_radial=0, max_angular=1, n_species=2, feature_size=3, n_centres=4,
            n_pairs=5, n_struct=6, spline_grid=7, n_sparse=8, is_invariants=9, precision=10,
            graph_captures=11, graph_launches=12, tile_pairs_fwd=13, tile_keys_ps=14,
            tile_pairs_adj=15, tile_keys_adj=16, nl_graph_launches=17, n_atoms=18,
            tile_adj_kind=19, half_list=20, fused=21, fwd_lds=22, fwd_group=23)

_lib = None


class RascalGPUError(RuntimeError):
    """Error raised by the MI355X library (carries the reference's message)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    # When PyTorch shares the process, its HIP runtime must be the one loaded
    # first (the library then binds to the same runtime instead of loading a
    # second copy).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "librascal_mi355x.so is not built (%s); run `python -m librascal_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    p, i, i64 = C.c_void_p, C.c_int, C.c_int64
    L.rgpu_last_error.restype = C.c_char_p
    L.rgpu_context_create.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.rgpu_context_destroy.argtypes = [p]
    L.rgpu_context_destroy.restype = None
    L.rgpu_set_stream.argtypes = [p, p]
    L.rgpu_set_precision.argtypes = [p, i]
    L.rgpu_set_half_list.argtypes = [p, i]
    L.rgpu_enable_graphs.argtypes = [p, i]
    L.rgpu_autotune.argtypes = [p]
    L.rgpu_set_tiles.argtypes = [p, i, i, i, i, i, i]
    L.rgpu_get_info.argtypes = [p, i, C.POINTER(C.c_int64)]
    L.rgpu_set_structures.argtypes = [p, i, p, p, p, p, p, i, p, p]
    L.rgpu_update_positions.argtypes = [p, p, p]
    L.rgpu_get_neighbours.argtypes = [p, p, p, p]
    L.rgpu_compute.argtypes = [p]
    L.rgpu_get_expansion.argtypes = [p, p, p]
    L.rgpu_get_features.argtypes = [p, p, p]
    L.rgpu_get_features_gradient.argtypes = [p, p]
    L.rgpu_features_backward.argtypes = [p, p, p, p]
    L.rgpu_set_sparse_points.argtypes = [p, i, i64, p, p, p]
    L.rgpu_gap_knm.argtypes = [p, i, p]
    L.rgpu_gap_kmm.argtypes = [p, p]
    L.rgpu_gram_f64.argtypes = [i64, i64, p, p]
    L.rgpu_gap_knm_gradient.argtypes = [p, i, p]
    L.rgpu_gap_predict.argtypes = [p, p, p, p]
    L.rgpu_gap_energy_noalloc.argtypes = [p, p]
    L.rgpu_eval_radial.argtypes = [p, i64, p, p, p]
    L.rgpu_eval_ylm.argtypes = [p, i64, p, p, p]
    L.rgpu_enable_timing.argtypes = [p, i]
    L.rgpu_get_timing.argtypes = [p, i, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.rgpu_reset_timing.argtypes = [p]
    for name in ("rgpu_context_create", "rgpu_set_stream", "rgpu_set_precision", "rgpu_get_info",
                 "rgpu_set_half_list",
                 "rgpu_enable_graphs", "rgpu_autotune", "rgpu_set_tiles",
                 "rgpu_set_structures", "rgpu_update_positions", "rgpu_get_neighbours",
                 "rgpu_compute", "rgpu_get_expansion", "rgpu_get_features", "rgpu_get_features_gradient",
                 "rgpu_features_backward", "rgpu_set_sparse_points", "rgpu_gap_knm", "rgpu_gap_kmm", "rgpu_gram_f64", "rgpu_gap_knm_gradient",
                 "rgpu_gap_predict", "rgpu_gap_energy_noalloc", "rgpu_eval_radial", "rgpu_eval_ylm", "rgpu_enable_timing",
                 "rgpu_get_timing", "rgpu_reset_timing"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def gram_f64(A):
    """A^T A (FP64) on the GPU's FP64 matrix cores, for a host array A."""
    A = np.ascontiguousarray(A, np.float64)
    if A.ndim != 2:
        raise ValueError("gram_f64 takes a 2-d array")
    out = np.empty((A.shape[1], A.shape[1]), np.float64)
    check(lib().rgpu_gram_f64(int(A.shape[0]), int(A.shape[1]), ptr(A), ptr(out)))
    return out


def check(rc):
    if rc != RGPU_OK:
        msg = lib().rgpu_last_error().decode()
        if rc == RGPU_ERR_INVALID:
            raise RascalGPUError(rc, msg)
        raise RascalGPUError(rc, msg)


def ptr(a):
    """Host numpy array or torch tensor (host or device) -> void pointer."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays passed to the library must be C-contiguous"
        return a.ctypes.data_as(C.c_void_p)
    # torch tensor
    assert a.is_contiguous(), "tensors passed to the library must be contiguous"
    return C.c_void_p(a.data_ptr())


def _check_out(a, shape):
    """Output arrays and tensors: float64, contiguous, of the given shape."""
    if tuple(a.shape) != tuple(shape):
        raise ValueError("output of shape %s expected, got %s" % (tuple(shape), tuple(a.shape)))
    if isinstance(a, np.ndarray):
        ok = a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    else:
        ok = str(a.dtype) == "torch.float64" and a.is_contiguous()
    if not ok:
        raise ValueError("outputs must be contiguous float64")


class Context:
    """Thin owner of an rgpu_context (one calculator + batch + GAP model)."""

    def __init__(self, hypers_json):
        self._p = C.c_void_p()
        check(lib().rgpu_context_create(hypers_json.encode(), C.byref(self._p)))

    def __del__(self):
        if getattr(self, "_p", None) and self._p.value and _lib is not None:
            _lib.rgpu_context_destroy(self._p)
            self._p = None

    @property
    def handle(self):
        return self._p

    def info(self, key):
        v = C.c_int64(0)
        check(lib().rgpu_get_info(self._p, INFO[key], C.byref(v)))
        return v.value

    def set_stream(self, stream_handle):
        check(lib().rgpu_set_stream(self._p, C.c_void_p(stream_handle)))

    def set_half_list(self, enable=True):
        """Expansion over the half neighbour list (results equal the full
        list's)."""
        check(lib().rgpu_set_half_list(self._p, int(bool(enable))))

    def set_precision(self, precision):
        check(lib().rgpu_set_precision(self._p, int(precision)))

    def enable_graphs(self, enable=True):
        check(lib().rgpu_enable_graphs(self._p, int(enable)))

    def autotune(self):
        """Pick the fastest work partitions for the current batch; returns
        them as a dict."""
        check(lib().rgpu_autotune(self._p))
        return self.tiles()

    def tiles(self):
        """Work partitions of the current precision (0 / -1: built-in)."""
        return dict(pairs_fwd=self.info("tile_pairs_fwd"), keys_ps=self.info("tile_keys_ps"),
                    pairs_adj=self.info("tile_pairs_adj"),
                    keys_adj=self.info("tile_keys_adj"), adj_kind=self.info("tile_adj_kind"),
                    half_list=self.info("half_list"),
                    fused=2 if self.info("fwd_lds") else (3 if self.info("fwd_group")
                                                          else self.info("fused")))

    def set_tiles(self, pairs_fwd=0, keys_ps=0, pairs_adj=0, keys_adj=0, adj_kind=3, fused=-1):
        """Set the work partitions of the current precision (see tiles())."""
        check(lib().rgpu_set_tiles(self._p, int(pairs_fwd), int(keys_ps), int(pairs_adj),
                                   int(keys_adj), int(adj_kind), int(fused)))

    def set_structures(self, atom_offsets, positions, types, cells, pbc, species,
                       centre_mask=None):
        """cells [n_struct][3][3] with lattice vectors as columns;
        centre_mask [n_atoms] (bool, None: every atom is a centre)."""
        ao = np.ascontiguousarray(atom_offsets, np.int64)
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
        ty = np.ascontiguousarray(types, np.int32)
        ce = np.ascontiguousarray(cells, np.float64).reshape(-1, 9)
        pb = np.ascontiguousarray(pbc, np.int32).reshape(-1, 3)
        sp = np.ascontiguousarray(sorted(species), np.int32)
        cm = None
        if centre_mask is not None:
            cm = np.ascontiguousarray(np.asarray(centre_mask, bool).ravel(), np.uint8)
            if cm.shape[0] != ty.shape[0]:
                raise ValueError("centre_mask needs one entry per atom")
        check(lib().rgpu_set_structures(self._p, len(ao) - 1, ptr(ao), ptr(pos), ptr(ty), ptr(ce),
                                        ptr(pb), len(sp), ptr(sp), ptr(cm)))

    def update_positions(self, positions, cells=None):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
        ce = None if cells is None else np.ascontiguousarray(cells, np.float64).reshape(-1, 9)
        check(lib().rgpu_update_positions(self._p, ptr(pos), ptr(ce)))

    def neighbours(self, with_vectors=True):
        """Strict pairs of every atom (self pair first): offsets [n_atoms+1],
        neighbour atom [n_pairs], pair vectors [n_pairs][3]."""
        n = self.info("n_atoms")
        npairs = self.info("n_pairs")
        off = np.zeros(n + 1, np.int64)
        pj = np.zeros(npairs, np.int32)
        vec = np.zeros((npairs, 3)) if with_vectors else None
        check(lib().rgpu_get_neighbours(self._p, ptr(off), ptr(pj), ptr(vec)))
        return off, pj, vec

    def compute(self):
        check(lib().rgpu_compute(self._p))

    def expansion(self):
        n = self.info("n_centres")
        S = self.info("n_species")
        N = self.info("max_radial")
        L1 = self.info("max_angular") + 1
        c = np.zeros((n, S, N, L1 * L1))
        km = np.zeros(n, np.uint64)
        check(lib().rgpu_get_expansion(self._p, ptr(c), ptr(km)))
        return c, km

    def key_masks(self):
        """Per-centre key masks (bit s: species s is an expansion key)."""
        km = np.zeros(self.info("n_centres"), np.uint64)
        check(lib().rgpu_get_expansion(self._p, None, ptr(km)))
        return km

    def features(self, out=None):
        """(X [n_centres][D], inverse norms [n_centres]).  With out (a float64
        array or tensor of that shape, host or device) X is written there and
        the norms are not returned."""
        n = self.info("n_centres")
        D = self.info("feature_size")
        if out is not None:
            _check_out(out, (n, D))
            check(lib().rgpu_get_features(self._p, ptr(out), None))
            return out, None
        X = np.zeros((n, D))
        inv = np.zeros(n)
        check(lib().rgpu_get_features(self._p, ptr(X), ptr(inv)))
        return X, inv

    def features_backward(self, dL_dX, want_stress=True, out=None):
        """dL/dr [n_atoms][3] (and the negative virial) for dL/dX given as a
        host array or a contiguous float64 torch tensor (host or device);
        dL/dr goes into out when given (float64, host or device)."""
        n = self.info("n_atoms")
        ns = self.info("n_struct")
        if hasattr(dL_dX, "data_ptr"):
            g = dL_dX
            if str(g.dtype) != "torch.float64" or not g.is_contiguous():
                raise ValueError("dL/dX tensors must be contiguous float64")
        else:
            g = np.ascontiguousarray(dL_dX, np.float64)
        if out is None:
            out = np.zeros((n, 3))
        else:
            _check_out(out, (n, 3))
        st = np.zeros((ns, 6)) if want_stress else None
        check(lib().rgpu_features_backward(self._p, ptr(g), ptr(out), ptr(st)))
        return out, st

    def features_gradient(self, out=None):
        """dX/dr [3 n_pairs][D]: rows 3 r + x for the pairs r of neighbours()
        (self pair first), FP64."""
        rows = 3 * self.info("n_pairs")
        D = self.info("feature_size")
        if out is None:
            out = np.zeros((rows, D))
        else:
            _check_out(out, (rows, D))
        check(lib().rgpu_get_features_gradient(self._p, ptr(out)))
        return out

    def set_sparse_points(self, zeta, species, values, weights=None):
        sp = np.ascontiguousarray(species, np.int32)
        v = np.ascontiguousarray(values, np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        check(lib().rgpu_set_sparse_points(self._p, int(zeta), len(sp), ptr(sp), ptr(v), ptr(w)))

    def gap_knm(self, target="Structure"):
        t = 0 if target == "Structure" else 1
        rows = self.info("n_struct") if t == 0 else self.info("n_centres")
        k = np.zeros((rows, self.info("n_sparse")))
        check(lib().rgpu_gap_knm(self._p, t, ptr(k)))
        return k

    def gap_knm_gradient(self, neg_stress=False):
        """K_NM gradient rows [3 n_atoms (+ 6 n_struct)][n_sparse] (FP64)."""
        n, ns, M = self.info("n_atoms"), self.info("n_struct"), self.info("n_sparse")
        out = np.zeros((3 * n + (6 * ns if neg_stress else 0), M))
        check(lib().rgpu_gap_knm_gradient(self._p, int(bool(neg_stress)), ptr(out)))
        return out

    def gap_kmm(self):
        M = self.info("n_sparse")
        k = np.zeros((M, M))
        check(lib().rgpu_gap_kmm(self._p, ptr(k)))
        return k

    def gap_predict(self, energy=True, gradient=True, neg_stress=True, out=None):
        """(energy [n_struct], gradient [n_atoms][3], neg_stress [n_struct][6]),
        None for the outputs not requested; out = (e, g, s) gives the
        destination arrays or tensors (float64, host or device, None to skip)."""
        ns = self.info("n_struct")
        n = self.info("n_atoms")
        if out is not None:
            e, g, s = out
            for a, shape in ((e, (ns,)), (g, (n, 3)), (s, (ns, 6))):
                if a is not None:
                    _check_out(a, shape)
        else:
            e = np.zeros(ns) if energy else None
            g = np.zeros((n, 3)) if gradient else None
            s = np.zeros((ns, 6)) if neg_stress else None
        check(lib().rgpu_gap_predict(self._p, ptr(e), ptr(g), ptr(s)))
        return e, g, s

    def enable_timing(self, enable=True):
        check(lib().rgpu_enable_timing(self._p, int(enable)))

    def reset_timing(self):
        check(lib().rgpu_reset_timing(self._p))

    def timing(self):
        """{stage: (total_ms, launches)} since the last reset."""
        out = {}
        for name, k in STAGES.items():
            ms = C.c_double(0)
            n = C.c_int64(0)
            check(lib().rgpu_get_timing(self._p, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def eval_radial(self, r, derivative=False):
        r = np.ascontiguousarray(r, np.float64)
        N = self.info("max_radial")
        L1 = self.info("max_angular") + 1
        R = np.zeros((len(r), N, L1))
        dR = np.zeros_like(R) if derivative else None
        check(lib().rgpu_eval_radial(self._p, len(r), ptr(r), ptr(R), ptr(dR)))
        return (R, dR) if derivative else R

    def eval_ylm(self, dirs, derivative=False):
        d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        L1 = self.info("max_angular") + 1
        Y = np.zeros((len(d), L1 * L1))
        dY = np.zeros((len(d), 3, L1 * L1)) if derivative else None
        check(lib().rgpu_eval_ylm(self._p, len(d), ptr(d), ptr(Y), ptr(dY)))
        return (Y, dY) if derivative else Y
