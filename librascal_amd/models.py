"""Sparse GAP models with the reference's Python API.

Mirrors bindings/rascal/models/sparse_points.py (SparsePoints),
kernels.py:16-170 (Kernel), krr.py:17-214 (KRR.predict / predict_forces /
predict_stress), krr.py:350-477 (train_gap_model) and the module functions
compute_sparse_kernel_gradients /
compute_sparse_kernel_neg_stress (bind_py_models.cc:150-200).  Every kernel
value, energy, force and stress is computed on the GPU through the
calculator's context (C ABI in include/rascal_mi355x.h).

Sparse points are stored per key (a <= b species pair), as the reference's
block-sparse sparse points (sparse_points.hh:144-500); when they meet a batch
with another species list they are re-laid out over that list, with the
blocks of absent keys dropped (they would contribute zero to every dot
product, as in the reference).
"""
import numpy as np

from . import _capi
from .neighbourlist import AtomsList, _as_structure
from .representations import SphericalInvariants


class SparsePoints:
    """Features of selected atomic environments (pseudo points)."""

    def __init__(self, representation):
        if not isinstance(representation, SphericalInvariants):
            raise ValueError("No pseudo point is appropiate for " + str(representation))
        self.representation = representation
        self._chunks = []   # (species tuple, centre numbers [m], features [m][K(species)*inner])
        self._cache = {}
        # bumped on every change: a context holding these points re-uploads
        # them when the version it loaded is no longer current
        self.version = 0

    def extend(self, atoms_list, selected_indices):
        """Add the environments selected_indices[s] (centre indices within
        structure s) of atoms_list, whose representation has been computed
        with this calculator."""
        if not isinstance(atoms_list, AtomsList):
            raise ValueError("atoms_list must be an AtomsList")
        if len(selected_indices) != len(atoms_list):
            raise ValueError("selected_indices needs one list per structure")
        # centres of each structure (all atoms, or those of the centre mask)
        cidx = np.arange(atoms_list.n_atoms)
        if atoms_list.center_atoms_mask is not None:
            cidx = cidx[atoms_list.center_atoms_mask]
        coff = np.searchsorted(cidx, atoms_list.atom_offsets)
        rows = []
        for s, ids in enumerate(selected_indices):
            n_s = coff[s + 1] - coff[s]
            for i in ids:
                if not 0 <= int(i) < n_s:
                    raise IndexError("centre index %d out of range for structure %d" % (i, s))
                rows.append(int(coff[s] + int(i)))
        if not rows:
            return self
        X = atoms_list.get_dense_features(self.representation)
        rows = np.asarray(rows)
        self._chunks.append((tuple(atoms_list.species), atoms_list.numbers[cidx[rows]].copy(),
                             np.ascontiguousarray(X[rows])))
        self._cache.clear()
        self.version += 1
        return self

    @classmethod
    def from_arrays(cls, representation, species, numbers, values):
        """Sparse points given as dense features over the species list
        `species` (one row per point, centre atomic numbers `numbers`)."""
        sp = cls(representation)
        values = np.ascontiguousarray(values, np.float64)
        if values.shape[1] != _feature_size(representation, len(species)):
            raise ValueError("values do not match the species list")
        sp._chunks.append((tuple(sorted(species)), np.asarray(numbers, np.int32).copy(), values))
        sp.version += 1
        return sp

    def size(self):
        return int(sum(len(c[1]) for c in self._chunks))

    def __len__(self):
        return self.size()

    @property
    def numbers(self):
        return np.concatenate([c[1] for c in self._chunks]) if self._chunks else np.zeros(0, np.int32)

    @property
    def all_species(self):
        out = set()
        for sp, nums, _ in self._chunks:
            out |= set(sp)
        return sorted(out)

    def values_for(self, species):
        """Features of all points laid out over the species list `species`."""
        key = tuple(sorted(species))
        if key in self._cache:
            return self._cache[key]
        rep = self.representation
        if _sparsified(rep):
            # one key holding the selected coefficients: no relayout
            vals = (np.concatenate([X for _, _, X in self._chunks]) if self._chunks
                    else np.zeros((0, _feature_size(rep, 0))))
            self._cache[key] = vals
            return vals
        inner = rep.get_num_coefficients(1)
        out_keys = rep.get_keys(key)
        parts = []
        for sp, nums, X in self._chunks:
            src = {k: i for i, k in enumerate(rep.get_keys(sp))}
            Y = np.zeros((len(nums), len(out_keys), inner))
            Xr = X.reshape(len(nums), -1, inner)
            for j, k in enumerate(out_keys):
                if k in src:
                    Y[:, j] = Xr[:, src[k]]
            parts.append(Y.reshape(len(nums), -1))
        vals = np.concatenate(parts) if parts else np.zeros((0, len(out_keys) * inner))
        self._cache[key] = vals
        return vals

    def get_features(self):
        """Dense features over the species of all the points."""
        return self.values_for(self.all_species)


def _sparsified(rep):
    return rep.hypers.get("coefficient_subselection") is not None


def _feature_size(rep, n_species):
    """Features per centre: the selected coefficients of a sparsified power
    spectrum, else the dense layout over n_species species."""
    if _sparsified(rep):
        return len(rep.hypers["coefficient_subselection"]["a"])
    return rep.get_num_coefficients(n_species)


def _set_model(rep, sparse_points, zeta, weights):
    """Load sparse points (and weights) into the calculator's context for the
    batch currently loaded, skipping the upload when nothing changed."""
    ctx = rep._context
    loaded = rep._loaded
    species = loaded.species if isinstance(loaded, AtomsList) else list(loaded)
    wkey = None if weights is None else np.ascontiguousarray(weights, np.float64).tobytes()
    # the token holds the points object itself (not its id, which Python
    # reuses once an object is freed) and its version
    token = (sparse_points, sparse_points.version, tuple(species), int(zeta), wkey)
    old = rep._sparse_token
    if (old is not None and old[0] is sparse_points and old[1:] == token[1:]):
        return
    ctx.set_sparse_points(zeta, sparse_points.numbers, sparse_points.values_for(species),
                          None if weights is None else np.asarray(weights, np.float64).ravel())
    rep._sparse_token = token


class Kernel:
    """Sparse GAP kernel (X . T)^zeta with the reference's arguments
    (kernels.py:61-108).  Only kernel_type "Sparse" with name "GAP" is
    provided on the GPU; full (dense) kernels are outside this library."""

    def __init__(self, representation, name="Cosine", kernel_type="Full", target_type="Structure",
                 **kwargs):
        if name in ("Cosine", "GAP") and "zeta" in kwargs:
            zeta = kwargs["zeta"]
            if not (isinstance(zeta, int) and zeta > 0):
                raise ValueError("The given zeta has to be a positive integer.")
        else:
            raise RuntimeError("Kernel name must be one of: Cosine, GAP.")
        if target_type not in ("Structure", "Atom"):
            raise ValueError("target_type must be 'Structure' or 'Atom'")
        self._rep = representation
        self.name = name
        self.kernel_type = kernel_type
        self.target_type = target_type
        self.zeta = int(kwargs["zeta"])
        self._kwargs = kwargs

    def _check_sparse(self):
        if self.kernel_type != "Sparse" or self.name != "GAP":
            raise NotImplementedError(
                "only the sparse GAP kernel (name='GAP', kernel_type='Sparse') runs on this "
                "library")

    def __call__(self, X, Y=None, grad=(False, False), compute_neg_stress=False):
        self._check_sparse()
        rep = self._rep
        if Y is None and tuple(grad) == (False, False):
            if not isinstance(X, SparsePoints):
                raise NotImplementedError("a sparse kernel between two batches is not defined")
            # K_MM
            if not (isinstance(rep._loaded, AtomsList)
                    and set(X.all_species) <= set(rep._loaded.species)):
                rep._load_species(X.all_species)
            _set_model(rep, X, self.zeta, None)
            return rep._context.gap_kmm()
        if tuple(grad) == (True, False) and isinstance(Y, SparsePoints):
            # K_NM gradient rows (SparseKernel::compute_derivative), in FP64
            if not isinstance(X, AtomsList):
                raise ValueError("X must be an AtomsList")
            X._require(rep)
            _set_model(rep, Y, self.zeta, None)
            ctx = rep._context
            prec = ctx.info("precision")
            if prec != _capi.FP64:
                ctx.set_precision(_capi.FP64)
            try:
                return ctx.gap_knm_gradient(compute_neg_stress)
            finally:
                if prec != _capi.FP64:
                    ctx.set_precision(prec)
        if tuple(grad) != (False, False) or not isinstance(Y, SparsePoints):
            raise NotImplementedError("The configuration: {} is not implemented for kernel {} "
                                      "in {} mode.".format(grad, self.name, self.kernel_type))
        if not isinstance(X, AtomsList):
            raise ValueError("X must be an AtomsList")
        X._require(rep)
        _set_model(rep, Y, self.zeta, None)
        return rep._context.gap_knm(self.target_type)


def _predict_raw(rep, kernel, managers, sparse_points, weights):
    managers._require(rep)
    _set_model(rep, sparse_points, kernel.zeta, weights)
    return rep._context


def compute_sparse_kernel_gradients(rep, kernel, managers, sparse_points, weights):
    """dE/dr_i [n_atoms][3] of E = sum_m w_m K_Nm (sparse_kernel_predict.hh:206-288)."""
    ctx = _predict_raw(rep, kernel, managers, sparse_points, weights)
    _, g, _ = ctx.gap_predict(energy=False, gradient=True, neg_stress=False)
    return g


def compute_sparse_kernel_neg_stress(rep, kernel, managers, sparse_points, weights):
    """Negative stress [n_struct][6] (Voigt xx, yy, zz, yz, xz, xy) of the
    model energy (sparse_kernel_predict.hh:291-377)."""
    ctx = _predict_raw(rep, kernel, managers, sparse_points, weights)
    _, _, s = ctx.gap_predict(energy=False, gradient=False, neg_stress=True)
    return s


class KRR:
    """Sparse kernel ridge regression model (krr.py:17-214)."""

    def __init__(self, weights, kernel, X_train, self_contributions, description="KRR potential model",
                 units=None):
        self.weights = np.asarray(weights, np.float64)
        self.kernel = kernel
        self.X_train = X_train
        self.self_contributions = self_contributions
        self.target_type = kernel.target_type
        self.description = description
        units = dict(units or {})
        units.setdefault("energy", "eV")
        units.setdefault("length", "AA")
        self.units = units

    def _get_property_baseline(self, managers):
        """Self contributions of the centres (krr.py:68-90)."""
        mask = managers.center_atoms_mask
        nums = np.asarray(managers.numbers).ravel()
        uniq = np.unique(nums)
        table = np.array([self.self_contributions[int(z)] for z in uniq], np.float64)
        per_atom = table[np.searchsorted(uniq, nums)] if nums.size else np.zeros(0)
        if self.target_type == "Structure":
            if mask is not None:
                per_atom = np.where(np.asarray(mask, bool), per_atom, 0.0)
            off = np.asarray(managers.atom_offsets, np.int64)
            Y0 = np.zeros(len(off) - 1)
            filled = off[1:] > off[:-1]
            if per_atom.size and filled.any():
                sums = np.add.reduceat(per_atom, np.minimum(off[:-1], per_atom.size - 1))
                Y0[filled] = sums[filled]
            return Y0
        return per_atom if mask is None else per_atom[np.asarray(mask, bool)]

    def predict(self, managers, KNM=None):
        rep = self.kernel._rep
        if KNM is None:
            if self.target_type == "Structure":
                ctx = _predict_raw(rep, self.kernel, managers, self.X_train, self.weights)
                e, _, _ = ctx.gap_predict(energy=True, gradient=False, neg_stress=False)
                pred = e
            else:
                pred = self.kernel(managers, self.X_train) @ self.weights
        else:
            if len(managers) != KNM.shape[0] and self.target_type == "Structure":
                raise ValueError("KNM size mismatch {}!={}".format(len(managers), KNM.shape[0]))
            if self.X_train.size() != KNM.shape[1]:
                raise ValueError("KNM size mismatch {}!={}".format(self.X_train.size(), KNM.shape[1]))
            pred = KNM @ self.weights
        return self._get_property_baseline(managers) + np.asarray(pred).reshape(-1)

    def predict_forces(self, managers, KNM=None):
        if self.kernel.kernel_type != "Sparse":
            raise NotImplementedError(
                "force prediction only implemented for kernels with kernel_type=='Sparse'")
        if KNM is None:
            g = compute_sparse_kernel_gradients(self.kernel._rep, self.kernel, managers,
                                                self.X_train, self.weights)
        else:
            if 3 * managers.n_atoms != KNM.shape[0]:
                raise ValueError("KNM size mismatch {}!={}".format(3 * managers.n_atoms,
                                                                   KNM.shape[0]))
            g = (KNM @ self.weights).reshape(-1, 3)
        return -g

    def predict_stress(self, managers, KNM=None):
        if self.kernel.kernel_type != "Sparse":
            raise NotImplementedError(
                "stress prediction only implemented for kernels with kernel_type=='Sparse'")
        if KNM is None:
            s = compute_sparse_kernel_neg_stress(self.kernel._rep, self.kernel, managers,
                                                 self.X_train, self.weights)
        else:
            if 6 * len(managers) != KNM.shape[0]:
                raise ValueError("KNM size mismatch {}!={}".format(6 * len(managers), KNM.shape[0]))
            s = (KNM @ self.weights).reshape(len(managers), 6)
        return -s

    def predict_energy_forces_stress(self, managers):
        """predict, predict_forces and predict_stress of a Structure-target
        sparse model from one GPU prediction (energy, gradient and virial
        together)."""
        if self.kernel.kernel_type != "Sparse" or self.target_type != "Structure":
            raise NotImplementedError("one-pass prediction needs a sparse Structure kernel")
        ctx = _predict_raw(self.kernel._rep, self.kernel, managers, self.X_train, self.weights)
        e, g, s = ctx.gap_predict(energy=True, gradient=True, neg_stress=True)
        return self._get_property_baseline(managers) + e, -g, -s

    def get_weights(self):
        return self.weights

    def get_representation_calculator(self):
        return self.kernel._rep


def compute_KNM(frames, X_sparse, kernel, soap):
    """GAP kernel of structures against the sparse points, stacked with its
    position gradients (krr.py:297-348): rows [n_frames] of summed
    (Structure) kernels, then 3 n_atoms gradient rows, frame by frame.  All
    frames are computed as one batch on the GPU (the reference loops over
    them); a tqdm-wrapped list is accepted."""
    if hasattr(frames, "iterable"):
        frames = frames.iterable
    frames = list(frames)
    al = soap.transform(frames)
    if kernel.target_type != "Structure":
        raise ValueError("compute_KNM builds Structure kernel rows")
    en = kernel(al, X_sparse)
    grad = kernel(al, X_sparse, grad=(True, False))
    return np.vstack([en, grad])


def _gram_on_device(A):
    """A^T A in FP64 on the GPU (the FP64 matrix-core GEMM of the GAP path,
    rgpu_gram_f64)."""
    return _capi.gram_f64(A)


def train_gap_model(kernel, frames, KNM_, X_sparse, y_train, self_contributions, grad_train=None,
                    lambdas=None, jitter=1e-8):
    """Sparse GAP training (krr.py:350-477): the weights solve
    (K_MM + K_MN Lambda^-2 K_NM) alpha = K_MN Lambda^-2 y, with Lambda^-1 =
    lambdas[0] / sigma_y * sqrt(N_atoms) on the rows of the per-structure
    property (the first len(y_train) rows of KNM_) and lambdas[1] / sigma_y on
    the gradient rows (grad_train, e.g. minus the forces), sigma_y the
    standard deviation of the property after the self contributions are
    subtracted; jitter on the diagonal of K_MM.  K_MM comes from the GPU
    kernel, K_NM^T K_NM from an FP64 GEMM on the GPU; the M x M solve is
    numpy's lstsq, as in the reference.  Returns a KRR model."""
    KMM = kernel(X_sparse)
    Y = np.asarray(y_train, np.float64).reshape((-1, 1)).copy()
    KNM = np.array(KNM_, np.float64, copy=True)
    n_centers = Y.shape[0]
    Natoms = np.zeros(n_centers)
    Y0 = np.zeros((n_centers, 1))
    for iframe, frame in enumerate(frames):
        numbers = _as_structure(frame)["numbers"]
        Natoms[iframe] = len(numbers)
        for sp in numbers:
            Y0[iframe] += self_contributions[int(sp)]
    Y = Y - Y0
    delta = np.std(Y)
    # lambdas[0] is per atom, hence sqrt(N_atoms) on the property rows
    KNM[:n_centers] /= lambdas[0] / delta * np.sqrt(Natoms)[:, None]
    Y /= lambdas[0] / delta * np.sqrt(Natoms)[:, None]
    if grad_train is not None:
        KNM[n_centers:] /= lambdas[1] / delta
        F = np.asarray(grad_train, np.float64).reshape((-1, 1)).copy()
        F /= lambdas[1] / delta
        Y = np.vstack([Y, F])
    KMM[np.diag_indices_from(KMM)] += jitter
    K = KMM + _gram_on_device(KNM)
    Y = KNM.T @ Y
    weights = np.linalg.lstsq(K, Y, rcond=None)[0]
    return KRR(weights, kernel, X_sparse, self_contributions)


__all__ = ["SparsePoints", "Kernel", "KRR", "compute_sparse_kernel_gradients",
           "compute_sparse_kernel_neg_stress", "compute_KNM", "train_gap_model"]
