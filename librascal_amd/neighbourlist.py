"""AtomsList: a batch of structures with its computed representation.

Mirrors bindings/rascal/neighbourlist/structure_manager.py:35-330 for the
SOAP path: structure conversion (ase.Atoms-like objects or dicts), centre
masks (center_atoms_mask, :337-340, and mask_center_atoms_by_id /
_by_species, :351-470), the non-periodic bounding-box cell
(sanitize_non_periodic_structure, :260-286), get_features /
get_features_gradient / get_gradients_info / get_representation_info with
the reference's dense layouts.
"""
import numpy as np

from . import _capi
from .synthetic import pack


def _as_structure(frame):
    """dict(positions, numbers, cell (columns), pbc, center_atoms_mask) from
    an ase.Atoms-like object (mask from its arrays, cell rows transposed to
    columns) or a dict (optional "center_atoms_mask")."""
    if isinstance(frame, dict):
        st = dict(positions=np.asarray(frame["positions"], np.float64).reshape(-1, 3),
                  numbers=np.asarray(frame.get("numbers", frame.get("atom_types")),
                                     np.int32).ravel(),
                  cell=np.asarray(frame["cell"], np.float64).reshape(3, 3),
                  pbc=np.asarray(frame["pbc"], np.int32).ravel())
        mask = frame.get("center_atoms_mask")
    elif all(hasattr(frame, a) for a in ("get_cell", "get_positions", "get_atomic_numbers",
                                        "get_pbc")):
        st = dict(positions=np.asarray(frame.get_positions(), np.float64),
                  numbers=np.asarray(frame.get_atomic_numbers(), np.int32),
                  cell=np.asarray(frame.get_cell(), np.float64).T.copy(),
                  pbc=np.asarray(frame.get_pbc(), np.int32))
        mask = getattr(frame, "arrays", {}).get("center_atoms_mask")
    else:
        raise RuntimeError("Cannot convert structure of type {}".format(type(frame)))
    if mask is None:
        mask = np.ones(len(st["numbers"]), bool)
    mask = np.asarray(mask, bool).ravel()
    if mask.shape[0] != len(st["numbers"]):
        raise ValueError("center_atoms_mask needs one entry per atom")
    st["center_atoms_mask"] = mask
    return sanitize_non_periodic_structure(st)


def _mask_of(frame):
    key = "center_atoms_mask"
    if isinstance(frame, dict):
        return frame.get(key)
    return getattr(frame, "arrays", {}).get(key)


def _set_mask(frame, mask):
    if isinstance(frame, dict):
        frame["center_atoms_mask"] = mask
    else:
        frame.arrays["center_atoms_mask"] = mask


def mask_center_atoms_by_id(frame, id_select=None, id_blacklist=None):
    """Centre mask by atom index (structure_manager.py:351-396): without a
    mask, start from all atoms (or none when id_select is given); then
    select id_select and exclude id_blacklist.  Modifies the frame (a
    structure dict or an ase.Atoms-like object with .arrays)."""
    mask = _mask_of(frame)
    n = len(frame["numbers"]) if isinstance(frame, dict) else len(frame)
    if mask is None:
        mask = np.zeros(n, bool) if id_select is not None else np.ones(n, bool)
    mask = np.asarray(mask, bool).copy()
    if id_select is not None:
        mask[id_select] = True
    if id_blacklist is not None:
        mask[id_blacklist] = False
    _set_mask(frame, mask)


def mask_center_atoms_by_species(frame, species_select=[], species_blacklist=[]):
    """Centre mask by atomic number (structure_manager.py:399-470): select
    the atoms of species_select (combined with an existing mask by 'or'),
    then exclude those of species_blacklist."""
    numbers = np.asarray(frame["numbers"] if isinstance(frame, dict)
                         else frame.get_atomic_numbers())
    if any(isinstance(x, str) for x in list(species_select) + list(species_blacklist)):
        raise ValueError("species must be given as atomic numbers")
    mask = _mask_of(frame)
    if mask is None:
        mask = np.zeros(len(numbers), bool) if len(species_select) else np.ones(len(numbers), bool)
    mask = np.asarray(mask, bool).copy()
    if len(species_select):
        mask |= np.isin(numbers, species_select)
    if len(species_blacklist):
        mask &= ~np.isin(numbers, species_blacklist)
    _set_mask(frame, mask)


def sanitize_non_periodic_structure(structure):
    """Zero cell + no periodicity: bounding box x1.05 and recentred atoms
    (structure_manager.py:260-286)."""
    if np.all(np.asarray(structure["pbc"]) == 0):
        cell = np.asarray(structure["cell"])
        if np.allclose(cell, np.zeros((3, 3))):
            pos = np.asarray(structure["positions"], np.float64)
            lo, hi = pos.min(axis=0), pos.max(axis=0)
            lengths = (hi - lo) * 1.05
            com = pos.mean(axis=0)
            structure = dict(structure)
            structure["positions"] = pos + (0.5 * lengths - com)[None, :]
            structure["cell"] = np.diag(lengths)
    return structure


class AtomsList:
    """A collection of structures and their representation on the GPU."""

    def __init__(self, frames, nl_options=None, start=None, length=None):
        if hasattr(frames, "get_positions") or isinstance(frames, dict):
            frames = [frames]
        frames = list(frames)
        if start is not None:
            frames = frames[start:start + length if length is not None else None]
        self.structures = [_as_structure(f) for f in frames]
        self.nl_options = nl_options
        (self.atom_offsets, self.positions, self.numbers, self.cells, self.pbc) = \
            pack(self.structures)
        mask = (np.concatenate([s["center_atoms_mask"] for s in self.structures])
                if self.structures else np.zeros(0, bool))
        # None when every atom is a centre
        self.center_atoms_mask = None if mask.all() else mask
        self.calculator = None
        self.species = None
        # incremented whenever set_positions changes the positions, so that a
        # result computed for earlier positions can tell that it is stale
        self.positions_version = 0

    def __len__(self):
        return len(self.structures)

    @property
    def n_atoms(self):
        return int(self.atom_offsets[-1])

    @property
    def n_centres(self):
        if self.center_atoms_mask is None:
            return self.n_atoms
        return int(np.count_nonzero(self.center_atoms_mask))

    def compute(self, calculator, species=None):
        sp = set(self.numbers.tolist()) | set(species or [])
        if calculator.hypers.get("expansion_by_species_method") == "user defined":
            sp |= set(int(z) for z in calculator.hypers.get("global_species", []))
        self.species = sorted(sp)
        self.calculator = calculator
        calculator._loaded = None
        calculator._load(self)
        return self

    def _upload(self, calculator):
        calculator._context.set_structures(self.atom_offsets, self.positions, self.numbers,
                                           self.cells, self.pbc, self.species,
                                           self.center_atoms_mask)

    def _require(self, calculator):
        if calculator is None:
            calculator = self.calculator
        if calculator is None or calculator is not self.calculator:
            raise ValueError("the representation of this AtomsList was computed with another "
                             "calculator")
        calculator._load(self)
        return calculator

    def set_positions(self, calculator, positions, cells=None):
        """New positions [n_atoms][3] (and optionally cells [n_struct][3][3],
        lattice vectors as columns) for the same atoms, e.g. one MD step:
        periodic directions are wrapped into the cell, then the neighbour
        list and the representation are rebuilt on the GPU with the batch's
        buffers kept (the reference's manager.update, structure_manager.py
        :231-258, followed by compute)."""
        calculator = self._require(calculator)
        from .synthetic import wrap_positions
        pos = np.array(positions, np.float64).reshape(-1, 3)
        if pos.shape[0] != self.n_atoms:
            raise ValueError("positions must have one row per atom of the batch")
        if cells is not None:
            self.cells = np.ascontiguousarray(np.asarray(cells, np.float64).reshape(-1, 3, 3))
        off = self.atom_offsets
        for s in range(len(off) - 1):
            if np.any(self.pbc[s]):
                pos[off[s]:off[s + 1]] = wrap_positions(pos[off[s]:off[s + 1]],
                                                        self.cells[s].reshape(3, 3),
                                                        self.pbc[s])
        self.positions = np.ascontiguousarray(pos)
        for k, st in enumerate(self.structures):
            st["positions"] = self.positions[off[k]:off[k + 1]]
            if cells is not None:
                st["cell"] = self.cells[k].reshape(3, 3)
        self.positions_version += 1
        calculator._context.update_positions(self.positions,
                                             None if cells is None else self.cells)
        return self

    @property
    def managers(self):
        """The reference's ManagerCollection handle; here the list itself."""
        return self

    def _present_keys(self, calculator):
        """Keys present in at least one centre (dense export drops the rest,
        property_block_sparse.hh:911-932 with the keys of the collection)."""
        mask = calculator._context.key_masks()
        S = len(self.species)
        present_species = np.zeros((len(mask), S), bool)
        for s in range(S):
            present_species[:, s] = (mask >> np.uint64(s)) & np.uint64(1)
        if "soap_type" not in calculator.hypers:
            return [s for s in range(S) if present_species[:, s].any()], present_species
        keys = []
        k = 0
        for a in range(S):
            for b in range(a, S):
                if (present_species[:, a] & present_species[:, b]).any():
                    keys.append(k)
                k += 1
        return keys, present_species

    def get_features(self, calculator=None, species=None):
        """Dense feature matrix (n_centres, n_keys * inner) with the keys of
        the batch (or of the requested species list)."""
        calculator = self._require(calculator)
        ctx = calculator._context
        N = calculator.hypers["max_radial"]
        L1 = calculator.hypers["max_angular"] + 1
        if "soap_type" in calculator.hypers:
            X, _ = ctx.features()
            inner = N * N * L1
        else:
            c, _ = ctx.expansion()
            X = c.reshape(c.shape[0], -1)
            inner = N * L1 * L1
        return self._select_keys(calculator, X, inner, species)

    def _select_keys(self, calculator, X, inner, species):
        """Columns of the keys present in the batch (or of the species list
        given) from a dense matrix over all keys of the batch species."""
        if calculator.hypers.get("coefficient_subselection") is not None:
            return X  # one sparsified key (0, 0) holding the selection
        nk = X.shape[1] // inner
        X = X.reshape(X.shape[0], nk, inner)
        if species is None:
            keys, _ = self._present_keys(calculator)
            return np.ascontiguousarray(X[:, keys, :].reshape(X.shape[0], -1))
        return self._features_for_species(calculator, X, species)

    def get_features_gradient(self, calculator=None, species=None):
        """Gradient of the features with respect to the atomic positions:
        (3 * (n_pairs incl. self pairs), n_features), rows x, y, z per pair in
        the order of get_gradients_info, columns as get_features
        (structure_manager.py:140-181).  Computed in FP64 on the GPU."""
        calculator = self._require(calculator)
        if self.center_atoms_mask is not None:
            raise RuntimeError("Can't compute spherical expansion gradients with masked center "
                               "atoms")
        if "soap_type" not in calculator.hypers:
            raise NotImplementedError("feature gradients are provided for SphericalInvariants")
        ctx = calculator._context
        prec = ctx.info("precision")
        if prec != _capi.FP64:
            ctx.set_precision(_capi.FP64)
        try:
            G = ctx.features_gradient()
        finally:
            if prec != _capi.FP64:
                ctx.set_precision(prec)
        N = calculator.hypers["max_radial"]
        L1 = calculator.hypers["max_angular"] + 1
        return self._select_keys(calculator, G, N * N * L1, species)

    def get_representation_info(self):
        """(n_centres, 3) rows [structure, centre index within the structure,
        atomic number] (structure_manager.py:194-204)."""
        idx = np.arange(self.n_atoms)
        if self.center_atoms_mask is not None:
            idx = idx[self.center_atoms_mask]
        sidx = np.searchsorted(self.atom_offsets, idx, side="right") - 1
        return np.stack([sidx, idx - self.atom_offsets[sidx], self.numbers[idx]],
                        axis=1).astype(np.int32)

    def _features_for_species(self, calculator, X, species):
        sp = list(self.species)
        out_keys = calculator.get_keys(species)
        inner = X.shape[2]
        out = np.zeros((X.shape[0], len(out_keys), inner))
        all_keys = calculator.get_keys(sp)
        index = {k: i for i, k in enumerate(all_keys)}
        for i, k in enumerate(out_keys):
            if k in index:
                out[:, i, :] = X[:, index[k], :]
        return out.reshape(X.shape[0], -1)

    def get_dense_features(self, calculator=None):
        """Features over all (a<=b) keys of the batch species, no dropping."""
        calculator = self._require(calculator)
        X, _ = calculator._context.features()
        return X

    def get_neighbours(self, calculator=None):
        calculator = self._require(calculator)
        return calculator._context.neighbours()

    def get_gradients_info(self, calculator=None):
        """(n_pairs, 5) rows [structure, i, j, type_i, type_j] with centre and
        neighbour indices within the structure, self pair first
        (bind_py_structure_manager.cc:622-744)."""
        calculator = self._require(calculator)
        if self.center_atoms_mask is not None:
            raise RuntimeError("Can't compute spherical expansion gradients with masked center "
                               "atoms")
        off, pj, _ = calculator._context.neighbours(with_vectors=False)
        n = len(off) - 1
        centre = np.repeat(np.arange(n), np.diff(off))
        sidx = np.searchsorted(self.atom_offsets, centre, side="right") - 1
        a0 = self.atom_offsets[sidx]
        info = np.stack([sidx, centre - a0, pj - a0, self.numbers[centre], self.numbers[pj]], axis=1)
        return info.astype(np.int32)
