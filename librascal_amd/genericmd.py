"""Generic calculator-style MD driver for a model file (models/genericmd.py
of the reference): load a KRR model saved in the reference's JSON format,
then per MD step take positions and a cell and return energy, forces and
the stress matrix, with the neighbour list, representation and model
evaluated on the GPU (one context reused across steps)."""
import numpy as np

from .io import load_obj
from .neighbourlist import AtomsList, sanitize_non_periodic_structure
from .synthetic import wrap_positions


class GenericMDCalculator:
    """Initialise with a model JSON and a structure template (or atomic
    numbers); calculate() assumes the order and identity of the atoms do not
    change between calls."""

    matrix_indices_in_voigt_notation = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))

    def __init__(self, model_json, is_periodic, structure_template=None, atomic_numbers=None):
        self.model_filename = model_json
        self.model = load_obj(model_json)
        self.representation = self.model.get_representation_calculator()
        self.manager = None
        self.is_periodic = is_periodic
        if structure_template is not None:
            numbers, pbc = _template(structure_template)
            if is_periodic is not None and bool(is_periodic) != bool(np.any(pbc)):
                raise ValueError("Structure template PBC flags: " + str(pbc)
                                 + " incompatible with 'is_periodic' setting")
        elif atomic_numbers is not None:
            numbers = np.asarray(atomic_numbers, np.int32).ravel()
            pbc = np.array([bool(is_periodic)] * 3)
        else:
            raise ValueError("Must specify one of 'structure_template' or 'atomic_numbers'")
        self.numbers = numbers
        self.pbc = np.asarray(pbc, bool).reshape(3)

    def calculate(self, positions, cell_matrix):
        """positions [N][3] and the cell with lattice vectors as rows (zero
        for non-periodic runs), in AA; returns (energy, forces [N][3],
        stress [3][3]) in eV and AA.  The stress matrix is the reference's
        (genericmd.py:123-129): the model's Voigt components (the
        volume-normalised energy gradient with respect to the cell) placed in
        the upper triangle, then the upper triangle including the diagonal
        added transposed, so the off-diagonal entries are mirrored and the
        diagonal entries doubled.

        The first call sets the structure on the GPU; later calls move its
        atoms (AtomsList.set_positions: the neighbour-list device pass
        replayed from a HIP graph, its size check deferred to the end of the
        prediction) and run one prediction for energy, forces and stress."""
        positions = np.asarray(positions, np.float64)
        cell_matrix = np.asarray(cell_matrix, np.float64)
        if positions.shape != (len(self.numbers), 3):
            raise ValueError("Improper shape of positions (is the number of atoms consistent?)")
        if cell_matrix.shape != (3, 3):
            raise ValueError("Improper shape of cell info (expected 3x3 matrix)")
        cell = cell_matrix.T.copy()  # lattice vectors as columns
        if self.manager is None:
            if np.any(self.pbc):
                positions = wrap_positions(positions, cell, self.pbc.astype(np.int32))
            frame = dict(positions=positions, numbers=self.numbers, cell=cell,
                         pbc=self.pbc.astype(np.int32))
            self.representation._context.enable_graphs(True)
            self.manager = self.representation.transform(
                AtomsList([frame], self.representation.nl_options))
        elif np.any(self.pbc):
            self.manager.set_positions(self.representation, positions, cell[None])
        else:
            # non-periodic: the atoms' bounding box (x1.05, atoms recentred)
            # as each step's cell, as for the first step
            st = sanitize_non_periodic_structure(
                dict(positions=positions, numbers=self.numbers, cell=np.zeros((3, 3)),
                     pbc=np.zeros(3, np.int32)))
            self.manager.set_positions(self.representation, st["positions"], st["cell"][None])
        if self.model.kernel.kernel_type == "Sparse" and self.model.target_type == "Structure":
            energy, forces, stress_voigt = self.model.predict_energy_forces_stress(self.manager)
        else:
            energy = self.model.predict(self.manager)
            forces = self.model.predict_forces(self.manager)
            stress_voigt = self.model.predict_stress(self.manager)
        stress = np.zeros((3, 3))
        stress[tuple(zip(*self.matrix_indices_in_voigt_notation))] = np.asarray(stress_voigt).ravel()
        stress += np.triu(stress).T
        return energy, forces, stress


def _template(t):
    """(atomic numbers, pbc) of a structure template: a structure dict, an
    ase.Atoms-like object, or a file ASE can read."""
    if isinstance(t, dict):
        return np.asarray(t["numbers"], np.int32).ravel(), np.asarray(t.get("pbc", [0, 0, 0]), bool)
    if hasattr(t, "get_atomic_numbers"):
        return np.asarray(t.get_atomic_numbers(), np.int32), np.asarray(t.get_pbc(), bool)
    try:
        import ase.io
    except ImportError as exc:
        raise ImportError("reading a structure template file needs ase; pass a structure dict "
                          "or atomic_numbers instead") from exc
    at = ase.io.read(t, 0)
    return np.asarray(at.get_atomic_numbers(), np.int32), np.asarray(at.get_pbc(), bool)
