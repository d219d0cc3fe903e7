"""PyTorch autograd bindings of the SOAP + GAP path.

SoapFeatures: positions -> features X (the reference's get_features); its
backward runs the hand-written adjoint on the GPU (dL/dX -> dL/dc -> per-pair
forces, never storing the per-pair gradient tensor).
GapEnergy: positions -> per-structure GAP energies of a sparse model; its
backward applies dL/dE to the forces computed with the energy.

Torch is only plumbing here: positions go to the library as host arrays
(they are validated and wrapped on the host), features and gradients come
back as tensors on the positions' device.  The library's stream is set to
torch's current stream, so device outputs are ordered with torch's work.
"""
import numpy as np
import torch

from .models import _set_model
from .synthetic import wrap_positions


def _sync_stream(calculator, device):
    if device.type == "cuda":
        calculator._context.set_stream(torch.cuda.current_stream(device).cuda_stream)


def _load_positions(calculator, atoms_list, positions):
    """New positions for the batch, wrapped into the cell along periodic
    directions (environments, and so every gradient, are unchanged)."""
    atoms_list._require(calculator)
    pos = positions.detach().to("cpu", torch.float64).numpy().reshape(-1, 3).copy()
    off = atoms_list.atom_offsets
    for s in range(len(off) - 1):
        if np.any(atoms_list.pbc[s]):
            pos[off[s]:off[s + 1]] = wrap_positions(pos[off[s]:off[s + 1]],
                                                    atoms_list.cells[s].reshape(3, 3),
                                                    atoms_list.pbc[s])
    atoms_list.positions = np.ascontiguousarray(pos)
    calculator._context.update_positions(atoms_list.positions)
    calculator._context.compute()


class SoapFeatures(torch.autograd.Function):
    """X = features(positions) for the structures of atoms_list, computed
    with calculator (a SphericalInvariants)."""

    @staticmethod
    def forward(fctx, positions, calculator, atoms_list):
        _sync_stream(calculator, positions.device)
        _load_positions(calculator, atoms_list, positions)
        X, _ = calculator._context.features()
        fctx.calculator = calculator
        fctx.atoms_list = atoms_list
        fctx.save_for_backward(positions)
        return torch.from_numpy(X).to(positions.device)

    @staticmethod
    def backward(fctx, dX):
        calculator, atoms_list = fctx.calculator, fctx.atoms_list
        (positions,) = fctx.saved_tensors
        if calculator._loaded is not atoms_list:
            _load_positions(calculator, atoms_list, positions)
        _sync_stream(calculator, dX.device)
        g, _ = calculator._context.features_backward(
            dX.detach().to(torch.float64).contiguous(), want_stress=False)
        return torch.from_numpy(g).to(positions.device, positions.dtype), None, None


class GapEnergy(torch.autograd.Function):
    """Per-structure energies sum_m w_m K_Nm (without self contributions) of
    the sparse GAP model `model` (a KRR with a GAP Kernel)."""

    @staticmethod
    def forward(fctx, positions, model, atoms_list):
        rep = model.kernel._rep
        _sync_stream(rep, positions.device)
        _load_positions(rep, atoms_list, positions)
        _set_model(rep, model.X_train, model.kernel.zeta, model.weights)
        e, g, _ = rep._context.gap_predict(energy=True, gradient=True, neg_stress=False)
        fctx.save_for_backward(torch.from_numpy(g))
        fctx.offsets = atoms_list.atom_offsets
        fctx.device, fctx.dtype = positions.device, positions.dtype
        return torch.from_numpy(e).to(positions.device)

    @staticmethod
    def backward(fctx, dE):
        (g,) = fctx.saved_tensors
        counts = torch.from_numpy(np.diff(fctx.offsets))
        scale = torch.repeat_interleave(dE.detach().cpu().to(torch.float64), counts)
        return (g * scale[:, None]).to(fctx.device, fctx.dtype), None, None


def soap_features(positions, calculator, atoms_list):
    return SoapFeatures.apply(positions, calculator, atoms_list)


def gap_energy(positions, model, atoms_list):
    return GapEnergy.apply(positions, model, atoms_list)


__all__ = ["SoapFeatures", "GapEnergy", "soap_features", "gap_energy"]
