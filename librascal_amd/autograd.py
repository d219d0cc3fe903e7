"""PyTorch autograd bindings of the SOAP + GAP path.

SoapFeatures: positions -> features X (the reference's get_features); its
backward runs the hand-written adjoint on the GPU (dL/dX -> dL/dc -> per-pair
forces, never storing the per-pair gradient tensor).
GapEnergy: positions -> per-structure GAP energies of a sparse model; its
backward applies dL/dE to the forces computed with the energy.

Torch is only plumbing here.  Positions go to the library as a host array
(n_atoms x 3 doubles: they are validated and wrapped on the host, as the
reference's structure checks do); features, dL/dX, forces and energies stay
on the positions' device: the library writes them straight into torch
tensors.  The library's stream is set to torch's current stream, so device
outputs are ordered with torch's work.
"""
import numpy as np
import torch

from .models import _set_model


def _sync_stream(calculator, device):
    if device.type == "cuda":
        calculator._context.set_stream(torch.cuda.current_stream(device).cuda_stream)


def _load_positions(calculator, atoms_list, positions):
    """New positions for the batch (periodic directions wrapped into the
    cell: environments, and so every gradient, are unchanged) and its
    representation; returns the positions version of atoms_list that the
    context now holds."""
    pos = positions.detach().to("cpu", torch.float64).numpy().reshape(-1, 3)
    atoms_list.set_positions(calculator, pos)
    calculator._context.compute()
    return atoms_list.positions_version


def _out(shape, device):
    return torch.empty(shape, dtype=torch.float64, device=device)


class SoapFeatures(torch.autograd.Function):
    """X = features(positions) for the structures of atoms_list, computed
    with calculator (a SphericalInvariants)."""

    @staticmethod
    def forward(fctx, positions, calculator, atoms_list):
        _sync_stream(calculator, positions.device)
        version = _load_positions(calculator, atoms_list, positions)
        ctx = calculator._context
        X = _out((ctx.info("n_centres"), ctx.info("feature_size")), positions.device)
        ctx.features(out=X)
        fctx.calculator = calculator
        fctx.atoms_list = atoms_list
        fctx.version = version
        fctx.save_for_backward(positions)
        return X

    @staticmethod
    def backward(fctx, dX):
        calculator, atoms_list = fctx.calculator, fctx.atoms_list
        (positions,) = fctx.saved_tensors
        _sync_stream(calculator, dX.device)
        # the context may hold another batch, or this batch at other
        # positions (a later forward): the adjoint needs the expansion of the
        # positions this X was computed from
        if calculator._loaded is not atoms_list or atoms_list.positions_version != fctx.version:
            fctx.version = _load_positions(calculator, atoms_list, positions)
        g = _out((atoms_list.n_centres, 3), dX.device)
        calculator._context.features_backward(dX.detach().to(torch.float64).contiguous(),
                                              want_stress=False, out=g)
        return g.to(positions.device, positions.dtype), None, None


class GapEnergy(torch.autograd.Function):
    """Per-structure energies sum_m w_m K_Nm (without self contributions) of
    the sparse GAP model `model` (a KRR with a GAP Kernel)."""

    @staticmethod
    def forward(fctx, positions, model, atoms_list):
        rep = model.kernel._rep
        _sync_stream(rep, positions.device)
        _load_positions(rep, atoms_list, positions)
        _set_model(rep, model.X_train, model.kernel.zeta, model.weights)
        ctx = rep._context
        e = _out((len(atoms_list),), positions.device)
        g = _out((atoms_list.n_centres, 3), positions.device)
        ctx.gap_predict(energy=True, gradient=True, neg_stress=False, out=(e, g, None))
        fctx.save_for_backward(g)
        fctx.counts = torch.from_numpy(np.diff(atoms_list.atom_offsets)).to(positions.device)
        fctx.dtype = positions.dtype
        return e

    @staticmethod
    def backward(fctx, dE):
        (g,) = fctx.saved_tensors
        scale = torch.repeat_interleave(dE.detach().to(g.device, torch.float64), fctx.counts)
        return (g * scale[:, None]).to(fctx.dtype), None, None


def soap_features(positions, calculator, atoms_list):
    return SoapFeatures.apply(positions, calculator, atoms_list)


def gap_energy(positions, model, atoms_list):
    return GapEnergy.apply(positions, model, atoms_list)


__all__ = ["SoapFeatures", "GapEnergy", "soap_features", "gap_energy"]
