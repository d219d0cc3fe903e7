"""The allocation-free GAP energy entry point (rgpu_gap_energy_noalloc)
through the C ABI, against rgpu_gap_predict's energies of the same batch
(the same kernels summing in the same order: equal), on two calls in a
row, the second with a smaller batch, in both precisions."""
import ctypes as C
import json

import numpy as np
import pytest

from conftest import soap_hypers
from librascal_amd import _capi
from librascal_amd.synthetic import pack, periodic_box, triclinic_cells

pytestmark = pytest.mark.gpu


def _batch(sts, precision, zeta=2, per_species=12, seed=0):
    ctx = _capi.Context(json.dumps(soap_hypers(rc=4.0, N=6, L=4, sigma=0.5, spline=1e-8)))
    ctx.set_precision(precision)
    off, pos, num, cells, pbc = pack(sts)
    ctx.set_structures(off, pos, num, cells, pbc, sorted(set(num.tolist())))
    X, _ = ctx.features()
    rng = np.random.default_rng(seed)
    spec, vals = [], []
    for t in sorted(set(num.tolist())):
        pick = rng.choice(np.where(num == t)[0], size=per_species)
        spec += [t] * per_species
        vals.append(X[pick])
    ctx.set_sparse_points(zeta, np.asarray(spec, np.int32), np.concatenate(vals),
                          rng.normal(size=len(spec)))
    return ctx


def _noalloc(ctx, n_struct):
    import torch
    out = torch.full((n_struct,), np.nan, dtype=torch.float64, device="cuda")
    _capi.check(_capi.lib().rgpu_gap_energy_noalloc(ctx.handle, C.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", [_capi.FP64, _capi.MIXED])
def test_noalloc_energy_matches_gap_predict(precision):
    big = triclinic_cells(n_cells=5, n_min=60, n_max=120, seed=31)
    small = periodic_box(96, 0.05, (1, 8), d_min=1.0, seed=32)
    for sts in (big, small):
        # 150 points per species: several column tiles of the energy GEMM
        # (partial sums kept in the scratch array)
        ctx = _batch(sts, precision, per_species=150)
        e_ref, _, _ = ctx.gap_predict()  # sizes every buffer of the batch
        for _ in range(2):
            e = _noalloc(ctx, len(sts))
            assert np.array_equal(e, e_ref)
        # moved atoms: the neighbour list and representation are rebuilt
        # within the existing buffers
        off, pos, _, _, _ = pack(sts)
        ctx.update_positions(pos)
        e2 = _noalloc(ctx, len(sts))
        assert np.array_equal(e2, e_ref)


def test_noalloc_energy_refusals():
    sts = periodic_box(64, 0.05, (1, 8), d_min=1.0, seed=33)
    ctx = _batch(sts, _capi.FP64)
    host = np.zeros(1)
    with pytest.raises(_capi.RascalGPUError, match="device pointer"):
        _capi.check(_capi.lib().rgpu_gap_energy_noalloc(ctx.handle, _capi.ptr(host)))
    # a context whose buffers were never sized for a prediction refuses
    # instead of allocating
    fresh = _capi.Context(json.dumps(soap_hypers(rc=4.0, N=6, L=4, sigma=0.5, spline=1e-8)))
    off, pos, num, cells, pbc = pack(sts)
    fresh.set_structures(off, pos, num, cells, pbc, [1, 8])
    X, _ = ctx.features()
    fresh.set_sparse_points(2, np.array([1, 8], np.int32), X[:2], np.ones(2))
    with pytest.raises(_capi.RascalGPUError, match="allocates no device memory"):
        _noalloc(fresh, 1)
    # the refused call left the context usable
    e, _, _ = fresh.gap_predict()
    assert np.isfinite(e).all()
    assert np.array_equal(_noalloc(fresh, 1), e)
