"""Host logic of the Python model layer (no device calls): sparse-point
re-layout over species lists, argument checks of Kernel, and KRR baselines."""
import json

import numpy as np
import pytest


class _NoDeviceContext:
    def __init__(self, hypers_json):
        self.hypers = json.loads(hypers_json)

    def set_precision(self, p):
        pass


@pytest.fixture
def rep(monkeypatch):
    from librascal_amd import representations as R
    monkeypatch.setattr(R._capi, "Context", _NoDeviceContext)
    return R.SphericalInvariants(interaction_cutoff=3.0, cutoff_smooth_width=0.5, max_radial=2,
                                 max_angular=1, gaussian_sigma_type="Constant",
                                 gaussian_sigma_constant=0.4)


def test_sparse_points_relayout_keeps_dot_products(rep):
    """Dense vectors over species (1, 8) re-laid over (1, 6, 8) keep every
    key block; dot products with any batch vector over (1, 6, 8) equal the
    block-sparse dot product (sparse_points.hh:180-225)."""
    from librascal_amd.models import SparsePoints
    inner = rep.get_num_coefficients(1)
    rng = np.random.default_rng(0)
    vals = rng.normal(size=(5, 3 * inner))
    sp = SparsePoints.from_arrays(rep, [1, 8], [1, 8, 8, 1, 1], vals)
    out = sp.values_for([1, 6, 8])
    assert out.shape == (5, 6 * inner)
    keys_old = rep.get_keys([1, 8])
    keys_new = rep.get_keys([1, 6, 8])
    for i, k in enumerate(keys_new):
        blk = out[:, i * inner:(i + 1) * inner]
        if k in keys_old:
            j = keys_old.index(k)
            assert np.array_equal(blk, vals[:, j * inner:(j + 1) * inner])
        else:
            assert not blk.any()
    x = rng.normal(size=6 * inner)
    ref = sum(vals[:, keys_old.index(k) * inner:(keys_old.index(k) + 1) * inner]
              @ x[i * inner:(i + 1) * inner] for i, k in enumerate(keys_new) if k in keys_old)
    assert np.allclose(out @ x, ref)
    # a layout without species 8 drops its keys
    assert sp.values_for([1]).shape == (5, inner)
    assert sp.size() == len(sp) == 5 and sp.all_species == [1, 8]


def test_sparse_points_shape_check(rep):
    from librascal_amd.models import SparsePoints
    with pytest.raises(ValueError):
        SparsePoints.from_arrays(rep, [1, 8], [1], np.zeros((1, 7)))


def test_kernel_argument_checks(rep):
    from librascal_amd.models import Kernel
    with pytest.raises(ValueError):
        Kernel(rep, name="GAP", zeta=1.5, kernel_type="Sparse")
    with pytest.raises(ValueError):
        Kernel(rep, name="GAP", zeta=0, kernel_type="Sparse")
    with pytest.raises(RuntimeError):
        Kernel(rep, name="RBF", zeta=2, kernel_type="Sparse")
    k = Kernel(rep, name="Cosine", zeta=2, kernel_type="Full")
    with pytest.raises(NotImplementedError):
        k(None)
    g = Kernel(rep, name="GAP", zeta=2, kernel_type="Sparse")
    with pytest.raises(NotImplementedError):
        g(None, None, grad=(True, False))


def test_krr_property_baseline(rep):
    from librascal_amd.models import KRR, Kernel
    from librascal_amd.neighbourlist import AtomsList
    frames = [dict(positions=np.zeros((2, 3)) + [[1, 1, 1], [2, 1, 1]], numbers=[1, 8],
                   cell=np.eye(3) * 5, pbc=[1, 1, 1]),
              dict(positions=np.ones((3, 3)), numbers=[8, 8, 1], cell=np.eye(3) * 5, pbc=[1, 1, 1])]
    al = AtomsList(frames)
    sc = {1: -0.5, 8: -2.0}
    kern = Kernel(rep, name="GAP", zeta=2, kernel_type="Sparse", target_type="Structure")
    m = KRR(np.zeros(1), kern, None, sc)
    assert np.allclose(m._get_property_baseline(al), [-2.5, -4.5])
    kern.target_type = "Atom"
    m = KRR(np.zeros(1), kern, None, sc)
    assert np.allclose(m._get_property_baseline(al), [-0.5, -2.0, -2.0, -2.0, -0.5])
