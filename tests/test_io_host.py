"""Model files in the reference's JSON format (utils/io.py:225-420), on the
host: the reference's own fixture simple_gap_model.json (a KRR model of
methane dimers: SphericalInvariants with RadialScaling, 30 sparse points,
GAP zeta 2) loads into this package's objects, and a model written by
dump_obj reads back unchanged.  The device context is replaced by a stand-in
that only records the hypers (no GPU here)."""
import json
import os

import numpy as np

from conftest import GOLDEN

MODEL = os.path.join(GOLDEN, "simple_gap_model.json")


class _HostContext:
    def __init__(self, hypers_json):
        self.hypers = json.loads(hypers_json)

    def set_precision(self, p):
        pass


def _patch(monkeypatch):
    from librascal_amd import representations as R
    monkeypatch.setattr(R._capi, "Context", _HostContext)


def test_load_reference_model_fixture(monkeypatch):
    _patch(monkeypatch)
    from librascal_amd.io import load_obj
    from librascal_amd.models import KRR, Kernel, SparsePoints
    from librascal_amd.representations import SphericalInvariants
    raw = json.load(open(MODEL))
    model = load_obj(MODEL)
    assert isinstance(model, KRR)
    assert isinstance(model.kernel, Kernel) and model.kernel.zeta == 2
    assert model.kernel.kernel_type == "Sparse" and model.kernel.target_type == "Structure"
    rep = model.get_representation_calculator()
    assert isinstance(rep, SphericalInvariants)
    fc = rep.hypers["cutoff_function"]
    assert fc["type"] == "RadialScaling" and fc["cutoff"]["value"] == 7.5
    assert rep.hypers["max_radial"] == 4 and rep.hypers["max_angular"] == 2
    sp = model.X_train
    assert isinstance(sp, SparsePoints) and sp.size() == 30
    assert sorted(set(sp.numbers.tolist())) == [1, 6]
    assert np.all(sp.numbers[:20] == 1) and np.all(sp.numbers[20:] == 6)
    T = sp.values_for([1, 6])
    assert T.shape == (30, 3 * 48)
    # every block equals the JSON values of its species and key
    data = raw["init_params"]["X_train"]["data"]["sparse_points"]
    keys = [tuple(k) for k in data["keys"]]
    for s, by_key in data["values"]:
        rows = np.where(sp.numbers == s)[0]
        for key, vals in by_key:
            k = keys.index(tuple(key))
            ref = np.asarray(vals).reshape(len(rows), 48)
            assert np.array_equal(T[rows, k * 48:(k + 1) * 48], ref)
    w = np.asarray(raw["init_params"]["weights"][1])
    assert np.array_equal(model.weights, w)
    assert model.self_contributions == {1: 0.0, 6: 0.0}


def test_model_round_trip(monkeypatch, tmp_path):
    _patch(monkeypatch)
    from librascal_amd.io import dump_obj, load_obj
    model = load_obj(MODEL)
    fn = str(tmp_path / "model.json")
    dump_obj(fn, model)
    again = load_obj(fn)
    assert np.array_equal(again.weights, model.weights)
    assert np.array_equal(again.X_train.values_for([1, 6]), model.X_train.values_for([1, 6]))
    assert np.array_equal(again.X_train.numbers, model.X_train.numbers)
    assert again.get_representation_calculator().hypers == model.get_representation_calculator().hypers
    assert again.kernel.zeta == model.kernel.zeta


def test_large_arrays_in_npy_files(monkeypatch, tmp_path):
    """Weights stored as a separate .npy file next to the JSON (the
    reference's layout for large arrays) are memory-mapped, never unpickled."""
    _patch(monkeypatch)
    from librascal_amd.io import load_obj
    raw = json.load(open(MODEL))
    w = np.asarray(raw["init_params"]["weights"][1])
    np.save(str(tmp_path / "model-krr-weights.npy"), w)
    raw["init_params"]["weights"] = "model-krr-weights.npy"
    fn = str(tmp_path / "model.json")
    json.dump(raw, open(fn, "w"))
    model = load_obj(fn)
    assert np.array_equal(model.weights, w)
