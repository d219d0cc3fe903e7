"""Model files in the reference's JSON format (utils/io.py:225-420).

An object is stored as {"class_name", "module_name", "version",
"init_params", "data"}; nested objects are stored the same way, small numpy
arrays as ["npy", nested list] and large ones as the name of an .npy file
next to the JSON file.  load_obj() rebuilds SphericalInvariants /
SphericalExpansion, SparsePoints, Kernel and KRR objects of this package from
such a file, so a model trained and saved with the reference predicts here
unchanged; dump_obj() writes the same format.  Arrays are read with
numpy.load(allow_pickle=False) only.
"""
import json
import os

import numpy as np

from .models import KRR, Kernel, SparsePoints
from .representations import SphericalExpansion, SphericalInvariants

CURRENT_VERSION = "0.1"
_KEYS = ("class_name", "module_name", "version", "init_params", "data")


def is_valid_object_dict(d):
    return isinstance(d, dict) and all(k in d for k in _KEYS)


def _is_npy_list(v):
    return isinstance(v, list) and len(v) == 2 and v[0] == "npy"


def _load_npy(data, path):
    """In place: ["npy", list] -> array, "<name>.npy" -> memory-mapped array."""
    for k, v in list(data.items()):
        if isinstance(v, dict):
            _load_npy(v, path)
        elif _is_npy_list(v):
            data[k] = np.asarray(v[1])
        elif isinstance(v, str) and v.endswith(".npy") and os.path.exists(os.path.join(path, v)):
            data[k] = np.load(os.path.join(path, v), mmap_mode="r", allow_pickle=False)


def _sparse_points_from_data(representation, data):
    """Dense per-species rows from the block-sparse JSON of
    SparsePointsBlockSparse (sparse_points.hh:456-500): per centre species,
    per key (a, b), the values of the points listed in `indices`, each
    inner_size = N^2 (L+1) long; keys absent for a point are zero."""
    sp = data["sparse_points"]
    inner = int(sp["inner_size"])
    keys = [tuple(k) for k in sp["keys"]]
    species = sorted({s for k in keys for s in k} | set(int(s) for s in sp["center_species"]))
    all_keys = [(a, b) for i, a in enumerate(species) for b in species[i:]]
    if representation.get_num_coefficients(len(species)) != len(all_keys) * inner:
        raise ValueError("sparse points do not match the representation's hypers")
    counts = dict((int(c[0]), int(c[1])) for c in sp["counters"])
    index_of = dict((int(s), dict((tuple(k), list(ix)) for k, ix in by_key))
                    for s, by_key in sp["indices"])
    numbers, rows = [], []
    for s, by_key in sp["values"]:
        s = int(s)
        T = np.zeros((counts[s], len(all_keys) * inner))
        for key, vals in by_key:
            key = tuple(key)
            col = all_keys.index(key) * inner
            ix = index_of.get(s, {}).get(key, list(range(counts[s])))
            T[np.asarray(ix, int), col:col + inner] = np.asarray(vals, np.float64).reshape(len(ix), inner)
        numbers.append(np.full(counts[s], s, np.int32))
        rows.append(T)
    if not rows:
        return SparsePoints(representation)
    return SparsePoints.from_arrays(representation, species, np.concatenate(numbers),
                                    np.concatenate(rows))


def _sparse_points_to_data(sp_obj):
    species = sp_obj.all_species
    rep = sp_obj.representation
    hp = rep.hypers
    N, L1 = hp["max_radial"], hp["max_angular"] + 1
    inner = N * N * L1
    keys = [[a, b] for i, a in enumerate(species) for b in species[i:]]
    T = sp_obj.values_for(species)
    nums = sp_obj.numbers
    values, indices, counters = [], [], []
    for s in sorted(set(int(z) for z in nums)):
        rows = np.where(nums == s)[0]
        counters.append([s, int(len(rows))])
        values.append([s, [[k, T[rows, j * inner:(j + 1) * inner].ravel().tolist()]
                           for j, k in enumerate(keys)]])
        indices.append([s, [[k, list(range(len(rows)))] for k in keys]])
    return dict(sparse_points=dict(center_species=sorted(set(int(z) for z in nums)),
                                   counters=counters, indices=indices, inner_size=inner,
                                   keys=keys, keys_sp=[[s, keys] for s, _ in counters],
                                   name="SparsePointsBlockSparse", values=values))


def from_dict(d):
    """Rebuild an object (and its nested objects) from its dict form."""
    if not is_valid_object_dict(d):
        raise RuntimeError("not a valid dictionary representation of an object")
    ip = {}
    for k, v in d["init_params"].items():
        if is_valid_object_dict(v):
            v = from_dict(v)
        elif isinstance(v, list):
            v = [from_dict(x) if is_valid_object_dict(x) else x for x in v]
        ip[k] = v
    name = d["class_name"]
    if name == "SphericalInvariants":
        return SphericalInvariants(**ip)
    if name == "SphericalExpansion":
        return SphericalExpansion(**ip)
    if name == "SparsePoints":
        return _sparse_points_from_data(ip["representation"], d["data"])
    if name == "Kernel":
        return Kernel(**ip)
    if name == "KRR":
        sc = ip.get("self_contributions") or {}
        ip["self_contributions"] = dict((int(k), float(v)) for k, v in sc.items())
        return KRR(**ip)
    raise NotImplementedError("objects of class %s are not provided by this library" % name)


def to_dict(obj):
    """The dict form of a SphericalInvariants/Expansion, SparsePoints,
    Kernel or KRR object (arrays as ["npy", list])."""
    def wrap(name, module, init_params, data=None):
        return dict(class_name=name, module_name=module, version=CURRENT_VERSION,
                    init_params=init_params, data=data or {})
    if isinstance(obj, (SphericalInvariants, SphericalExpansion)):
        mod = ("rascal.representations.spherical_invariants" if isinstance(obj, SphericalInvariants)
               else "rascal.representations.spherical_expansion")
        ip = dict(obj._init_params)
        if ip.get("global_species") is None:
            ip["global_species"] = []
        return wrap(type(obj).__name__, mod, ip)
    if isinstance(obj, SparsePoints):
        return wrap("SparsePoints", "rascal.models.sparse_points",
                    dict(representation=to_dict(obj.representation)), _sparse_points_to_data(obj))
    if isinstance(obj, Kernel):
        ip = dict(representation=to_dict(obj._rep), name=obj.name, kernel_type=obj.kernel_type,
                  target_type=obj.target_type, zeta=obj.zeta)
        return wrap("Kernel", "rascal.models.kernels", ip)
    if isinstance(obj, KRR):
        ip = dict(weights=["npy", np.asarray(obj.weights).tolist()], kernel=to_dict(obj.kernel),
                  X_train=to_dict(obj.X_train),
                  self_contributions=dict((str(k), v) for k, v in obj.self_contributions.items()),
                  description=obj.description, units=dict(obj.units))
        return wrap("KRR", "rascal.models.krr", ip)
    raise ValueError("cannot serialise %r" % (obj,))


def load_obj(fn):
    """Load an object saved with the reference's dump_obj / to_file."""
    fn = os.path.abspath(fn)
    if os.path.splitext(fn)[1] != ".json":
        raise NotImplementedError("Unknown file extention: {}".format(os.path.splitext(fn)[1]))
    with open(fn) as f:
        data = json.load(f)
    if not is_valid_object_dict(data):
        raise RuntimeError("The file: {}; does not contain a valid dictionary".format(fn)
                           + " representation of an object.")
    _load_npy(data, os.path.dirname(fn))
    return from_dict(data)


def dump_obj(fn, obj):
    """Save an object in the reference's JSON format (arrays inline)."""
    if os.path.splitext(fn)[1] != ".json":
        raise NotImplementedError("Unknown file extention: {}".format(os.path.splitext(fn)[1]))
    with open(fn, "w") as f:
        json.dump(to_dict(obj), f)
