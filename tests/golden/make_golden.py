"""Write the committed golden fixtures from librascal's own test data.

Reads only DATA files of the reference tree (UBJSON/JSON/XYZ fixtures and
input structures) and stores compact subsets under tests/golden/.  Run it in
a checkout that has the reference at REFERENCE_ROOT (default
/root/reference); the tests read only the fixtures it writes.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ubjson_read  # noqa: E402

REF = os.environ.get("REFERENCE_ROOT", "/root/reference")
DATA = os.path.join(REF, "reference_data")

_SYMBOLS = {"H": 1, "C": 6, "N": 7, "O": 8, "Si": 14, "P": 15, "S": 16,
            "Ca": 20, "Cr": 24}


def ase_json_structures(path, limit=None):
    """ASE-db JSON: cell rows are lattice vectors; stored here transposed so
    that lattice vectors are columns (the reference AtomicStructure layout)."""
    d = json.load(open(path))
    keys = [k for k in d if k not in ("ids", "nextid")]
    keys = sorted(keys, key=int)
    out = []
    for k in keys[:limit]:
        e = d[k]
        out.append(dict(positions=np.asarray(e["positions"], float).tolist(),
                        numbers=[int(x) for x in e["numbers"]],
                        cell=np.asarray(e["cell"], float).T.tolist(),
                        pbc=[int(bool(x)) for x in e["pbc"]]))
    return out


def extxyz_structures(path, limit=None):
    """Extended XYZ frames (Lattice rows = lattice vectors; transposed)."""
    out = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines) and (limit is None or len(out) < limit):
        if not lines[i].strip():
            i += 1
            continue
        n = int(lines[i].split()[0])
        hdr = lines[i + 1]
        lat = hdr.split('Lattice="')[1].split('"')[0].split()
        cell = np.asarray([float(x) for x in lat]).reshape(3, 3)
        pbc = [1, 1, 1]
        if 'pbc="' in hdr:
            pbc = [int(x == "T") for x in hdr.split('pbc="')[1].split('"')[0].split()]
        energy = None
        for tok in hdr.split():
            if tok.startswith("energy="):
                energy = float(tok.split("=")[1])
        pos, num = [], []
        for ln in lines[i + 2:i + 2 + n]:
            f = ln.split()
            num.append(_SYMBOLS[f[0]])
            pos.append([float(f[1]), float(f[2]), float(f[3])])
        out.append(dict(positions=pos, numbers=num, cell=cell.T.tolist(), pbc=pbc,
                        energy=energy))
        i += 2 + n
    return out


def main():
    tests_only = os.path.join(DATA, "tests_only")
    inputs = os.path.join(DATA, "inputs")

    # Hyp1f1 vs mpmath at 200 digits (tests/test_math_hyp1f1.cc:42-88)
    h = ubjson_read.load(os.path.join(tests_only, "hyp1f1_reference.ubjson"))
    arr = np.array([[e["a"], e["b"], e["z"], e["val"], e["der"]] for e in h])
    np.savez_compressed(os.path.join(HERE, "hyp1f1_reference.npz"), data=arr)

    # real spherical harmonics + ALPs (test_math_spherical_harmonics.cc:50-213)
    y = ubjson_read.load(os.path.join(tests_only, "spherical_harmonics_reference.ubjson"))
    np.savez_compressed(
        os.path.join(HERE, "spherical_harmonics_reference.npz"),
        max_angular=np.array([e["max_angular_l"] for e in y]),
        directions=np.array([e["unit_vector"] for e in y]),
        harmonics=np.array([e["harmonics"] for e in y]),
        alps=np.array([e["alps"] for e in y]))

    # spherical expansion features, GTO rows (test_calculator.cc:460-493)
    se = ubjson_read.load(os.path.join(tests_only, "spherical_expansion_reference.ubjson"))
    # GTO rows in one pair of files, DVR rows (direct and Spline) in another,
    # the RadialDimReduction rows (GTO and DVR) in a third
    sel = {"GTO": ([], {}), "DVR": ([], {}), "RDR": ([], {})}
    idx = 0
    for fn in se["filenames"]:
        for rc in se["cutoffs"]:
            for k, e in enumerate(se["rep_info"][idx]):
                hp = e["hypers"]
                opt = hp["radial_contribution"].get("optimization", {})
                basis = hp["radial_contribution"]["type"]
                if "RadialDimReduction" in opt:
                    basis = "RDR"
                if basis not in sel:
                    continue
                rows, feats = sel[basis]
                name = "f%d_rc%s_%d" % (idx // len(se["cutoffs"]), rc, k)
                rows.append(dict(name=name, structure=os.path.basename(fn), hypers=hp))
                feats[name] = np.asarray(e["feature_matrix"])
            idx += 1
    for basis, stem in (("GTO", "spherical_expansion_reference"),
                        ("DVR", "spherical_expansion_dvr_reference"),
                        ("RDR", "spherical_expansion_rdr_reference")):
        rows, feats = sel[basis]
        np.savez_compressed(os.path.join(HERE, stem + ".npz"), **feats)
        if basis == "RDR":
            # every row carries the same projection matrices: kept once
            pms = set()
            for r in rows:
                red = r["hypers"]["radial_contribution"]["optimization"]["RadialDimReduction"]
                pms.add(json.dumps(red.pop("projection_matrices"), sort_keys=True))
            assert len(pms) == 1
            rows = dict(projection_matrices=json.loads(pms.pop()), rows=rows)
        json.dump(rows, open(os.path.join(HERE, stem + ".json"), "w"), indent=1)

    # cosine kernels of the power spectrum (test_kernels.cc:62-101)
    kr = ubjson_read.load(os.path.join(tests_only, "kernel_reference.ubjson"))
    krows, kmats = [], {}
    for k, e in enumerate(kr["rep_info"]["spherical_invariants"][0]):
        if e["hypers_rep"]["soap_type"] != "PowerSpectrum":
            continue
        name = "k%d" % k
        krows.append(dict(name=name, hypers=e["hypers_rep"], kernel=e["hypers_kernel"]))
        kmats[name] = np.asarray(e["kernel_matrix"])
    np.savez_compressed(os.path.join(HERE, "kernel_reference.npz"), **kmats)
    json.dump(dict(start=kr["start"], length=kr["length"], rows=krows),
              open(os.path.join(HERE, "kernel_reference.json"), "w"), indent=1)

    # GAP model fixture (sparse points, weights, hypers) and its structures
    gm = json.load(open(os.path.join(tests_only, "simple_gap_model.json")))
    json.dump(gm, open(os.path.join(HERE, "simple_gap_model.json"), "w"))

    structures = dict(
        CaCrP2O7=ase_json_structures(os.path.join(inputs, "CaCrP2O7_mvc-11955_symmetrized.json")),
        small_molecule=ase_json_structures(os.path.join(inputs, "small_molecule.json")),
        small_molecules_20=ase_json_structures(os.path.join(inputs, "small_molecules-20.json"),
                                               limit=kr["start"] + kr["length"]),
        diamond_cubic_distorted=ase_json_structures(
            os.path.join(inputs, "diamond_cubic_distorted.json")),
        SiC_moissanite_supercell=ase_json_structures(
            os.path.join(inputs, "SiC_moissanite_supercell.json")),
        methane_dimers=extxyz_structures(os.path.join(inputs, "methane_dimer_sample.xyz"),
                                         limit=100),
    )
    json.dump(structures, open(os.path.join(HERE, "structures.json"), "w"))

    # sparse-kernel gradient test inputs (test_sparse_kernels.cc)
    ski = json.load(open(os.path.join(tests_only, "sparse_kernel_inputs.json")))
    json.dump(ski, open(os.path.join(HERE, "sparse_kernel_inputs.json"), "w"), indent=1)
    print("golden fixtures written to", HERE)
    sparsification()


def sparsification():
    """Sparsified power spectrum fixture (test_calculator.cc:622-748): the
    full and coefficient_subselection hypers per case, the global ids of the
    selected coefficients in the reference's per-centre full vector, and the
    input structures they name."""
    tests_only = os.path.join(DATA, "tests_only")
    inputs = os.path.join(DATA, "inputs")
    cases = json.load(open(os.path.join(tests_only, "sparsification_inputs.json")))
    files = sorted(set(os.path.basename(c["filename"]) for c in cases))
    out = dict(cases=[dict(structure=os.path.basename(c["filename"]), rep=c["hypers"]["rep"],
                           rep_sparse=c["hypers"]["rep_sparse"],
                           adaptors=c["hypers"]["adaptors"]) for c in cases],
               structures={f: ase_json_structures(os.path.join(inputs, f)) for f in files})
    json.dump(out, open(os.path.join(HERE, "sparsification_inputs.json"), "w"))
    print("sparsification fixture written")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "sparsification":
        sparsification()
    else:
        main()
