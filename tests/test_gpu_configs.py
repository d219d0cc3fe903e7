"""Parity at the shapes of the benchmark configurations (synthetic seeded
inputs), against the FP64 oracle where it finishes in seconds and against
size-independent properties (finite differences, FP64 vs mixed) elsewhere.

  diamond  8 diamond cells of 64 Si atoms, rc 5, N 8, L 6 (direct Hyp1f1 and
           Spline 1e-8), FP64 expansion + invariants <= 1e-10
  ragged   non-periodic molecules of 8..64 atoms, 4 species, rc 4, N 8, L 6,
           sigma 0.4: FP64 <= 1e-10 (block-relative); FP32 expansion <= 1e-5
           block-relative, FP32 normalised invariants <= 1e-5 absolute (|X| = 1)
  forces   periodic 2-species box (1024 atoms, rho 0.05), rc 5, N 10, L 8,
           zeta 2, 1000 sparse points per species: GAP forces of the GPU
           adjoint vs central differences (h = 1e-4 A) of the GPU energy,
           K_NM vs the oracle
  triclinic  random triclinic cells, 3 species, rc 6, N 12, L 9: FP64
           invariants <= 1e-10 on a subset of cells, mixed GAP forces vs FP64
"""
import json

import numpy as np
import pytest

import oracle as O
from conftest import expansion_rel, ps_rel, rel_err, soap_hypers
from librascal_amd import _capi
from librascal_amd.synthetic import (diamond_cells, molecules, pack, periodic_box, triclinic_cells,
                                     wrap_positions)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def make_ctx(hp, structures, precision=_capi.FP64):
    ctx = _capi.Context(json.dumps(hp))
    ctx.set_precision(precision)
    off, pos, num, cells, pbc = pack(structures)
    sp = sorted(set(num.tolist()))
    ctx.set_structures(off, pos, num, cells, pbc, sp)
    return ctx, sp


@pytest.mark.parametrize("spline", [None, 1e-8])
def test_diamond_cells_fp64(spline):
    sts = diamond_cells(n_cells=8)
    hp = soap_hypers(rc=5.0, sw=0.5, sigma=0.5, N=8, L=6, spline=spline)
    ctx, sp = make_ctx(hp, sts)
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=sp)
    c_ref, _ = calc.expansion(b)
    c, _ = ctx.expansion()
    assert expansion_rel(c_ref, c, 6) < FP64_TOL
    p_ref, _ = calc.power_spectrum(b)
    X, _ = ctx.features()
    assert ps_rel(p_ref, X, len(sp), 8, 6) < FP64_TOL


def test_ragged_molecules_fp32_and_fp64():
    sts = molecules(n_mol=48)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=8, L=6, spline=1e-8)
    calc = O.Calculator(hp)
    ctx, sp = make_ctx(hp, sts)
    b = O.BatchData(sts, species=sp)
    c_ref, _ = calc.expansion(b)
    p_ref, _ = calc.power_spectrum(b)
    c64, _ = ctx.expansion()
    X64, _ = ctx.features()
    assert expansion_rel(c_ref, c64, 6) < FP64_TOL
    assert ps_rel(p_ref, X64, len(sp), 8, 6) < FP64_TOL
    ctx.set_precision(_capi.MIXED)
    c32, _ = ctx.expansion()
    X32, _ = ctx.features()
    assert expansion_rel(c_ref, c32, 6) < 1e-5
    assert np.abs(X32 - p_ref).max() < 1e-5


def _gap_box(seed=20262):
    sts = periodic_box(1024, 0.05, (1, 8), d_min=1.0, seed=seed)
    hp = soap_hypers(rc=5.0, sw=0.5, sigma=0.5, N=10, L=8, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    X, _ = ctx.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(seed)
    spec, vals = [], []
    for t in sp:
        idx = np.where(num == t)[0]
        pick = rng.choice(idx, size=1000, replace=True)  # 512 atoms per species
        spec += [t] * 1000
        vals.append(X[pick])
    spec = np.asarray(spec, np.int32)
    vals = np.concatenate(vals)
    w = rng.normal(size=len(spec))
    ctx.set_sparse_points(2, spec, vals, w)
    return sts, hp, ctx, sp, spec, vals, w


def test_forces_via_adjoint_vs_finite_differences():
    sts, hp, ctx, sp, spec, vals, w = _gap_box()
    e0, g, s = ctx.gap_predict()
    pos = pack(sts)[1]
    h = 1e-4
    rng = np.random.default_rng(1)
    checked = 0
    for atom in rng.choice(len(pos), 6, replace=False):
        for c in range(3):
            es = []
            for sgn in (1, -1):
                p = pos.copy()
                p[atom, c] += sgn * h
                ctx.update_positions(p)
                es.append(ctx.gap_predict(True, False, False)[0][0])
            fd = (es[0] - es[1]) / (2 * h)
            assert abs(fd - g[atom, c]) < 1e-5 * np.abs(g).max() + 1e-7 * abs(fd)
            checked += 1
    ctx.update_positions(pos)
    assert checked == 18


def test_forces_box_knm_vs_oracle():
    sts, hp, ctx, sp, spec, vals, w = _gap_box()
    calc = O.Calculator(hp)
    b = O.BatchData(sts, species=sp)
    ss = O.SparseSet(spec, vals, 2)
    k_ref = np.zeros((1, len(spec)))
    k_ref[:, ss.order] = O.gap_knm(calc, b, ss, "Structure")
    k = ctx.gap_knm("Structure")
    assert rel_err(k_ref, k).max() < FP64_TOL


def test_triclinic_cells_fp64_subset_and_mixed_forces():
    sts = triclinic_cells(n_cells=6, n_min=32, n_max=96)
    hp = soap_hypers(rc=6.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    calc = O.Calculator(hp)
    sub = sts[:2]
    b = O.BatchData(sub, species=sp)
    p_ref, _ = calc.power_spectrum(b)
    X, _ = ctx.features()
    n_sub = b.n_atoms
    assert ps_rel(p_ref, X[:n_sub], len(sp), 12, 9) < FP64_TOL
    rng = np.random.default_rng(4)
    num = pack(sts)[2]
    spec = np.concatenate([np.full(64, t, np.int32) for t in sp])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 64)] for t in sp])
    w = rng.normal(size=len(spec))
    ctx.set_sparse_points(2, spec, vals, w)
    e64, g64, s64 = ctx.gap_predict()
    ctx.set_precision(_capi.MIXED)
    e32, g32, s32 = ctx.gap_predict()
    assert rel_err(e64, e32).max() < 1e-4
    rms = np.sqrt(np.mean((g32 - g64) ** 2)) / np.sqrt(np.mean(g64 ** 2))
    assert rms < 1e-4


def test_ragged_batch_graph_replays_match_eager():
    """The representation stage replayed from a captured HIP graph, across
    position updates (new neighbour lists), equals eager launches; each replay
    is checked, with eager work and synchronisation between replays."""
    sts = molecules(n_mol=64, seed=77)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=8, L=6, spline=1e-8)
    g_ctx, sp = make_ctx(hp, sts, precision=_capi.MIXED)
    g_ctx.enable_graphs(True)
    e_ctx, _ = make_ctx(hp, sts, precision=_capi.MIXED)
    pos0 = pack(sts)[1]
    rng = np.random.default_rng(9)
    for it in range(4):
        pos = pos0 + (0.0 if it == 0 else 0.02) * rng.normal(size=pos0.shape)
        # keep the jittered molecules inside their (non-periodic) boxes
        off, _, _, cells, _ = pack(sts)
        for s in range(len(sts)):
            sl = slice(off[s], off[s + 1])
            pos[sl] = np.clip(pos[sl], 1e-6, np.diag(cells[s]) - 1e-6)
        g_ctx.update_positions(pos)
        e_ctx.update_positions(pos)
        g_ctx.compute()
        Xg, ng = g_ctx.features()
        Xe, ne = e_ctx.features()  # eager work between replays
        assert np.array_equal(Xg, Xe) and np.array_equal(ng, ne)
    assert g_ctx.info("graph_captures") >= 1
    assert g_ctx.info("graph_launches") == 4


def test_autotuned_tiles_keep_results():
    """The autotuner picks work partitions on the triclinic 3-species batch
    (rc 6, N 12, L 9); features and the adjoint's forces agree to rounding
    with the default partitions (the norm sums change order with the key
    blocks per pass)."""
    sts = triclinic_cells(n_cells=8, n_min=32, n_max=128, seed=31)
    hp = soap_hypers(rc=6.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8)
    for precision in (_capi.FP64, _capi.MIXED):
        ctx, sp = make_ctx(hp, sts, precision=precision)
        X0, _ = ctx.features()
        rng = np.random.default_rng(2)
        G = rng.normal(size=X0.shape)
        g0, s0 = ctx.features_backward(G)
        tiles = ctx.autotune()
        # pairs_adj 0: as many pairs per batch as fit (or, adj_kind 3, the
        # pair phase in a kernel of its own)
        assert tiles["pairs_fwd"] > 0 and tiles["keys_ps"] > 0 and tiles["pairs_adj"] >= 0
        assert tiles["adj_kind"] in (0, 1, 2, 3)
        X1, _ = ctx.features()
        tol = 1e-12 if precision == _capi.FP64 else 1e-5
        assert np.abs(X1 - X0).max() < tol * np.abs(X0).max()
        g1, s1 = ctx.features_backward(G)
        assert np.abs(g1 - g0).max() < tol * np.abs(g0).max()
        assert np.abs(s1 - s0).max() < tol * np.abs(s0).max()


@pytest.mark.parametrize("N,L", [(12, 9), (8, 6), (5, 4), (6, 3)])
def test_mixed_adjoint_partitions_agree(N, L):
    """Mixed-precision adjoint (packed dL/dX and invariants, dL/dp staged in
    LDS one, three or all key blocks per pass, several pair batch sizes) vs
    the FP64 adjoint (pinned against the oracle elsewhere), on ragged
    molecules with 4 species and a random dL/dX that is not symmetric in the
    diagonal-key blocks.  N = 5, L = 4 gives key blocks of 125 and 75
    elements (element-wise staging) and a partial group of radial channels;
    N = 6 another partial group.  adj_kind 1 is the matrix-core kernel
    (no invariants read; dL/dc = alpha f_l W_l C_l - 2 beta f_l^2 C_l C_l^T C_l),
    adj_kind 2 the register-tile kernel with the same formula (built for
    N, L = 12, 9 and 8, 6; the staged kernel elsewhere), keys_adj its key
    blocks per staged chunk of rounds; adj_kind 3 the same with the pair
    phase as a kernel of its own."""
    sts = molecules(n_mol=24, seed=5)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=N, L=L, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    X64, _ = ctx.features()
    rng = np.random.default_rng(N)
    G = rng.normal(size=X64.shape)
    g64, s64 = ctx.features_backward(G)
    ctx.set_precision(_capi.MIXED)
    out = {}
    for kind in (0, 1, 2, 3):
        for kch, pb in ((0, 0), (1, 4), (3, 16), (10, 6)):
            ctx.set_tiles(pairs_adj=pb, keys_adj=kch, adj_kind=kind)
            t = ctx.tiles()
            assert t["keys_adj"] == kch and t["pairs_adj"] == pb and t["adj_kind"] == kind
            ctx.features()
            out[kind, kch, pb] = ctx.features_backward(G)
    scale = np.abs(g64).max()
    for key, (g, s) in out.items():
        assert np.abs(g - g64).max() < 2e-5 * scale, key
        assert np.abs(s - s64).max() < 2e-5 * np.abs(s64).max(), key
        assert np.abs(g - out[0, 0, 0][0]).max() < 1e-5 * scale, key


@pytest.mark.parametrize("n_species", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("N,L", [(12, 9), (8, 6)])
def test_tile_adjoint_species_counts(n_species, N, L):
    """The register-tile adjoint (adj_kind 2; 3 with the pair phase as a
    kernel of its own) over 1-8 species: the circle-
    method rounds with a bye (odd counts), one or two species per wavefront
    (more than 4 species), chunks of one round (keys_adj 1) and of all
    rounds; mixed forces and stress vs the FP64 adjoint on a random dL/dX,
    with and without normalisation (beta = 0)."""
    species = (1, 6, 7, 8, 9, 14, 15, 16)[:n_species]
    probs = np.full(n_species, 1.0 / n_species)
    sts = molecules(n_mol=16, n_min=12, n_max=40, species=species, probs=probs, seed=40 + n_species)
    for normalize in (True, False):
        hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=N, L=L, spline=1e-8)
        hp["normalize"] = normalize
        ctx, sp = make_ctx(hp, sts)
        X64, _ = ctx.features()
        rng = np.random.default_rng(n_species)
        G = rng.normal(size=X64.shape)
        g64, s64 = ctx.features_backward(G)
        ctx.set_precision(_capi.MIXED)
        scale = np.abs(g64).max()
        for kind in (2, 3):
            for kch in (0, 1, 2 * n_species):
                ctx.set_tiles(pairs_adj=0, keys_adj=kch, adj_kind=kind)
                ctx.features()
                g, s = ctx.features_backward(G)
                assert np.abs(g - g64).max() < 2e-5 * scale, (normalize, kind, kch)
                assert np.abs(s - s64).max() < 2e-5 * np.abs(s64).max(), (normalize, kind, kch)


def test_large_gemm_tiles_mixed_vs_fp64():
    """A larger mixed batch (2 species x ~2048 centres, 2048 sparse points
    per species, N 12, L 9; a few hundred GEMM tiles per product): mixed
    energies, forces and stress vs the FP64 GPU path."""
    sts = periodic_box(4096, 0.05, (1, 8), d_min=1.0, seed=99)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    X, _ = ctx.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(3)
    spec = np.concatenate([np.full(2048, t, np.int32) for t in sp])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 2048)] for t in sp])
    w = rng.normal(size=len(spec))
    ctx.set_sparse_points(3, spec, vals, w)
    e64, g64, s64 = ctx.gap_predict()
    ctx.set_precision(_capi.MIXED)
    e32, g32, s32 = ctx.gap_predict()
    assert rel_err(e64, e32).max() < 1e-4
    assert np.sqrt(np.mean((g32 - g64) ** 2)) / np.sqrt(np.mean(g64 ** 2)) < 1e-4
    assert np.abs(s32 - s64).max() < 1e-4 * np.abs(s64).max()


@pytest.mark.parametrize("normalize", [True, False])
def test_mixed_prediction_adjoint_repeatable(normalize):
    """The prediction hands the adjoint X_i . G_i = zeta E_i (no in-kernel
    reduction, so none of its barriers): repeated mixed predictions agree to
    rounding, and their forces equal features_backward of the same dE/dX
    (zeta 1: dE/dX_i = sum over the species' sparse points of w_m T_m), where
    the adjoint reduces X . G itself.  Guards the ordering of the adjoint's
    LDS tables against their first use."""
    sts = periodic_box(4096, 0.05, (1, 8), d_min=1.0, seed=99)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.5, N=12, L=9, spline=1e-8, normalize=normalize)
    ctx, sp = make_ctx(hp, sts)
    X, _ = ctx.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(5)
    spec = np.concatenate([np.full(2048, t, np.int32) for t in sp])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 2048)] for t in sp])
    w = rng.normal(size=len(spec))
    ctx.set_sparse_points(1, spec, vals, w)
    G = np.zeros_like(X)
    for t in sp:
        G[num == t] = (w[spec == t][:, None] * vals[spec == t]).sum(0)
    ctx.set_precision(_capi.MIXED)
    ref = ctx.features_backward(G)[0]
    scale = np.abs(ref).max()
    for _ in range(4):
        g = ctx.gap_predict()[1]
        assert np.abs(g - ref).max() < 1e-5 * scale


@pytest.mark.parametrize("graphs", [False, True])
def test_deferred_neighbour_check_rebuilds_a_short_list(graphs):
    """update_positions leaves the size check of the neighbour list to the
    prediction, which checks after its own work.  Compressing a dilute box
    (about 10 neighbours per atom within rc 5) to about 63 per atom exceeds
    the pair capacity kept from the first build (48 per atom + 1024): the
    prediction has to build the list again and repeat itself, and then
    matches a context built on the dense box directly (FP64).  Two updates
    without a prediction between them are checked the same way.  With
    graphs, the list passes of the updates are replayed from a captured
    graph; every replay's list equals the eagerly built one."""
    n = 512
    dense = periodic_box(n, 0.12, (1, 8), d_min=1.0, seed=7)
    hp = soap_hypers(rc=5.0, sw=0.5, sigma=0.5, N=4, L=3)
    off, pos, num, cells, pbc = pack(dense)
    sp = sorted(set(num.tolist()))
    s = (0.12 / 0.02) ** (1.0 / 3.0)
    ctx = _capi.Context(json.dumps(hp))
    if graphs:
        ctx.enable_graphs(True)
    ctx.set_structures(off, pos * s, num, cells * s, pbc, sp)
    n_dilute = ctx.info("n_pairs")
    ref = _capi.Context(json.dumps(hp))
    ref.set_structures(off, pos, num, cells, pbc, sp)
    n_dense = ref.info("n_pairs")
    assert n_dilute < 48 * n + 1024 < n_dense
    nl_dense = ref.neighbours()
    dil = _capi.Context(json.dumps(hp))
    dil.set_structures(off, pos * s, num, cells * s, pbc, sp)
    nl_dilute = dil.neighbours()
    X, _ = ref.features()
    rng = np.random.default_rng(1)
    idx = rng.choice(n, 96, replace=False)
    spec = num[idx].astype(np.int32)
    w = rng.normal(size=len(idx))
    for c in (ctx, ref):
        c.set_sparse_points(2, spec, X[idx], w)
    e0, g0, s0 = ref.gap_predict()

    def same(out):
        e, g, st = out
        np.testing.assert_allclose(e, e0, rtol=1e-12)
        np.testing.assert_allclose(g, g0, rtol=0, atol=1e-12 * np.abs(g0).max())
        np.testing.assert_allclose(st, s0, rtol=0, atol=1e-12 * np.abs(s0).max())

    def same_list(a, b):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)

    ctx.update_positions(pos, cells)
    same(ctx.gap_predict())
    assert ctx.info("n_pairs") == n_dense
    same_list(ctx.neighbours(), nl_dense)
    ctx.update_positions(pos * s, cells * s)
    ctx.update_positions(pos, cells)
    same(ctx.gap_predict())
    same_list(ctx.neighbours(), nl_dense)
    # the dilute box again, then the dense one twice: every list fits
    for p, c, nl in ((pos * s, cells * s, nl_dilute), (pos, cells, nl_dense), (pos, cells, nl_dense)):
        ctx.update_positions(p, c)
        out = ctx.gap_predict()
        same_list(ctx.neighbours(), nl)
    same(out)
    if graphs:
        assert ctx.info("nl_graph_launches") >= 4
    else:
        assert ctx.info("nl_graph_launches") == 0


@pytest.mark.parametrize("precision", [_capi.FP64, _capi.MIXED])
def test_reductions_are_repeatable(precision):
    """Every sum of the prediction runs in a fixed order (the energy GEMM's
    per-tile partials added in tile order, per-pair forces gathered per
    atom in pair order, virials and structure sums per segment), so two
    predictions, and one after the neighbour list and its reverse lists are
    rebuilt, agree bit for bit; the K_NM Structure rows equal the Atom rows
    summed in centre order."""
    sts = triclinic_cells(n_cells=4, n_min=40, n_max=90, seed=41)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.5, N=8, L=6, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    ctx.set_precision(precision)
    X, _ = ctx.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(42)
    # 300 points per species: several column tiles of both energy GEMMs
    spec = np.concatenate([np.full(300, t, np.int32) for t in sp])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 300)] for t in sp])
    ctx.set_sparse_points(2, spec, vals, rng.normal(size=len(spec)))
    first = ctx.gap_predict()
    again = ctx.gap_predict()
    ctx.update_positions(pack(sts)[1])
    rebuilt = ctx.gap_predict()
    for a, b, c in zip(first, again, rebuilt):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    k_struct, k_atom = ctx.gap_knm("Structure"), ctx.gap_knm("Atom")
    assert np.array_equal(ctx.gap_knm("Structure"), k_struct)
    off = pack(sts)[0]
    for s in range(len(sts)):
        t = np.zeros(k_atom.shape[1])
        for a in range(off[s], off[s + 1]):
            t += k_atom[a]
        assert np.array_equal(k_struct[s], t)


def test_caller_stream_with_graphs_matches_default_stream():
    """A context on a caller's stream (a torch stream), with graph replay,
    over two position updates: the reverse pair lists built on the context's
    second stream are joined into the caller's stream, and every prediction
    equals that of a context on the default stream, bit for bit."""
    import torch
    sts = triclinic_cells(n_cells=3, n_min=40, n_max=90, seed=45)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.5, N=8, L=6, spline=1e-8)
    a, sp = make_ctx(hp, sts, precision=_capi.MIXED)
    b, _ = make_ctx(hp, sts, precision=_capi.MIXED)
    stream = torch.cuda.Stream()
    a.set_stream(stream.cuda_stream)
    a.enable_graphs(True)
    X, _ = b.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(46)
    spec = np.concatenate([np.full(40, t, np.int32) for t in sp])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 40)] for t in sp])
    w = rng.normal(size=len(spec))
    a.set_sparse_points(2, spec, vals, w)
    b.set_sparse_points(2, spec, vals, w)
    off, pos0, _, cells, pbc = pack(sts)
    for it in range(3):
        pos = pos0 + (0.0 if it == 0 else 0.01) * rng.normal(size=pos0.shape)
        for k in range(len(sts)):  # moved atoms wrapped back into their cells
            seg = slice(off[k], off[k + 1])
            pos[seg] = wrap_positions(pos[seg], cells[k], pbc[k])
        a.update_positions(pos)
        b.update_positions(pos)
        for x, y in zip(a.gap_predict(), b.gap_predict()):
            assert np.array_equal(x, y)
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", [_capi.FP64, _capi.MIXED])
def test_species_without_sparse_points(precision):
    """A species of the batch with no sparse points adds nothing to the
    energy (its kernel rows are empty) but its atoms still move the other
    centres' environments: energies, forces and stress against the oracle on
    a 3-species box whose sparse points cover two species, 200 per species
    (several column tiles), and bit-for-bit repeatable."""
    sts = periodic_box(300, 0.06, (1, 6, 8), d_min=1.0, seed=43)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.5, N=6, L=4, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    X, _ = ctx.features()
    num = pack(sts)[2]
    rng = np.random.default_rng(44)
    spec = np.concatenate([np.full(200, t, np.int32) for t in (1, 8)])
    vals = np.concatenate([X[rng.choice(np.where(num == t)[0], 200)] for t in (1, 8)])
    w = rng.normal(size=len(spec))
    ctx.set_sparse_points(2, spec, vals, w)
    ctx.set_precision(precision)
    e, g, s = ctx.gap_predict()
    e_ref, g_ref, s_ref = O.gap_predict(O.Calculator(hp), O.BatchData(sts, species=sp),
                                        O.SparseSet(spec, vals, 2), w)
    tol = FP64_TOL if precision == _capi.FP64 else 1e-4
    assert rel_err(e_ref, e).max() < tol
    assert np.abs(g - g_ref).max() < tol * np.abs(g_ref).max()
    if precision == _capi.FP64:
        assert np.abs(s - s_ref).max() < tol * np.abs(s_ref).max()
    for a, b in zip((e, g, s), ctx.gap_predict()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [_capi.FP64, _capi.MIXED])
def test_forces_with_long_reverse_lists(precision):
    """Atoms named by more than 256 pairs (dense jittered diamond, a 3.57,
    rc 7.5: about 300 neighbours) have their pair forces summed on the
    long-list path of the force gather: forces (and the FP64 stress)
    against the oracle, and bit-for-bit repeatable."""
    sts = diamond_cells(n_cells=1, reps=2, a=3.57, jitter=0.05, species=6, seed=3)
    hp = soap_hypers(rc=7.5, sw=0.5, sigma=0.5, N=6, L=4, spline=1e-8)
    ctx, sp = make_ctx(hp, sts)
    off, pj, _ = ctx.neighbours(with_vectors=False)
    strict = np.concatenate([pj[off[i] + 1:off[i + 1]] for i in range(len(off) - 1)])
    assert np.bincount(strict).min() > 256
    X, _ = ctx.features()
    rng = np.random.default_rng(5)
    spec, vals, w = np.full(40, 6, np.int32), X[rng.choice(len(X), 40)], rng.normal(size=40)
    ctx.set_sparse_points(2, spec, vals, w)
    ctx.set_precision(precision)
    e, g, s = ctx.gap_predict()
    calc = O.Calculator(hp)
    e_ref, g_ref, s_ref = O.gap_predict(calc, O.BatchData(sts, species=sp), O.SparseSet(spec, vals, 2), w)
    # mixed: FP32 pair sums over ~300 neighbours with strong cancellation
    # (1.3e-4 of the largest component measured)
    tol = FP64_TOL if precision == _capi.FP64 else 1e-3
    assert np.abs(g - g_ref).max() < tol * np.abs(g_ref).max()
    if precision == _capi.FP64:  # the stress takes no part in the gather
        assert np.abs(s - s_ref).max() < tol * np.abs(s_ref).max()
    again = ctx.gap_predict()
    for a, b in zip((e, g, s), again):
        assert np.array_equal(a, b)


def test_prediction_outputs_to_device_memory():
    """Outputs given as device pointers (contiguous torch tensors on the GPU)
    receive exactly the values written to host arrays (the prediction's
    sums run in a fixed order)."""
    import torch
    sts, hp, ctx, sp, spec, vals, w = _gap_box()
    e, g, s = ctx.gap_predict()

    def close(a, b):
        return np.array_equal(a, b)

    ed = torch.zeros(e.shape, dtype=torch.float64, device="cuda")
    gd = torch.zeros(g.shape, dtype=torch.float64, device="cuda")
    sd = torch.zeros(s.shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _capi.check(_capi.lib().rgpu_gap_predict(ctx._p, _capi.ptr(ed), _capi.ptr(gd), _capi.ptr(sd)))
    torch.cuda.synchronize()
    assert close(ed.cpu().numpy(), e) and close(gd.cpu().numpy(), g) and close(sd.cpu().numpy(), s)
    # host energy, device gradient and stress in one call
    eh = np.zeros_like(e)
    gd.zero_()
    sd.zero_()
    torch.cuda.synchronize()
    _capi.check(_capi.lib().rgpu_gap_predict(ctx._p, _capi.ptr(eh), _capi.ptr(gd), _capi.ptr(sd)))
    torch.cuda.synchronize()
    assert close(eh, e) and close(gd.cpu().numpy(), g) and close(sd.cpu().numpy(), s)


@pytest.mark.parametrize("spline", [None, 1e-8])
@pytest.mark.parametrize("N,L", [(12, 9), (5, 4), (8, 6)])
def test_fused_expansion_invariants_equal_two_launches(N, L, spline):
    """The persistent fused expansion + power spectrum launch (c kept in LDS
    between the stages), the expansion writing c once from LDS (fused 2) and
    the FP32 expansion on pair groups (fused 3; per-entry in FP64 and for
    N, L it is not built for), against the separate kernels: expansion, key masks
    and invariants equal in FP64 (to rounding) and in the mixed path, on
    ragged molecules with 4 species (more centres than resident
    workgroups for the larger shapes)."""
    sts = molecules(n_mol=64, seed=9)
    hp = soap_hypers(rc=4.0, sw=0.5, sigma=0.4, N=N, L=L, spline=spline)
    for prec, tol in ((_capi.FP64, 1e-13), (_capi.MIXED, 1e-6)):
        ctx, sp = make_ctx(hp, sts)
        ctx.set_precision(prec)
        out = {}
        for fused in (0, 1, 2, 3):
            ctx.set_tiles(fused=fused)
            assert ctx.tiles()["fused"] == fused
            ctx.compute()
            c, km = ctx.expansion()
            X, inv = ctx.features()
            out[fused] = (c, km, X, inv)
        c0, km0, X0, inv0 = out[0]
        for variant in (1, 2, 3):
            c1, km1, X1, inv1 = out[variant]
            assert np.array_equal(km0, km1)
            assert np.abs(c1 - c0).max() <= tol * np.abs(c0).max()
            assert np.abs(X1 - X0).max() <= tol * np.abs(X0).max()
            assert np.abs(inv1 - inv0).max() <= tol * np.abs(inv0).max()
