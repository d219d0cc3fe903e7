"""Benchmark: SOAP+GAP energy & force throughput on one MI355X.

Default workload (the headline metric): one periodic box of 8192 atoms, 4
species in equal fractions, density 0.06 atoms/A^3, minimum separation 1 A;
SOAP cutoff 5 A (cosine smooth width 0.5 A), Gaussian sigma 0.5 A, GTO radial
basis splined to 1e-8, max_radial 12, max_angular 9, normalised power
spectrum; sparse GAP with 4096 sparse points (1024 per species) taken as
random environments of the same box, zeta 4, weights ~ N(0, 1).

A step is one pass of the hot path over the batch: neighbour list, spherical
expansion, power spectrum, GAP kernel energy, and forces + virial through the
adjoint.  The headline value is the mixed-precision path (FP32 expansion and
invariants, BF16 hi+lo matrix-core GEMMs, FP64 reductions); the FP64 GPU path
is timed on the same input and the force error of the mixed path against it
is reported.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12
PEAK_BF16 = 2.5e15
PEAK_FP64_MATRIX = 78.6e12

WORKLOAD = dict(name="periodic box 8192 atoms, 4 species, rho 0.06 A^-3, rc 5 A, "
                     "n_max 12, l_max 9, 4096 sparse points (1024/species), zeta 4",
                n_atoms=8192, density=0.06, species=(1, 6, 7, 8), rc=5.0, sw=0.5, sigma=0.5,
                N=12, L=9, n_sparse_per_species=1024, zeta=4, seed=20263)


def hypers(w):
    return dict(max_radial=w["N"], max_angular=w["L"], soap_type="PowerSpectrum", normalize=True,
                inversion_symmetry=True, expansion_by_species_method="environment wise",
                global_species=[], compute_gradients=False,
                cutoff_function=dict(type="ShiftedCosine", cutoff=dict(value=w["rc"], unit="AA"),
                                     smooth_width=dict(value=w["sw"], unit="AA")),
                gaussian_density=dict(type="Constant",
                                      gaussian_sigma=dict(value=w["sigma"], unit="AA")),
                radial_contribution=dict(type="GTO", optimization=dict(Spline=dict(accuracy=1e-8))))


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_steps(ctx, positions, steps):
    """One step = NL rebuild + expansion + power spectrum + GAP E/F/virial."""
    out = None
    for _ in range(steps):
        ctx.update_positions(positions)
        out = ctx.gap_predict(energy=True, gradient=True, neg_stress=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-centres", type=int, default=720)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 256 MiB -> 256 MiB elementwise copy, the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box of this build has one MI355X, and the "
              "multi-GPU design is described (not run) in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only (synchronisation)
    from librascal_amd import _capi
    from librascal_amd.synthetic import pack, periodic_box

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 3 x (256 MiB read + 256 MiB write)")

    w = WORKLOAD
    t0 = time.time()
    sts = periodic_box(w["n_atoms"], w["density"], w["species"], d_min=1.0, seed=w["seed"])
    off, pos, num, cells, pbc = pack(sts)
    species = sorted(set(num.tolist()))
    log("structure built in %.1fs" % (time.time() - t0))
    hp = json.dumps(hypers(w))

    ctx64 = _capi.Context(hp)
    ctx64.set_structures(off, pos, num, cells, pbc, species)
    n_pairs = ctx64.info("n_pairs")
    D = ctx64.info("feature_size")
    X, _ = ctx64.features()
    rng = np.random.default_rng(w["seed"])
    sp_species, sp_values = [], []
    for t in species:
        idx = np.where(num == t)[0]
        pick = rng.choice(idx, size=w["n_sparse_per_species"], replace=False)
        sp_species += [t] * len(pick)
        sp_values.append(X[pick])
    sp_species = np.asarray(sp_species, np.int32)
    sp_values = np.concatenate(sp_values)
    weights = rng.normal(size=len(sp_species))
    del X
    ctx64.set_sparse_points(w["zeta"], sp_species, sp_values, weights)
    log("fp64 context ready: %d pairs, D=%d, M=%d" % (n_pairs, D, len(sp_species)))

    ctxm = _capi.Context(hp)
    ctxm.set_precision(_capi.MIXED)
    # neighbour-list pass and representation stage replayed from HIP graphs
    ctxm.enable_graphs(True)
    ctxm.set_structures(off, pos, num, cells, pbc, species)
    ctxm.set_sparse_points(w["zeta"], sp_species, sp_values, weights)
    # work partitions picked once per batch shape (untimed set-up)
    tiles64 = ctx64.autotune()
    tilesm = ctxm.autotune()
    log("autotuned tiles: fp64 %s, mixed %s" % (tiles64, tilesm))

    n_centres = int(off[-1])

    def timed(ctx, steps, warmup, label):
        run_steps(ctx, pos, warmup)
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = run_steps(ctx, pos, steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        # stage times: HIP events on the context's stream around each stage,
        # over a second pass of the same steps (the events, and the eager
        # launches they need, stay out of the pass timed above)
        ctx.enable_timing(True)
        ctx.reset_timing()
        run_steps(ctx, pos, steps)
        torch.cuda.synchronize()
        stages = ctx.timing()
        ctx.enable_timing(False)
        log("%s: %.3f ms/step" % (label, 1e3 * el / steps))
        return el, out, stages

    steps64 = max(2, args.steps // 4)
    el64, (e64, g64, s64), st64 = timed(ctx64, steps64, 1, "fp64")
    elm, (em, gm, sm), stm = timed(ctxm, args.steps, args.warmup, "mixed")

    force_rel_rms = float(np.sqrt(np.mean((gm - g64) ** 2)) / np.sqrt(np.mean(g64 ** 2)))
    energy_rel = float(abs(em[0] - e64[0]) / abs(e64[0]))
    stress_rel = float(np.max(np.abs(sm - s64)) / np.max(np.abs(s64)))

    ms_step = 1e3 * elm / args.steps
    value = n_centres * args.steps / elm
    value64 = n_centres * steps64 / el64

    # roofline of the dominant kernel: the K = (X T^T)^zeta GEMM.  The mixed
    # path multiplies the packed invariants (diagonal species blocks stored
    # once, n1 <= n2): Dp columns instead of D, the same products.
    M_s = w["n_sparse_per_species"]
    rows = [int(np.sum(num == t)) for t in species]
    S_, N_, L1_ = len(species), w["N"], w["L"] + 1
    Dp = S_ * N_ * (N_ + 1) // 2 * L1_ + (S_ * (S_ + 1) // 2 - S_) * N_ * N_ * L1_
    # one grouped launch covers every species block: 2 * rows_s * M_s * Dp flop each
    flop_gemm = sum(2.0 * r * M_s * Dp for r in rows)
    k_ms, k_n = stm["gap_kernel"]
    k_avg_ms = k_ms / max(k_n, 1)
    achieved = flop_gemm / (k_avg_ms * 1e-3)
    peak = PEAK_BF16 / 3.0  # three BF16 products per algorithmic multiply-add
    traffic = profiled_traffic(GEMM_KERNEL)
    roofline = dict(bound="mfma", achieved=achieved / 1e12, peak=peak / 1e12, unit="TFLOP/s",
                    frac=achieved / peak, traffic=traffic,
                    kernel="gap_kernel (%s: K = (X T^T)^zeta, all species blocks in one launch)"
                    % GEMM_KERNEL,
                    avg_launch_ms=k_avg_ms, flop_per_launch=flop_gemm, packed_columns=Dp,
                    bf16_mfma_utilisation=3.0 * achieved / PEAK_BF16,
                    traffic_source=TRAFFIC_FILE if traffic is not None else None)
    # HBM-bound stages: algorithmic bytes per step / stage time (FP32 path)
    n_p = n_pairs
    c_bytes = 4.0 * n_centres * S_ * N_ * L1_ * L1_
    p_bytes = 4.0 * n_centres * Dp  # packed invariants (hi + lo planes) or dL/dX
    pair_bytes = 56.0 * n_p
    stage_bytes = dict(expansion=c_bytes + pair_bytes, power_spectrum=c_bytes + p_bytes,
                       backward=2 * p_bytes + c_bytes + pair_bytes)
    hbm = {}
    for k, b in stage_bytes.items():
        ms = stm[k][0] / max(args.steps, 1)
        hbm[k] = dict(bytes=b, ms=ms, achieved_GBps=b / (ms * 1e-3) / 1e9,
                      frac=b / (ms * 1e-3) / PEAK_HBM)
    roofline["hbm_stages"] = hbm

    stage_ms = {k: round(v[0] / args.steps, 4) for k, v in stm.items()}
    stage_ms64 = {k: round(v[0] / steps64, 4) for k, v in st64.items()}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(sts, hypers(w), sp_species, sp_values, weights, w["zeta"],
                               args.cpu_centres)
        except Exception as exc:  # the baseline is reported, not required
            log("cpu baseline failed: %r" % (exc,))

    line = dict(
        metric="SOAP+GAP energy&force envs/s, 1 MI355X: mixed vs FP64 at ≤1e-4 force err",
        value=value, unit="envs/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms_step, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="fp32 expansion/invariants + bf16x3 MFMA, fp64 reductions", data="synthetic",
        config=dict(workload=w["name"], n_centres=n_centres, n_pairs=n_pairs, feature_size=D,
                    n_sparse=int(len(sp_species)), precision="mixed", parallelism="single GPU",
                    tiles=tilesm, hip_graphs=True),
        fp64_path=dict(value=value64, unit="envs/s", ms_per_step=1e3 * el64 / steps64,
                       stage_ms=stage_ms64, tiles=tiles64),
        errors_vs_fp64=dict(force_rel_rms=force_rel_rms, energy_rel=energy_rel,
                            neg_stress_max_rel=stress_rel),
        stage_ms=stage_ms, roofline=roofline, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


TRAFFIC_FILE = "profiles/hbm_traffic_periodic_box_8192.md"
# the energy GEMM's kernel (the packed layout always takes the direct-to-LDS
# variant)
GEMM_KERNEL = "bf16x3_gemm_dma<0, 128, 2>"


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration) from the committed rocprofv3 --pmc summary, or
    None when it is absent."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].strip("`") == kernel:
            try:
                return float(cells[3]) + float(cells[4])
            except ValueError:
                return None
    return None


def cpu_baseline(structures, hp, sp_species, sp_values, weights, zeta, n_centres):
    """The FP64 CPU oracle (serial restatement of the reference algorithm,
    stored per-pair feature gradients) timed on one host core over the first
    n_centres environments of the same box."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    calc = O.Calculator(hp)
    b = O.BatchData(structures)
    ss = O.SparseSet(sp_species, sp_values, zeta)
    t = time.perf_counter()
    O.gap_predict(calc, b, ss, weights, max_centres=n_centres)
    el = time.perf_counter() - t
    return dict(value=n_centres / el, unit="envs/s", cores=1, kind="port",
                sample="first %d of %d environments of the benchmark box, energy + forces + "
                       "virial, FP64" % (n_centres, b.n_atoms), seconds=el)


if __name__ == "__main__":
    main()
