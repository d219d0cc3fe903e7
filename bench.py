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
        print("bench.py measures one GPU only (--gpus 1): the library runs the whole batch on one "
              "MI355X; sharding the centres over several GPUs is not implemented",
              file=sys.stderr)
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

    # rooflines of the step's stages from their HIP-event times; the stage
    # with the longest time per step leads, the others are "secondary"
    S_, N_, L1_ = len(species), w["N"], w["L"] + 1
    Dp = S_ * N_ * (N_ + 1) // 2 * L1_ + (S_ * (S_ + 1) // 2 - S_) * N_ * N_ * L1_
    Dq = (Dp + 31) // 32 * 32
    n_p = n_pairs
    c_bytes = 4.0 * n_centres * S_ * N_ * L1_ * L1_
    g_bytes = 4.0 * n_centres * Dq                     # dL/dX, FP32 packed rows
    p_bytes = 2.0 * 2.0 * n_centres * Dq               # invariants, BF16 hi + lo planes
    pair_bytes = 56.0 * n_p                            # vector, distance, fc, dfc, j, species
    out_bytes = 24.0 * n_p + 48.0 * n_centres          # force updates, centre virials
    adj_kind = tilesm.get("adj_kind", 0)
    adj_kernels = ADJOINT_KERNELS[adj_kind]
    # the staged kernel (kind 0) reads the invariants; kind 3 passes dL/dc
    # (the size of c) through HBM between its two kernels
    p_read = p_bytes if adj_kind == 0 else 0.0
    dc_trip = 2.0 * c_bytes if adj_kind == 3 else 0.0
    adj_bytes = g_bytes + p_read + c_bytes + dc_trip + pair_bytes + out_bytes

    def per_launch(stage):
        ms, nl = stm[stage]
        return ms / max(nl, 1)

    def valu_issue(kernels, avg_ms):
        vals = [profiled_counter(SQ_FILE, k, "SQ_INSTS_VALU") for k in kernels]
        if any(v is None for v in vals):
            return None
        # a wave64 VALU instruction occupies its SIMD 4 cycles; 1024 SIMDs
        busy_s = sum(vals) * 4.0 / (1024 * CLOCK_HZ)
        return dict(insts_valu_per_launch=sum(vals), issue_ms=busy_s * 1e3,
                    frac_of_launch=busy_s / (avg_ms * 1e-3), clock_hz=CLOCK_HZ, source=SQ_FILE)

    def traffic_of(kernels):
        t = [profiled_traffic(k) for k in kernels]
        return None if any(v is None for v in t) else sum(t)

    lines = {}
    a_avg = per_launch("backward")
    lines["backward"] = dict(
        bound="hbm", achieved=adj_bytes / (a_avg * 1e-3) / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s",
        frac=adj_bytes / (a_avg * 1e-3) / PEAK_HBM, traffic=traffic_of(adj_kernels),
        kernel="adjoint (%s: dL/dX -> dL/dc -> pair forces and virials)" % " + ".join(adj_kernels),
        avg_launch_ms=a_avg, bytes_per_launch=adj_bytes,
        bytes_per_centre=dict(dL_dX=g_bytes / n_centres, invariant_planes=p_read / n_centres,
                              expansion=c_bytes / n_centres, dL_dc_round_trip=dc_trip / n_centres,
                              pair_records=pair_bytes / n_centres, outputs=out_bytes / n_centres),
        valu_issue=valu_issue(adj_kernels, a_avg), traffic_source=TRAFFIC_FILE)
    M_s = w["n_sparse_per_species"]
    rows = [int(np.sum(num == t)) for t in species]
    flop_gemm = sum(2.0 * r * M_s * Dp for r in rows)  # packed columns, all species blocks
    peak_mfma = PEAK_BF16 / 3.0  # three BF16 products per algorithmic multiply-add
    for stage, kern in (("gap_kernel", GEMM_KERNEL), ("gap_force", FORCE_GEMM_KERNEL)):
        avg = per_launch(stage)
        ach = flop_gemm / (avg * 1e-3)
        lines[stage] = dict(bound="mfma", kernel=kern, achieved=ach / 1e12, peak=peak_mfma / 1e12,
                            unit="TFLOP/s", frac=ach / peak_mfma, avg_launch_ms=avg,
                            flop_per_launch=flop_gemm, traffic=profiled_traffic(kern),
                            traffic_source=TRAFFIC_FILE)
    stage_bytes = dict(expansion=(c_bytes + pair_bytes, EXPANSION_KERNEL),
                       power_spectrum=(c_bytes + p_bytes, PS_KERNEL))
    for stage, (b, kern) in stage_bytes.items():
        avg = per_launch(stage)
        lines[stage] = dict(bound="hbm", kernel=kern, bytes_per_launch=b, avg_launch_ms=avg,
                            achieved=b / (avg * 1e-3) / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s",
                            frac=b / (avg * 1e-3) / PEAK_HBM, traffic=profiled_traffic(kern),
                            traffic_source=TRAFFIC_FILE)
    lead = max(lines, key=lambda k: lines[k]["avg_launch_ms"])
    roofline = dict(lines[lead], stage=lead)
    roofline["secondary"] = {k: v for k, v in lines.items() if k != lead}

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
SQ_FILE = "profiles/sq_counters_periodic_box_8192.md"
# kernels of the headline step (the packed layout always takes the
# direct-to-LDS GEMMs; the adjoint kernel is the autotuner's choice)
GEMM_KERNEL = "bf16x3_gemm_dma<0, 128, 2>"
FORCE_GEMM_KERNEL = "bf16x3_gemm_dma<1, 128, 2>"
EXPANSION_KERNEL = "expansion_kernel<float, 5, false, 12, 9>"
PS_KERNEL = "power_spectrum_kernel<float, true, 12, 9>"
ADJOINT_KERNELS = {0: ["adjoint_staged_kernel<false, 5>"], 1: ["adjoint_mfma_kernel<false, 12>"],
                   2: ["adjoint_tile_kernel<12, 9, 6, 1, 10, false, true>"],
                   3: ["adjoint_tile_kernel<12, 9, 6, 1, 6, false, false>", "adjoint_pair_kernel<12, 9>"]}
# shader clock under this load (SQ_BUSY_CYCLES over the kernel durations)
CLOCK_HZ = 2.0e9


def _profile_rows(path):
    path = os.path.join(ROOT, path)
    if not os.path.exists(path):
        return None, []
    header, rows = None, []
    for line in open(path):
        if not line.startswith("|"):
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if header is None:
            header = cells
        elif not set("".join(cells)) <= set("-: "):
            rows.append(cells)
    return header, rows


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration) from the committed rocprofv3 --pmc summary, or
    None when it is absent."""
    header, rows = _profile_rows(TRAFFIC_FILE)
    for cells in rows:
        if len(cells) >= 5 and cells[0].strip("`") == kernel:
            try:
                return float(cells[3]) + float(cells[4])
            except ValueError:
                return None
    return None


def profiled_counter(path, kernel, counter):
    """Per-launch value of an SQ counter of `kernel` (summed over its
    wavefronts) from a committed rocprofv3 --pmc summary, or None."""
    header, rows = _profile_rows(path)
    if not header or counter not in header:
        return None
    k = header.index(counter)
    for cells in rows:
        if cells[0].strip("`") == kernel and len(cells) > k:
            try:
                return float(cells[k].split()[0])
            except ValueError:
                return None
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# worker state of the all-cores baseline (one oracle calculator per process)
_W = {}


def _worker_init(structures, hp, sp_species, sp_path, weights, zeta):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    _W["O"] = O
    _W["calc"] = O.Calculator(hp)
    _W["batch"] = O.BatchData(structures)
    _W["sparse"] = O.SparseSet(sp_species, np.load(sp_path, mmap_mode="r"), zeta)
    _W["w"] = weights


def _worker_run(rng):
    first, count = rng
    O = _W["O"]
    t = time.perf_counter()
    e, g, s = O.gap_predict(_W["calc"], _W["batch"], _W["sparse"], _W["w"], first_centre=first,
                            max_centres=count)
    return e, g, s, time.perf_counter() - t


def cpu_baseline(structures, hp, sp_species, sp_values, weights, zeta, n_centres, per_worker=240):
    """The FP64 CPU oracle (serial restatement of the reference algorithm,
    stored per-pair feature gradients), energy + forces + virial of
    environments of the same box: on one host core over the first
    n_centres environments, and on all the cores given to this process
    (one oracle process per core, consecutive ranges of per_worker
    environments each)."""
    import multiprocessing as mp
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    calc = O.Calculator(hp)
    b = O.BatchData(structures)
    ss = O.SparseSet(sp_species, sp_values, zeta)
    t = time.perf_counter()
    O.gap_predict(calc, b, ss, weights, max_centres=n_centres)
    el1 = time.perf_counter() - t
    single = dict(value=n_centres / el1, cores=1, centres=n_centres, seconds=el1)
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = nproc
    # this process's share of the host: the cores it may run on, at most 16
    # (the CPU share of one GPU on the benchmark boxes)
    workers = max(1, min(allowed, 16))
    multi = None
    tmp = tempfile.mkdtemp(prefix="rascal_cpu_")
    try:
        sp_path = os.path.join(tmp, "sparse.npy")
        np.save(sp_path, np.ascontiguousarray(sp_values, np.float64))
        ranges = [(k * per_worker, per_worker) for k in range(workers)]
        ctx = mp.get_context("spawn")
        with ctx.Pool(workers, initializer=_worker_init,
                      initargs=(structures, hp, np.asarray(sp_species), sp_path,
                                np.asarray(weights), zeta)) as pool:
            pool.map(_worker_run, [(0, 1)] * workers)  # every worker set up
            t = time.perf_counter()
            res = pool.map(_worker_run, ranges, chunksize=1)
            elw = time.perf_counter() - t
        total = per_worker * workers
        multi = dict(value=total / elw, cores=workers, centres=total, seconds=elw,
                     max_worker_seconds=max(r[3] for r in res))
    finally:
        for f in os.listdir(tmp):
            os.remove(os.path.join(tmp, f))
        os.rmdir(tmp)
    head = multi if multi is not None else single
    return dict(value=head["value"], unit="envs/s", cores=head["cores"], kind="port",
                sample="%d environments of the benchmark box (%d of %d per process, consecutive "
                       "ranges) on %d cores, energy + forces + virial, FP64; one core: the first %d"
                       % (head["centres"], per_worker if multi else n_centres, b.n_atoms,
                          head["cores"], n_centres),
                seconds=head["seconds"], single_core=single, all_cores=multi, nproc=nproc,
                cpus_allowed=allowed, cpu_model=_cpu_model())


if __name__ == "__main__":
    main()
