#!/usr/bin/env python3
"""Throughput benchmark of the photon-history engine on one MI355X.

Workload (default): the three-layer sample with a Kramers continuum, 1-50 keV
in 0.1 keV steps (490 intervals) plus two W L lines, 1e7 photons per interval
and per line, 4 interactions, FP64 tallies in 2048 channels x 4 orders
(xmimsim_amd.workloads.three_layer_kramers).  One step = one xmi_main_msim
call over the whole workload (~4.9e9 photon histories), tables and the
solid-angle grid already resident on the GPU.

Printed: one JSON line with histories/s, the roofline of the detection-weight
kernel (measured with HIP events during the warm-up step) and the CPU baseline
(the oracle, i.e. the C/OpenMP restatement of the reference algorithm, timed
on a bounded sample of the same workload on this host's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="three_layer_kramers",
                    choices=["three_layer_kramers", "single_layer_steel", "quantify", "batch", "tables", "escape",
                             "brute", "train"])
    ap.add_argument("--batch-samples", type=int, default=1024, help="batch: independent samples")
    ap.add_argument("--streams", type=int, default=4, help="batch: engines (HIP streams) serving the sample queue")
    ap.add_argument("--train-samples", type=int, default=32768, help="quantify: Monte Carlo training samples")
    ap.add_argument("--train-steps", type=int, default=8000, help="quantify: Adam steps")
    ap.add_argument("--quant-samples", type=int, default=64, help="quantify: unknown samples")
    ap.add_argument("--plain-samples", type=int, default=0,
                    help="quantify: run the plain loop on the first N unknowns only (0: all)")
    ap.add_argument("--train-photons", type=int, default=100_000, help="quantify: photons per training sample")
    ap.add_argument("--plain", default="pymca", choices=["pymca", "gn"],
                    help="quantify: the plain Monte Carlo loop (the reference's, or Gauss-Newton)")
    ap.add_argument("--photons", type=float, default=1e7, help="photons per interval and per line")
    ap.add_argument("--grid-hits", type=int, default=5000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default="", metavar="K=V,...",
                    help="launch shape (xmi_msim_set_launch): batch_log2, use_graph, lines_kp, wg_per_cu")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shape with xmimsim_amd.autotune on a sample of the workload first")
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once (a reference kernel of known traffic for the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB: reads and writes exactly that
    many bytes, so the HBM counters of this kernel give their scale factor."""
    import torch
    n = int(mib * 2 ** 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "check": float(b[n - 1].item())}))


def make_input(args):
    from xmimsim_amd import workloads
    n = int(args.photons)
    if args.workload == "single_layer_steel":
        return workloads.single_layer_steel(n_photons_line=n)
    return workloads.three_layer_kramers(n_photons_interval=n, n_photons_line=n)


def estimator_work(workload):
    """Per-event work of the detection-weight estimator, counted by the oracle
    (the CPU restatement, same algorithm and random streams) on a sample of the
    workload and frozen in profiles/estimator_work_<workload>.json by
    oracle/count_work.py: E (elements of the event's layer), E*L and E*E*L
    (L = layers on the escape path), Doppler trials, fluorescence lines
    evaluated and their sum of L."""
    path = os.path.join(ROOT, "profiles", f"estimator_work_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["per_event"]


def vr_bytes_per_event(w):
    """Algorithmic bytes of one detection-weight evaluation (8 B per table
    scalar needed, 8 B per tally add), split by estimator kernel:
      geometry  32 (solid-angle bilinear lookup)
      scatter   per element 16 (form factor) + 16 (scattering function)
                + 16*L*E (attenuation at the Compton energy) + 48 per Doppler
                draw (two profile-CDF pairs bounding the accepted momenta and
                the inverse-table pair) + 16 (Rayleigh and Compton tallies)
      lines     per element 72 (shell cross sections); per line 16 (line
                data) + 8*L (attenuation per crossed layer) + 16 (tallies)"""
    E, E2L, tr = w["elements"], w["element2_layers"], w["doppler_trials"]
    A, AL = w["lines"], w["line_layers"]
    return {"geom": 32.0,
            "scatter": E * (16 + 16 + 16) + 16 * E2L + 48 * tr,
            "lines": 72 * E + A * (16 + 16) + 8 * AL}


def _fp64_weights():
    with open(os.path.join(ROOT, "profiles", "fp64_op_weights.json")) as f:
        return json.load(f)["weights"]


def vr_flops_per_event(w):
    """Algorithmic FP64 flops of the scatter and line phases of one
    detection-weight evaluation (FMA = 2, add/mul = 1, math routines weighted
    as in profiles/fp64_op_weights.json):
      lines    per line: 2L (attenuation sum) + exp + 3 (wf * P_shell * g *
               e^-mu) + 1 (sum over photons); per event 3 + div (g)
      scatter  per event: sin, cos (theta), cos (phi), sin (theta/2) and 12 +
               div (Thomson and Klein-Nishina factors); per element:
               Rayleigh 5 div + 11 (form factor, cross section, weight,
               channel); Compton per Doppler draw 3 div + 20 (bounds of the
               accepted momenta, their CDF values, the draw, the energy),
               then log + exp + 4 (energy context) + 7L (attenuation) + exp
               + 4 div + 12 (scattering function, cross section, weight,
               channel)"""
    wt = _fp64_weights()
    ex, lg, dv, sn, cs = wt["exp"], wt["log"], wt["div"], wt["sin"], wt["cos"]
    E, EL, tr = w["elements"], w["element_layers"], w["doppler_trials"]
    A, AL = w["lines"], w["line_layers"]
    return {"lines": A * (ex + 4) + 2 * AL + 3 + dv,
            "scatter": (2 * sn + 2 * cs + 12 + dv + E * (5 * dv + 11) + tr * (3 * dv + 20)
                        + E * (lg + 2 * ex + 4 * dv + 16) + 7 * EL)}


def cpu_baseline(args, prep_gpu):
    """Oracle (reference structure: one RNG stream per thread, thread 0 takes
    the remainder) on a bounded sample of the workload, grown until its
    transport takes about --cpu-seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from xmimsim_amd import api
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # a job's CPU share is what OMP_NUM_THREADS grants it, when set
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores)
    n = 50 if args.workload == "three_layer_kramers" else 5000
    for attempt in range(4):
        sample = make_input(argparse.Namespace(**{**vars(args), "photons": n}))
        prep = api.Prepared(sample)
        prep.solid_angle = prep_gpu.solid_angle
        prep._sa_owner = "borrowed"
        _, _, _, st = oracle.run(prep, seed=args.seed, rng_mode=oracle.RNG_PER_THREAD, n_threads=cores)
        prep.solid_angle = None
        prep.close()
        if st.seconds >= 0.5 * args.cpu_seconds or attempt == 3:
            break
        n = max(n + 1, int(n * min(8.0, args.cpu_seconds / max(st.seconds, 1e-3))))
    return {
        "value": st.histories / st.seconds,
        "unit": "histories/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{args.workload} with {n} photons per interval/line ({st.histories} histories, "
                  f"{st.seconds:.1f} s, transport only)",
    }


def _measured_traffic(workload):
    """HBM bytes per launch of every kernel, measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel of known size), from the
    committed profile summary of this workload, if there is one."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {k: v["bytes_per_launch"] for k, v in json.load(f)["kernels"].items()}


def apply_launch(args, prep=None):
    """Sets the launch shape of --launch, or the autotuner's for prep's
    workload with --autotune (outside the timed region); returns the shape
    and the tuning record."""
    from xmimsim_amd import api, autotune
    shape, tuned = dict(autotune.DEFAULT), None
    for kv in filter(None, args.launch.split(",")):
        k, v = kv.split("=")
        shape[k.strip()] = int(v)
    if args.autotune and prep is not None:
        t = time.time()
        shape, table = autotune.tune(prep, cache=None, seed=args.seed)
        tuned = {"seconds": round(time.time() - t, 2), "timed": table}
        print(f"[bench] autotune: {shape} in {tuned['seconds']} s", file=sys.stderr, flush=True)
    api.set_launch(**shape)
    return shape, tuned


def batch(args):
    """Independent samples (1-5 layers, 2-30 elements, budgets 1e5-1e7, one
    line each) through xmi_main_msim_batch: aggregate histories/s and the
    per-sample completion times, with one engine and with --streams engines.
    All samples share the default geometry, hence one solid-angle grid
    (1024 x 1024 x --grid-hits) whose range covers 5 cm of a light matrix."""
    import numpy as np
    from xmimsim_amd import api, workloads
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    inps = workloads.batch_samples(args.batch_samples, seed=args.seed % 100000)
    Zs = sorted({z for inp in inps for l in inp.layers for z in l.Z})
    t = time.time()
    tables = api.SharedTables(Zs)
    log(f"tables of {len(Zs)} elements in {time.time() - t:.1f} s")
    deep = api.Input(layers=[api.Layer([Zs[0]], [1.0], 1.0, 1.0)] * 5, discrete=[api.DiscreteLine(40.0, 1e9)],
                     n_photons_line=10)
    grid = api.Prepared(deep, tables=tables)
    grid.compute_solid_angle(hits_per_single=args.grid_hits, seed=args.seed)
    preps = []
    for inp in inps:
        p = api.Prepared(inp, tables=tables)
        p.solid_angle = grid.solid_angle
        p._sa_owner = "borrowed"
        preps.append(p)
    hist = sum(inp.n_photons_line for inp in inps)
    api.run_batch(preps[:8], tables, seed=args.seed, n_streams=args.streams)  # warm-up: engines, table upload
    # one shape for every sample: the autotuner's for the largest budget
    launch, tuned = apply_launch(args, max(preps, key=lambda p: p.inp.n_photons_line))
    runs = {}
    for ns in (1, args.streams):
        t = time.perf_counter()
        _, done = api.run_batch(preps, tables, seed=args.seed, n_streams=ns)
        dt = time.perf_counter() - t
        runs[ns] = {"seconds": round(dt, 3), "histories_per_s": round(hist / dt, 1),
                    "p50_ms": round(float(np.percentile(done, 50)), 1), "p99_ms": round(float(np.percentile(done, 99)), 1)}
        log(f"{ns} stream(s): {runs[ns]}")
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()
    tables.close()
    best = runs[args.streams]
    out = {
        "metric": "photon histories/s on 1 GPU (batched samples)",
        "value": best["histories_per_s"],
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(best["seconds"] * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded independent samples)",
        "config": {"workload": "batch", "samples": len(inps), "histories": int(hist), "streams": args.streams,
                   "elements": len(Zs), "solid_angle_grid": f"shared, 1024x1024x{args.grid_hits}",
                   "launch": launch, "autotune": tuned},
        "per_sample_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"]},
        "one_stream": runs[1],
    }
    print(json.dumps(out))


def tables(args):
    """Interaction tables of Z 11-92 (the element set of the batch and
    quantification workloads): the angle, Compton-profile and azimuth inverse
    CDFs on the GPU, with the host's light parts; elements/s.  The host
    generator is timed beside it on a sample of four elements."""
    from xmimsim_amd import api
    lib = api.load_library()
    Zs = list(range(11, 93))
    lib.xmi_tables_set_device(1)
    try:
        warm = api.SharedTables([26])  # device start-up outside the timed region
        warm.close()
        t0 = time.perf_counter()
        T = api.SharedTables(Zs)
        dt = time.perf_counter() - t0
        T.close()
    finally:
        lib.xmi_tables_set_device(-1)
    cpu = None
    if not args.no_cpu:
        sample = [14, 26, 50, 82]
        lib.xmi_tables_set_device(0)
        try:
            t0 = time.perf_counter()
            T = api.SharedTables(sample)
            dc = time.perf_counter() - t0
            T.close()
        finally:
            lib.xmi_tables_set_device(-1)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores)
        cpu = {"value": round(len(sample) / dc, 3), "unit": "elements/s", "cores": cores, "kind": "port",
               "sample": f"host generator on Z {sample} ({dc:.1f} s)"}
    print(json.dumps({
        "metric": "interaction tables generated (elements/s)", "value": round(len(Zs) / dt, 3),
        "unit": "elements/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(dt * 1e3, 1),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "analytic physics provider, Z 11-92",
        "config": {"workload": "tables", "elements": len(Zs), "energies": 400, "random_numbers": 2000,
                   "profile_points": 10001},
        "roofline": None, "cpu_baseline": cpu}))


def escape(args):
    """Escape ratios of a 5 mm Ge crystal with the reference's default options
    (1990 input energies from 1 keV by 0.1 keV, 5e5 photons each, 1999 Compton
    output energies): photons/s of the whole call (tables, set-up, transport,
    read-back) and of the device transport.  The oracle's restatement is timed
    beside it on a sample of the same run."""
    from xmimsim_amd import api
    lib = api.load_library()
    inp = api.Input(layers=[api.Layer([26], [1.0], 7.9, 0.5)], discrete=[api.DiscreteLine(20.0, 1e9, 0.0)],
                    crystal=[api.Layer([32], [1.0], 5.32, 0.5)])
    warm = api.escape_ratios(inp, n_photons=1000, n_input_energies=10, seed=args.seed)
    api.free_escape_ratios(warm[3])
    n_e, n_ph = 1990, 500_000
    t0 = time.perf_counter()
    Z, fluo, comp, h = api.escape_ratios(inp, n_photons=n_ph, n_input_energies=n_e, seed=args.seed)
    dt = time.perf_counter() - t0
    api.free_escape_ratios(h)
    st = api.get_stats()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = api.default_options(use_M_lines=0, use_cascade_auger=0, use_cascade_radiative=0,
                                use_variance_reduction=0, escape_ratios_mode=1)
        prep = api.Prepared(api.escape_input(inp), options=o)
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores)
        n_s, e_s = 20_000, 64
        t0 = time.perf_counter()
        oracle.escape_ratios(prep, n_s, e_s, 1.0, 3.1, 1999, 0.1, 0.1, seed=args.seed, n_threads=cores)
        dc = time.perf_counter() - t0
        prep.close()
        cpu = {"value": round(n_s * e_s / dc, 1), "unit": "photons/s", "cores": cores, "kind": "port",
               "sample": f"{e_s} input energies 1-196 keV x {n_s} photons ({dc:.1f} s)"}
    print(json.dumps({
        "metric": "escape-ratio photons/s (detector crystal)", "value": round(n_e * n_ph / dt, 1),
        "unit": "photons/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(dt * 1e3, 1),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "Ge crystal 5 mm, analytic physics provider",
        "config": {"workload": "escape", "input_energies": n_e, "photons_per_energy": n_ph,
                   "compton_output_energies": 1999, "device_transport_ms": round(st.transport_ms, 1),
                   "device_photons_per_s": round(n_e * n_ph / (st.transport_ms * 1e-3), 1) if st.transport_ms else None},
        "roofline": None, "cpu_baseline": cpu}))


# brute-force step kernel, per photon step: the record it reads (layer,
# energy, history id, direction, position, the attenuation of one layer) and
# writes (position, layer, element)
BRUTE_STEP_BYTES = 4 + 8 + 8 + 24 + 24 + 8 + 24 + 4 + 4


def brute(args):
    """The three-layer Kramers workload without variance reduction (the
    reference's --disable-variance-reduction): analogue transport, detector
    hits, Auger and radiative cascades.  histories/s of whole calls; the
    roofline of the analogue step kernel from its per-step record bytes over
    its HIP-event time; the oracle's brute-force restatement timed beside it."""
    from xmimsim_amd import api
    lib = api.load_library()
    inp = make_input(argparse.Namespace(**{**vars(args), "workload": "three_layer_kramers"}))
    opts = api.default_options(use_variance_reduction=0)
    prep = api.Prepared(inp, options=opts)
    launch, tuned = apply_launch(args, prep)
    prof = None
    for w in range(args.warmup):
        lib.xmi_msim_set_profiling(1 if w == 0 else 0)
        _, _, _, st = prep.run(seed=args.seed + 1000 + w)
        if w == 0:
            prof = (st.step_kernel_ms, st.interaction_kernel_ms, st.launches, st.photon_steps, st.detector_hits)
        print(f"[bench] warmup {w}: {st.histories} histories, {st.transport_ms:.1f} ms", file=sys.stderr, flush=True)
    lib.xmi_msim_set_profiling(0)
    hist = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, _, _, st = prep.run(seed=args.seed + k)
        hist += st.histories
        print(f"[bench] step {k}: {st.histories} histories, {st.photon_steps} steps, {st.detector_hits} hits, "
              f"{st.transport_ms:.1f} ms device, {st.total_ms:.1f} ms wall", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    roofline = None
    if prof is not None and prof[2] > 0:
        step_ms, inter_ms, launches, steps, hits = prof
        sec = step_ms * 1e-3 / launches
        bytes_launch = steps / launches * BRUTE_STEP_BYTES
        roofline = {"bound": "hbm", "achieved": round(bytes_launch / sec / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(bytes_launch / sec / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic": _measured_traffic("brute").get("k_step_brute"),
                    "kernel": "k_step_brute (analogue step, detector tests, hit tallies)",
                    "avg_launch_ms": round(sec * 1e3, 4), "algorithmic_bytes_per_launch": round(bytes_launch),
                    "bytes_per_step": BRUTE_STEP_BYTES, "steps_per_launch": round(steps / launches),
                    "launches": int(launches), "detector_hits": int(hits),
                    "kernel_ms": {"step": round(step_ms, 2), "interactions": round(inter_ms, 2)}}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores)
        n = 2000
        for attempt in range(4):
            sample = make_input(argparse.Namespace(**{**vars(args), "workload": "three_layer_kramers", "photons": n}))
            p2 = api.Prepared(sample, options=opts)
            _, _, _, ost = oracle.run(p2, seed=args.seed, rng_mode=oracle.RNG_PER_THREAD, n_threads=cores)
            p2.close()
            if ost.seconds >= 0.5 * args.cpu_seconds or attempt == 3:
                break
            n = max(n + 1, int(n * min(8.0, args.cpu_seconds / max(ost.seconds, 1e-3))))
        cpu = {"value": round(ost.histories / ost.seconds, 1), "unit": "histories/s", "cores": cores, "kind": "port",
               "sample": f"brute force, three_layer_kramers with {n} photons per interval/line "
                         f"({ost.histories} histories, {ost.seconds:.1f} s)"}
    prep.close()
    print(json.dumps({
        "metric": "photon histories/s on 1 GPU (brute force)", "value": round(hist / elapsed, 1),
        "unit": "histories/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / max(args.steps, 1), 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded sample composition and spectrum)",
        "config": {"workload": "brute", "sample": "three_layer_kramers", "photons_per_interval_and_line": int(args.photons),
                   "histories_per_step": int(hist / max(args.steps, 1)), "n_interactions": inp.n_interactions_trajectory,
                   "cascades": "auger+radiative", "launch": launch, "autotune": tuned},
        "roofline": roofline, "cpu_baseline": cpu}))


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def train(args):
    """Training throughput of the net-line surrogate at its training
    configuration: an ensemble of 4 MLPs 85 -> 4 x 512 (GELU) -> 256, batch
    4096, Adam, on device-resident synthetic data (32768 rows).  Reported:
    samples/s (rows x ensemble members per second) and, as the roofline, the
    algorithmic matrix flops of forward, input-gradient and weight-gradient
    products per step over the step time against the BF16 matrix peak."""
    import numpy as np
    from xmimsim_amd import surrogate as S
    dims = S.DEFAULT_DIMS
    nets, batch, n = 4, 4096, 32768
    rng = np.random.default_rng(args.seed)
    X = rng.normal(0, 1, (n, dims[0])).astype(np.float32)
    T = rng.normal(0, 1, (n, dims[-1])).astype(np.float32)
    M = (rng.random((n, dims[-1])) > 0.3).astype(np.float32)
    sg = S.Surrogate(n_nets=nets, dims=dims, max_batch=batch, seed=args.seed)
    sg.set_data(X, T, M)
    steps = max(args.steps, 1) * 100
    lr = np.full(steps, 1e-3, np.float32)
    for w in range(args.warmup):
        sg.train_resident(rng.integers(0, n, (20, batch)).astype(np.int32), lr[:20])
    idx = rng.integers(0, n, (steps, batch)).astype(np.int32)
    t0 = time.perf_counter()
    loss = sg.train_resident(idx, lr)
    dt = time.perf_counter() - t0
    sg.close()
    # matrix flops per step: forward B*N*K, weight gradient N*K*B, input
    # gradient B*K*N (not for the first layer), 2 flops each, per member
    fl = 0.0
    for l in range(len(dims) - 1):
        k, nn = dims[l], dims[l + 1]
        fl += 2.0 * batch * nn * k * (3 if l > 0 else 2)
    fl *= nets
    achieved = fl * steps / dt / 1e12
    print(json.dumps({
        "metric": "surrogate training samples/s (ensemble of 4)", "value": round(batch * nets * steps / dt, 1),
        "unit": "samples/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / steps, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulation and master weights)", "data": "synthetic (random rows, resident)",
        "config": {"workload": "train", "dims": list(dims), "ensemble": nets, "batch": batch, "rows": n,
                   "loss_first": float(loss[0].mean()), "loss_last": float(loss[-1].mean())},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": None,
                     "flops_per_step": fl, "kernel": "whole training step (GEMMs, epilogues, loss, Adam)"},
        "cpu_baseline": None}))


def quantify(args):
    """Surrogate-screened vs plain Monte Carlo quantification of single-layer
    samples: training data from the GPU Monte Carlo, training of the BF16
    ensemble on the GPU, then both quantification loops on the same unknowns."""
    from xmimsim_amd import quant
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    r = quant.run_pipeline(n_train=args.train_samples, photons_train=args.train_photons, steps=args.train_steps,
                           n_quant=args.quant_samples, photons_quant=1_000_000, seed=args.seed % 100000,
                           plain=args.plain, log=log, n_plain=args.plain_samples)
    out = {
        "metric": "quantified samples/s (surrogate-screened vs plain MC)",
        "value": round(r["screened"]["samples_per_s"], 3),
        "unit": "samples/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "bf16 (surrogate) / f64 (Monte Carlo)",
        "data": "synthetic (seeded single-layer samples, 3-12 elements, Z 11-92)",
        "config": {"workload": "quantify", "train_samples": args.train_samples, "train_photons": args.train_photons,
                   "train_steps": args.train_steps, "ensemble": 4, "quant_samples": args.quant_samples,
                   "quant_photons": 1_000_000, "plain": args.plain, "plain_samples": r["plain_samples"],
                   "train_histories": r["train_histories"]},
        "plain_mc": r["plain"],
        "screened": r["screened"],
        "speedup": round(r["speedup"], 2),
        "surrogate": {k: r[k] for k in ("tables_s", "generate_s", "train_s", "loss_first", "loss_last",
                                        "holdout_err_median", "holdout_err_p90")},
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "quantify":
        return quantify(args)
    if args.workload == "batch":
        return batch(args)
    if args.workload == "tables":
        return tables(args)
    if args.workload == "escape":
        return escape(args)
    if args.workload == "brute":
        return brute(args)
    if args.workload == "train":
        return train(args)
    from xmimsim_amd import api
    inp = make_input(args)
    prep = api.Prepared(inp)
    lib = api.load_library()
    t = time.time()
    prep.compute_solid_angle(hits_per_single=args.grid_hits, seed=args.seed)
    grid_s = time.time() - t
    print(f"[bench] solid-angle grid {grid_s * 1e3:.1f} ms ({args.grid_hits} trials/cell)", file=sys.stderr, flush=True)
    launch, tuned = apply_launch(args, prep)

    prof = None
    for w in range(args.warmup):
        lib.xmi_msim_set_profiling(1 if w == 0 else 0)
        _, _, _, st = prep.run(seed=args.seed + 1000 + w)
        if w == 0:
            prof = {"geom": st.vr_geom_ms, "cone": st.vr_cone_ms, "scatter": st.vr_scatter_ms,
                    "lines": st.vr_lines_ms, "step": st.step_kernel_ms, "interactions": st.interaction_kernel_ms,
                    "launches": st.launches, "layer_events": list(prep.layer_events),
                    "fallbacks": st.solid_angle_fallbacks}
        print(f"[bench] warmup {w}: {st.histories} histories, {st.transport_ms:.1f} ms", file=sys.stderr, flush=True)
    lib.xmi_msim_set_profiling(0)

    hist = 0
    t0 = time.perf_counter()
    dev_ms = 0.0
    for k in range(args.steps):
        _, _, _, st = prep.run(seed=args.seed + k)
        hist += st.histories
        dev_ms += st.transport_ms
        print(f"[bench] step {k}: {st.histories} histories, {st.transport_ms:.1f} ms device, {st.total_ms:.1f} ms wall",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    value = hist / elapsed

    roofline = None
    work = estimator_work(args.workload)
    if prof is not None and prof["scatter"] > 0 and work is not None:
        launches = max(prof["launches"], 1)
        events = int(sum(prof["layer_events"]))
        ev_launch = events / launches
        bpe = vr_bytes_per_event(work)
        fpe = vr_flops_per_event(work)
        kernel_ms = {k: round(prof[k], 2) for k in ("geom", "cone", "scatter", "lines", "step", "interactions")}
        names = {"scatter": "k_vr_scatter", "lines": "k_vr_lines", "geom": "k_vr_list", "cone": "k_vr_cone",
                 "step": "k_step", "interactions": "k_interact"}
        traffic = _measured_traffic(args.workload)
        # the dominant kernel of the step: the scatter and line kernels are
        # bound by FP64 vector issue and gather latency (their table reads hit
        # the caches: the measured HBM traffic is far below the algorithmic
        # bytes), so their roofline is FP64 at the vector peak
        dom = max(("scatter", "lines"), key=lambda k: prof[k])
        sec = prof[dom] * 1e-3 / launches
        achieved = fpe[dom] * ev_launch / sec / 1e12
        # measured HBM traffic of the four kernels with the most time per step
        hbm = {}
        for k in sorted(kernel_ms, key=lambda k: -prof[k])[:4]:
            b = traffic.get(names[k])
            # the last order launches no interactions
            n_k = launches * (inp.n_interactions_trajectory - 1) / inp.n_interactions_trajectory \
                if k == "interactions" else launches
            t = prof[k] * 1e-3 / max(n_k, 1)
            hbm[names[k]] = {"avg_launch_ms": round(t * 1e3, 4), "bytes_per_launch": b,
                             "GBps": round(b / t / 1e9, 1) if b else None,
                             "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4) if b else None}
        roofline = {"bound": "fp64 valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                    "traffic": traffic.get(names[dom]),
                    "kernel": names[dom] + " (detection weight: " + dom + ")",
                    "flops_per_launch": round(fpe[dom] * ev_launch),
                    "avg_launch_ms": round(sec * 1e3, 4),
                    "algorithmic_bytes_per_launch": round(bpe[dom] * ev_launch),
                    "algorithmic_bytes_rate_GBps": round(bpe[dom] * ev_launch / sec / 1e9, 1),
                    "events": events, "events_per_launch": round(ev_launch), "launches": int(launches),
                    "work_per_event": {k: round(v, 4) for k, v in work.items()},
                    "hbm_measured": hbm,
                    # both estimator kernels against the FP64 peak, by their counted flops
                    "fp64_frac_estimator": {names[k]: round(fpe[k] * ev_launch / (prof[k] * 1e-3 / launches)
                                                            / 1e12 / FP64_PEAK_TFLOPS, 4)
                                            for k in ("scatter", "lines")},
                    "solid_angle_fallbacks": int(prof["fallbacks"]), "kernel_ms": kernel_ms}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, prep)

    out = {
        "metric": "photon histories/s on 1 GPU",
        "value": round(value, 1),
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / max(args.steps, 1), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sample composition and spectrum)",
        "config": {"workload": args.workload, "photons_per_interval_and_line": int(args.photons),
                   "histories_per_step": int(hist / max(args.steps, 1)), "n_interactions": inp.n_interactions_trajectory,
                   "channels": inp.nchannels, "solid_angle_grid_ms": round(grid_s * 1e3, 1),
                   "device_ms_per_step": round(dev_ms / max(args.steps, 1), 2), "launch": launch,
                   "autotune": tuned},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
