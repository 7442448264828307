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
                    choices=["three_layer_kramers", "single_layer_steel", "quantify", "batch", "tables", "escape"])
    ap.add_argument("--batch-samples", type=int, default=1024, help="batch: independent samples")
    ap.add_argument("--streams", type=int, default=4, help="batch: engines (HIP streams) serving the sample queue")
    ap.add_argument("--train-samples", type=int, default=32768, help="quantify: Monte Carlo training samples")
    ap.add_argument("--train-steps", type=int, default=8000, help="quantify: Adam steps")
    ap.add_argument("--quant-samples", type=int, default=64, help="quantify: unknown samples")
    ap.add_argument("--photons", type=float, default=1e7, help="photons per interval and per line")
    ap.add_argument("--grid-hits", type=int, default=5000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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


def vr_bytes_per_event(inp, layer_events, n_trials=2.0):
    """Algorithmic bytes of one detection-weight evaluation (8 B per table
    scalar needed, 8 B per tally add), summed over the layers it happens in,
    split by estimator kernel:
      geometry  32 (solid-angle bilinear lookup)
      scatter   E*[16 (form factor) + 16*L*E (attenuation at the Compton
                energy) + 16 (scattering function) + 16*n_trials (Doppler
                ICDF)] + 8*2E (Rayleigh and Compton tallies)
      lines     E*[72 (shell cross sections) + A*(16 + 8*L)] + 8*2E*A
    E elements of the layer, L layers on the escape path, A active lines per
    element.  Returns {kernel: total bytes} and the per-event total."""
    total = {"geom": 0.0, "scatter": 0.0, "lines": 0.0}
    n_ev = sum(layer_events)
    if n_ev == 0:
        return total, 0.0
    active = _active_lines(inp)
    for l, ev in enumerate(layer_events):
        if ev == 0:
            continue
        layer = inp.layers[l]
        E = len(layer.Z)
        L = l + 1  # reflection geometry: the escape path crosses the layers above
        A = sum(active[z] for z in layer.Z) / E
        total["geom"] += ev * 32
        total["scatter"] += ev * (E * (16 + 16 * L * E + 16 + 16 * n_trials) + 8 * 2 * E)
        total["lines"] += ev * (E * (72 + A * (16 + 8 * L)) + 8 * 2 * E * A)
    return total, sum(total.values()) / n_ev


def _fp64_weights():
    with open(os.path.join(ROOT, "profiles", "fp64_op_weights.json")) as f:
        return json.load(f)["weights"]


def vr_flops_per_event(inp, layer_events, n_trials=2.0):
    """Algorithmic FP64 flops of the scatter and line phases of one
    detection-weight evaluation (FMA = 2, add/mul = 1, math routines weighted
    as in profiles/fp64_op_weights.json), summed over the layers the events
    happen in.  E elements of the layer, L layers on the escape path, A_tot
    active lines of the layer, n_trials Doppler trials:
      lines    per line: 2L (attenuation sum) + exp + 3 (wf * P_shell * g *
               e^-mu) + 1 (sum over photons); per event 3 + div (g)
      scatter  per event: sin, cos (theta), cos (phi), sin (theta/2) and 12 +
               div (Thomson and Klein-Nishina factors); per element:
               Rayleigh 5 div + 11 (form factor, cross section, weight,
               channel); Compton n_trials (4 div + 11) (Doppler) + log + exp
               + 4 (energy context) + 7L (attenuation) + exp + 4 div + 12
               (scattering function, cross section, weight, channel)
    Returns {kernel: total flops}."""
    w = _fp64_weights()
    ex, lg, dv, sn, cs = w["exp"], w["log"], w["div"], w["sin"], w["cos"]
    active = _active_lines(inp)
    tot = {"scatter": 0.0, "lines": 0.0}
    for l, ev in enumerate(layer_events):
        if ev == 0:
            continue
        layer = inp.layers[l]
        E = len(layer.Z)
        L = l + 1
        A_tot = sum(active[z] for z in layer.Z)
        tot["lines"] += ev * (A_tot * (2 * L + ex + 4) + 3 + dv)
        compton = n_trials * (4 * dv + 11) + lg + ex + 4 + 7 * L + ex + 4 * dv + 12
        tot["scatter"] += ev * (2 * sn + 2 * cs + 12 + dv + E * (5 * dv + 11 + compton))
    return tot


def _active_lines(inp):
    """Fluorescence lines the estimator evaluates per element (the engine's own
    active-line count, from the physics provider)."""
    from xmimsim_amd import api
    lib = api.load_library()
    return {z: lib.xmi_provider_active_lines(z, 219) for z in inp.unique_Z()}


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


def _measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel of known size), from the
    committed profile summary of this workload, if there is one."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("bytes_per_launch")


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
                   "elements": len(Zs), "solid_angle_grid": f"shared, 1024x1024x{args.grid_hits}"},
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


def quantify(args):
    """Surrogate-screened vs plain Monte Carlo quantification of single-layer
    samples: training data from the GPU Monte Carlo, training of the BF16
    ensemble on the GPU, then both quantification loops on the same unknowns."""
    from xmimsim_amd import quant
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    r = quant.run_pipeline(n_train=args.train_samples, photons_train=100_000, steps=args.train_steps,
                           n_quant=args.quant_samples, photons_quant=1_000_000, seed=args.seed % 100000, log=log)
    out = {
        "metric": "quantified samples/s (surrogate-screened vs plain MC)",
        "value": round(r["screened"]["samples_per_s"], 3),
        "unit": "samples/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "bf16 (surrogate) / f64 (Monte Carlo)",
        "data": "synthetic (seeded single-layer samples, 3-12 elements, Z 11-92)",
        "config": {"workload": "quantify", "train_samples": args.train_samples, "train_photons": 100_000,
                   "train_steps": args.train_steps, "ensemble": 4, "quant_samples": args.quant_samples,
                   "quant_photons": 1_000_000},
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
    from xmimsim_amd import api
    inp = make_input(args)
    prep = api.Prepared(inp)
    lib = api.load_library()
    t = time.time()
    prep.compute_solid_angle(hits_per_single=args.grid_hits, seed=args.seed)
    grid_s = time.time() - t
    print(f"[bench] solid-angle grid {grid_s * 1e3:.1f} ms ({args.grid_hits} trials/cell)", file=sys.stderr, flush=True)

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
    if prof is not None and prof["scatter"] > 0:
        tot_bytes, per_event = vr_bytes_per_event(inp, prof["layer_events"])
        tot_flops = vr_flops_per_event(inp, prof["layer_events"])
        launches = max(prof["launches"], 1)
        kernel_ms = {k: round(prof[k], 2) for k in ("geom", "cone", "scatter", "lines", "step", "interactions")}
        # the dominant kernel of the step; the scatter and line kernels are
        # bound by FP64 vector issue (their table reads hit the caches: the
        # measured HBM traffic is far below the algorithmic bytes), so their
        # roofline is FP64 at the vector peak
        dom = max(("scatter", "lines"), key=lambda k: prof[k])
        names = {"scatter": "k_vr_scatter", "lines": "k_vr_lines"}
        sec = prof[dom] * 1e-3
        achieved = tot_flops[dom] / sec / 1e12
        # "mfma" names the compute roofline of the contract; the pipe is FP64
        # vector issue, whose peak equals the FP64 matrix peak on MI355X
        roofline = {"bound": "mfma", "pipe": "fp64 vector (VALU); FP64 vector and matrix peaks are both 78.6 TFLOP/s",
                    "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                    "traffic": _measured_traffic(names[dom], args.workload),
                    "kernel": names[dom] + " (detection weight: " + dom + ")",
                    "flops_per_launch": round(tot_flops[dom] / launches),
                    "avg_launch_ms": round(prof[dom] / launches, 4),
                    "algorithmic_bytes_per_launch": round(tot_bytes[dom] / launches),
                    "algorithmic_bytes_rate_GBps": round(tot_bytes[dom] / sec / 1e9, 1),
                    "estimator_bytes_per_event": round(per_event, 1), "events": int(sum(prof["layer_events"])),
                    "launches": int(launches), "solid_angle_fallbacks": int(prof["fallbacks"]),
                    "kernel_ms": kernel_ms}

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
                   "device_ms_per_step": round(dev_ms / max(args.steps, 1), 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
