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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="three_layer_kramers",
                    choices=["three_layer_kramers", "single_layer_steel"])
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
    scalar needed, 8 B per tally add), summed over the layers it happens in:
    32 (solid angle) + E*[16 (form factor) + 16*L*E (attenuation at the
    Compton energy) + 16 (scattering function) + 16*n_trials (Doppler ICDF)
    + 72 (shell cross sections) + A*(16 + 8*L)] + 8*(2E + 2E*A)."""
    total_events = sum(layer_events)
    if total_events == 0:
        return 0.0, 0.0
    # active lines per element: lines up to M5P5 with E >= 1 keV and a non-zero rate
    active = _active_lines(inp)
    total = 0.0
    for l, ev in enumerate(layer_events):
        if ev == 0:
            continue
        layer = inp.layers[l]
        E = len(layer.Z)
        L = l + 1  # reflection geometry: the escape path crosses the layers above
        A = sum(active[z] for z in layer.Z) / E
        b = 32 + E * (16 + 16 * L * E + 16 + 16 * n_trials + 72 + A * (16 + 8 * L)) + 8 * (2 * E + 2 * E * A)
        total += ev * b
    return total, total / total_events


def _active_lines(inp):
    """Fluorescence lines the estimator evaluates per element (the engine's own
    active-line count, from the physics provider)."""
    from xmimsim_amd import api
    lib = api.load_library()
    return {z: lib.xmi_provider_active_lines(z, 219) for z in inp.unique_Z()}


def cpu_baseline(args, prep_gpu):
    """Oracle (reference structure: one RNG stream per thread, thread 0 takes
    the remainder) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from xmimsim_amd import api
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # a job's CPU share is what OMP_NUM_THREADS grants it, when set
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cores)
    # pilot to size the sample
    pilot_n = 50 if args.workload == "three_layer_kramers" else 5000
    sample = make_input(argparse.Namespace(**{**vars(args), "photons": pilot_n}))
    prep = api.Prepared(sample)
    prep.solid_angle = prep_gpu.solid_angle
    prep._sa_owner = "borrowed"
    _, _, _, st = oracle.run(prep, seed=args.seed, rng_mode=oracle.RNG_PER_THREAD, n_threads=cores)
    # size from the pilot's transport time (its set-up time does not scale)
    pilot_s = max(st.seconds, 1e-3)
    n = max(pilot_n, int(pilot_n * args.cpu_seconds / pilot_s))
    sample = make_input(argparse.Namespace(**{**vars(args), "photons": n}))
    prep2 = api.Prepared(sample)
    prep2.solid_angle = prep_gpu.solid_angle
    prep2._sa_owner = "borrowed"
    _, _, _, st = oracle.run(prep2, seed=args.seed, rng_mode=oracle.RNG_PER_THREAD, n_threads=cores)
    prep.solid_angle = None
    prep2.solid_angle = None
    return {
        "value": st.histories / st.seconds,
        "unit": "histories/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{args.workload} with {n} photons per interval/line ({st.histories} histories, "
                  f"{st.seconds:.1f} s, transport only)",
    }


def main():
    args = parse()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
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
            prof = (st.vr_kernel_ms, st.step_kernel_ms, st.interaction_kernel_ms, st.launches,
                    list(prep.layer_events), st.histories, st.solid_angle_fallbacks)
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
    if prof is not None and prof[0] > 0:
        vr_ms, step_ms, int_ms, launches, layer_events, _, fallbacks = prof
        tot_bytes, per_event = vr_bytes_per_event(inp, layer_events)
        achieved = tot_bytes / (vr_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_varred (detection-weight estimator)",
                    "bytes_per_event": round(per_event, 1), "events": int(sum(layer_events)),
                    "launches": int(launches), "avg_launch_ms": round(vr_ms / max(launches, 1), 4),
                    "solid_angle_fallbacks": int(fallbacks),
                    "kernel_ms": {"k_varred": round(vr_ms, 2), "k_step": round(step_ms, 2),
                                  "interactions": round(int_ms, 2)}}

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
