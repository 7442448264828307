#!/usr/bin/env python3
"""Counts the detection-weight estimator's work on a sample of a benchmark
workload with the oracle (the CPU restatement, same algorithm and same random
streams as the GPU engine) and freezes the per-event averages that the
benchmark's algorithmic byte and flop accounting uses:

    python3 oracle/count_work.py [--workload three_layer_kramers] [--photons 200]

writes profiles/estimator_work_<workload>.json.  Test infrastructure: the
benchmark reads the JSON, never this script or the oracle."""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="three_layer_kramers", choices=["three_layer_kramers", "single_layer_steel"])
    ap.add_argument("--photons", type=int, default=200, help="photons per interval and per line of the sample")
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--grid-hits", type=int, default=200)
    args = ap.parse_args()
    import oracle
    from xmimsim_amd import api, workloads
    if args.workload == "single_layer_steel":
        inp = workloads.single_layer_steel(n_photons_line=args.photons)
    else:
        inp = workloads.three_layer_kramers(n_photons_interval=args.photons, n_photons_line=args.photons)
    prep = api.Prepared(inp)
    oracle.solid_angle(prep, hits=args.grid_hits, seed=3)
    t = time.time()
    _, _, _, st = oracle.run(prep, seed=args.seed, rng_mode=oracle.RNG_PER_HISTORY)
    w = oracle.work()
    prep.close()
    ev = w["events"]
    per_event = {
        "elements": w["elements"] / ev,
        "element_layers": w["element_layers"] / ev,
        "element2_layers": w["element2_layers"] / ev,
        "doppler_trials": w["doppler_trials"] / ev,
        "doppler_trials_per_compton_term": w["doppler_trials"] / w["elements"],
        "lines": w["lines"] / ev,
        "line_layers": w["line_layers"] / ev,
    }
    out = {
        "workload": args.workload,
        "command": "python3 oracle/count_work.py --workload %s --photons %d" % (args.workload, args.photons),
        "sample": {"photons_per_interval_and_line": args.photons, "histories": st.histories, "seed": args.seed,
                   "estimator_calls": st.vr_calls, "events_scored": ev, "seconds": round(time.time() - t, 1)},
        "totals": w,
        "per_event": per_event,
        "meaning": {
            "elements": "E, elements of the event's layer",
            "element_layers": "E * L, L = layers on the escape path towards the detector",
            "element2_layers": "E * E * L",
            "doppler_trials": "Doppler words drawn by the event's Compton terms",
            "lines": "fluorescence lines evaluated (line energy above 1 keV, shell cross section non-zero)",
            "line_layers": "sum of L over those lines",
        },
    }
    path = os.path.join(ROOT, "profiles", "estimator_work_%s.json" % args.workload)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out["per_event"]))


if __name__ == "__main__":
    main()
