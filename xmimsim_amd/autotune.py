"""Launch-shape autotuner of the transport engine (no counterpart in the
reference, whose OpenMP loop has no launch shape).

It times the engine on a bounded sample of the caller's own workload (the
same sample, the same tables and solid-angle grid, fewer photons) for a few
launch shapes of xmi_msim_set_launch and keeps the fastest:

- ``lines_kp``: photons per lane of the estimator's line kernel, so 64 kp
  histories per wavefront share each line's wave sum;
- ``wg_per_cu``: resident workgroups per CU of the estimator's grid-stride
  kernels (0: as many as the occupancy allows);
- ``batch_log2``: histories per batch, the granularity at which histories are
  regrouped by interaction type and by layer;
- ``use_graph``: a batch's launches one by one, or as replays of one captured
  HIP graph whose first node pops the batch from the device batch queue.

The search is a coordinate descent from the default shape in that order, each
parameter over its candidates with the others fixed, every candidate timed
twice (device time of the transport, the minimum kept).  The choice is cached
per workload signature (layers, elements per layer, lines and intervals,
interactions, options) in a JSON file.  Shapes change the estimator's sums
only in their last bits (the order of the per-wave sums), so a tuned run
agrees with the default shape's to about 1e-12 relative.
"""
import copy
import json
import os
from typing import Dict, List, Optional, Sequence, Tuple

from . import api

DEFAULT = {"batch_log2": 0, "use_graph": 0, "lines_kp": 0, "wg_per_cu": 0}
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "xmimsim_amd", "launch.json")


def _nl_bound(n_layers: int) -> int:
    for b in (1, 2, 4, 8):
        if n_layers <= b:
            return b
    return 16


def lines_kp_candidates(n_layers: int) -> Tuple[int, ...]:
    """Photons per lane, besides the default, that the line kernel is built
    for at this layer count (default 3 up to four layers, 2 up to eight, 1)."""
    nl = _nl_bound(n_layers)
    return (2, 4) if nl <= 4 else ((1,) if nl <= 8 else ())


def signature(prep: "api.Prepared") -> str:
    inp, o = prep.inp, prep.options
    layers = "/".join(str(len(l.Z)) for l in inp.layers)
    return (f"layers={layers};lines={len(inp.discrete)};intervals={max(len(inp.continuous) - 1, 0)};"
            f"n_int={inp.n_interactions_trajectory};vr={o.use_variance_reduction};"
            f"casc={o.use_cascade_auger}{o.use_cascade_radiative};adv={o.use_advanced_compton}")


def histories(inp: "api.Input") -> int:
    return inp.n_photons_line * len(inp.discrete) + inp.n_photons_interval * max(len(inp.continuous) - 1, 0)


def sample_input(inp: "api.Input", n: int) -> "api.Input":
    """The input with its photon budgets scaled to about n histories."""
    total = histories(inp)
    if total <= n:
        return inp
    f = n / total
    out = copy.copy(inp)
    out.n_photons_line = max(1, int(round(inp.n_photons_line * f)))
    out.n_photons_interval = max(1, int(round(inp.n_photons_interval * f)))
    return out


class _Borrowed:
    """Table handle of another Prepared (Prepared(tables=...) only reads hdf5F)."""

    def __init__(self, hdf5F):
        self.hdf5F = hdf5F


def _load_cache(path: Optional[str]) -> Dict:
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _store_cache(path: Optional[str], data: Dict):
    if not path:
        return
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def descend(measure, steps) -> Dict[str, int]:
    """Coordinate descent from DEFAULT: each (name, candidates) step in turn
    keeps the fastest value with the others fixed; measure(shape) -> ms."""
    shape = dict(DEFAULT)
    for name, cands in steps:
        best_v, best_t = shape[name], measure(shape)
        for v in cands:
            t = measure(dict(shape, **{name: v}))
            if t < best_t:
                best_v, best_t = v, t
        shape[name] = best_v
    return shape


def tune(prep: "api.Prepared", sample_histories: int = 1 << 23, seed: int = 1,
         batch_sizes: Sequence[int] = (20, 19), wg_per_cu: Sequence[int] = (2, 3),
         cache: Optional[str] = CACHE, repeats: int = 2, timer=None) -> Tuple[Dict[str, int], List[Dict]]:
    """The fastest launch shape for prep's workload and the table of timed
    shapes ({"launch": shape, "ms": device time}); (shape, []) when the
    cache holds the signature.  Candidates are tried besides the defaults
    (2^21 histories per batch, kernel by kernel, the default photons per lane
    and the occupancy's workgroups).  The engine's launch shape in force
    before the call is restored; apply the result with
    api.set_launch(**shape).  timer(shape) -> ms replaces the timed runs
    (tests of the search)."""
    key = f"{signature(prep)};sample={sample_histories}"
    data = _load_cache(cache)
    if key in data:
        return data[key]["launch"], []
    timed: Dict[Tuple, float] = {}
    table: List[Dict] = []
    sample = None
    if timer is None:
        sample = api.Prepared(sample_input(prep.inp, sample_histories), options=prep.options,
                              tables=_Borrowed(prep.hdf5F))
        if prep.solid_angle is not None:
            sample.solid_angle = prep.solid_angle
            sample._sa_owner = "borrowed"

        previous = api.get_launch()

        def timer(shape):
            best = float("inf")
            try:
                api.set_launch(**shape)
                for _ in range(repeats):
                    best = min(best, sample.run(seed=seed)[3].transport_ms)
            finally:
                api.set_launch(**previous)
            return best

    def measure(shape: Dict[str, int]) -> float:
        k = tuple(sorted(shape.items()))
        if k not in timed:
            timed[k] = timer(shape)
            table.append({"launch": dict(shape), "ms": timed[k]})
        return timed[k]

    steps = [("lines_kp", lines_kp_candidates(len(prep.inp.layers))), ("wg_per_cu", tuple(wg_per_cu)),
             ("batch_log2", tuple(batch_sizes)), ("use_graph", (1,))]
    if not prep.options.use_variance_reduction:
        steps = [s for s in steps if s[0] in ("batch_log2", "use_graph")]
    try:
        if sample is not None:
            timer(DEFAULT)  # the first run uploads the tables and the grid: not a timing
        shape = descend(measure, steps)
    finally:
        if sample is not None:
            sample.solid_angle = None
            sample.close()
    data = _load_cache(cache)
    data[key] = {"launch": shape, "ms": timed[tuple(sorted(shape.items()))]}
    _store_cache(cache, data)
    return shape, table
