"""CPU tests of the launch-shape interface and the autotuner's search: range
checks of xmi_msim_set_launch, the sampled input, the coordinate descent and
its cache (timings from a stand-in function, no device)."""
import pytest

from xmimsim_amd import api, autotune, workloads


def test_set_launch_ranges(lib):
    try:
        api.set_launch(batch_log2=18, use_graph=1, lines_kp=4, wg_per_cu=3)
        assert api.get_launch() == {"batch_log2": 18, "use_graph": 1, "lines_kp": 4, "wg_per_cu": 3}
        for bad in ({"batch_log2": 11}, {"batch_log2": 22}, {"lines_kp": 5}, {"use_graph": 2}, {"wg_per_cu": -1}):
            with pytest.raises(ValueError):
                api.set_launch(**bad)
        assert api.get_launch()["batch_log2"] == 18  # a refused shape leaves the last one
    finally:
        api.set_launch()
    assert api.get_launch() == autotune.DEFAULT


def test_sample_input_scales_budgets():
    inp = workloads.three_layer_kramers(n_photons_interval=10_000_000, n_photons_line=10_000_000)
    total = autotune.histories(inp)
    assert total == 492 * 10_000_000
    s = autotune.sample_input(inp, 1 << 23)
    assert abs(autotune.histories(s) - (1 << 23)) < 0.01 * (1 << 23)
    assert s.layers is inp.layers and inp.n_photons_line == 10_000_000
    small = workloads.single_layer_steel(n_photons_line=1000)
    assert autotune.sample_input(small, 1 << 23) is small


def test_lines_kp_candidates():
    assert autotune.lines_kp_candidates(1) == autotune.lines_kp_candidates(4) == (2, 4)
    assert autotune.lines_kp_candidates(5) == (1,)
    assert autotune.lines_kp_candidates(12) == ()


def test_descent_finds_the_separable_optimum():
    """With a time that is a sum of one term per parameter, coordinate
    descent reaches the best value of every parameter."""
    cost = {"lines_kp": {0: 3.0, 2: 5.0, 4: 1.0}, "wg_per_cu": {0: 2.0, 2: 0.5, 3: 4.0},
            "batch_log2": {0: 1.0, 20: 0.2, 19: 3.0}, "use_graph": {0: 1.0, 1: 0.9}}
    calls = []

    def measure(shape):
        calls.append(dict(shape))
        return sum(cost[k][v] for k, v in shape.items())

    steps = [("lines_kp", (2, 4)), ("wg_per_cu", (2, 3)), ("batch_log2", (20, 19)), ("use_graph", (1,))]
    best = autotune.descend(measure, steps)
    assert best == {"lines_kp": 4, "wg_per_cu": 2, "batch_log2": 20, "use_graph": 1}
    assert len(calls) == 4 + 2 + 2 + 2 + 1


def test_tune_caches_by_signature(lib, tmp_path):
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=1000))
    path = str(tmp_path / "launch.json")
    seen = []

    def timer(shape):
        seen.append(dict(shape))
        return 10.0 - shape["lines_kp"] - (shape["batch_log2"] == 19)

    best, table = autotune.tune(prep, cache=path, timer=timer)
    assert best == {"batch_log2": 19, "use_graph": 0, "lines_kp": 4, "wg_per_cu": 0}
    assert len(table) == len(seen) and best in [r["launch"] for r in table]
    again, rows = autotune.tune(prep, cache=path, timer=timer)
    assert again == best and rows == [] and len(seen) == len(table)
    brute = api.Prepared(workloads.single_layer_steel(n_photons_line=1000),
                         options=api.default_options(use_variance_reduction=0))
    b, t = autotune.tune(brute, cache=path, timer=timer)
    assert t and all(r["launch"]["lines_kp"] == 0 and r["launch"]["wg_per_cu"] == 0 for r in t)
    assert b["batch_log2"] == 19
    prep.close()
    brute.close()
