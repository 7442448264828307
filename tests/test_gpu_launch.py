"""GPU tests of the launch shapes: a run whose batches replay one captured
HIP graph, each replay popping its batch from the device batch queue, equals
the run launched kernel by kernel bit for bit (estimator, brute force with
cascades, several engines); the batch size, the photons per lane of the line
kernel and the resident workgroups per CU change the estimator's sums only in
their last bits; the autotuner returns a shape in its search space."""
import numpy as np
import pytest

from xmimsim_amd import api, autotune, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def default_launch():
    api.set_launch()
    yield
    api.set_launch()


def _steel(n):
    return workloads.single_layer_steel(n_photons_line=n)


def _run(prep, seed, **launch):
    api.set_launch(**launch)
    out = prep.run(seed=seed)
    st = api.get_stats()
    api.set_launch()
    return out, st


def test_graph_equals_eager_estimator():
    """4.6 batches of 2^14 histories, the last partial: graph replays equal
    the launches one by one, bit for bit, and a replay per batch ran."""
    prep = api.Prepared(workloads.three_layer_kramers(n_photons_interval=150, n_photons_line=150))
    prep.compute_solid_angle(hits_per_single=200, seed=3)
    print("graph vs eager, estimator", flush=True)
    (ch0, vr0, _, s0), st0 = _run(prep, 9, batch_log2=14)
    (ch1, vr1, _, s1), st1 = _run(prep, 9, batch_log2=14, use_graph=1)
    assert s0.histories == s1.histories > 4 * 2 ** 14
    assert st0.batches == st1.batches == -(-s0.histories // 2 ** 14) and st0.graph_replays == 0
    assert st1.graph_replays == st1.batches
    assert np.array_equal(ch0, ch1) and np.array_equal(vr0, vr1)
    assert vr0.sum() > 0
    prep.close()


def test_graph_equals_eager_brute():
    """Brute force with both cascades: the queue sets and offspring lists
    swapped per order are part of the captured graph."""
    inp = api.Input(layers=[api.Layer([26, 50], [0.5, 0.5], 5.0, 0.001)],
                    discrete=[api.DiscreteLine(40.0, 5e8, 5e8)], n_photons_line=50000, area_detector=3.0,
                    gain=0.02, nchannels=2500)
    prep = api.Prepared(inp, options=api.default_options(use_variance_reduction=0))
    print("graph vs eager, brute force", flush=True)
    (ch0, _, bf0, s0), _ = _run(prep, 4, batch_log2=13)
    (ch1, _, bf1, s1), st1 = _run(prep, 4, batch_log2=13, use_graph=1)
    assert st1.graph_replays == st1.batches == 7
    assert s0.histories == s1.histories
    assert np.array_equal(ch0, ch1) and np.array_equal(bf0, bf1)
    assert bf0.sum() > 0
    prep.close()


def test_graph_batch_engines():
    """Several engines capturing and replaying their samples' graphs at once
    give each sample's single-run result."""
    inps = workloads.batch_samples(5, seed=3, budget=(3e4, 8e4), n_elements=(2, 5), zrange=(20, 40))
    Zs = sorted({z for inp in inps for l in inp.layers for z in l.Z})
    tables = api.SharedTables(Zs)
    grid = api.Prepared(inps[0], tables=tables)
    grid.compute_solid_angle(hits_per_single=200, seed=3)
    preps = []
    for inp in inps:
        p = api.Prepared(inp, tables=tables)
        p.solid_angle = grid.solid_angle
        p._sa_owner = "borrowed"
        preps.append(p)
    print("graph, batch of engines", flush=True)
    api.set_launch(batch_log2=14, use_graph=1)
    out, _ = api.run_batch(preps, tables, seed=21, n_streams=3)
    api.set_launch(batch_log2=14)
    for p, (ch, vr) in zip(preps, out):
        c1, v1, _, _ = p.run(seed=21)
        assert np.array_equal(ch, c1) and np.array_equal(vr, v1)
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()


@pytest.mark.parametrize("launch", [{"batch_log2": 15}, {"lines_kp": 2}, {"lines_kp": 4}, {"wg_per_cu": 2},
                                    {"batch_log2": 16, "lines_kp": 4, "wg_per_cu": 3, "use_graph": 1}])
def test_launch_shapes_agree(launch):
    """Other shapes regroup the estimator's per-wave sums: every channel and
    history row within 1e-12 of the default shape's."""
    prep = api.Prepared(_steel(100000))
    prep.compute_solid_angle(hits_per_single=200, seed=3)
    print("launch shape", launch, flush=True)
    (ch0, vr0, _, _), _ = _run(prep, 13)
    (ch1, vr1, _, _), _ = _run(prep, 13, **launch)
    for a, b in ((ch1, ch0), (vr1, vr0)):
        scale = np.abs(b).max()
        assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1e-3 * scale)), launch
    prep.close()


def test_autotune_picks_a_shape_of_its_space(tmp_path):
    prep = api.Prepared(_steel(20000))
    prep.compute_solid_angle(hits_per_single=200, seed=3)
    print("autotune", flush=True)
    best, table = autotune.tune(prep, sample_histories=2 ** 16, batch_sizes=(14, 15), seed=5,
                                cache=str(tmp_path / "launch.json"))
    assert best in [row["launch"] for row in table]
    assert all(row["ms"] > 0 for row in table)
    assert api.get_launch() == autotune.DEFAULT
    again, rows = autotune.tune(prep, sample_histories=2 ** 16, batch_sizes=(14, 15), seed=5,
                                cache=str(tmp_path / "launch.json"))
    assert again == best and rows == []  # from the cache
    prep.close()
