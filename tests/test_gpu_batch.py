"""GPU test of the batched engine: every sample's result from
xmi_main_msim_batch (several engines sharing the sample queue) equals, bit for
bit, xmi_main_msim of that sample alone with the same seed."""
import numpy as np
import pytest

from xmimsim_amd import api, workloads

pytestmark = pytest.mark.gpu


def test_batch_equals_single_runs():
    inps = workloads.batch_samples(6, seed=11, budget=(2e4, 1e5), n_elements=(2, 6), zrange=(20, 40))
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
    out, done = api.run_batch(preps, tables, seed=77, n_streams=3)
    assert np.all(done > 0)
    for p, (ch, vr) in zip(preps, out):
        c1, v1, _, _ = p.run(seed=77)
        assert np.array_equal(ch, c1)
        assert np.array_equal(vr, v1)
        assert vr.sum() > 0
    # one seed per sample
    seeds = [1000 + 17 * k for k in range(len(preps))]
    out, _ = api.run_batch(preps, tables, seed=0, n_streams=3, seeds=seeds)
    for p, sd, (ch, vr) in zip(preps, seeds, out):
        c1, v1, _, _ = p.run(seed=sd)
        assert np.array_equal(ch, c1) and np.array_equal(vr, v1)
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()
    tables.close()


def _deep_grid(tables, Zs, hits=200, seed=3):
    """one grid whose depth range covers every sample (5 cm of a light matrix)"""
    deep = api.Input(layers=[api.Layer([Zs[0]], [1.0], 1.0, 1.0)] * 5, discrete=[api.DiscreteLine(40.0, 1e9)],
                     n_photons_line=10)
    grid = api.Prepared(deep, tables=tables)
    grid.compute_solid_angle(hits_per_single=hits, seed=seed)
    return grid


def test_batch_of_full_distribution_samples_matches_oracle(orc):
    """Samples drawn from the full batched-workload distribution (1-5 layers,
    2-30 elements of Z 11-92, Dirichlet(0.5), 1-20 g/cm3, 1 um - 1 cm,
    10-40 keV) at reduced budgets: every sample of the batch equals the
    oracle on the same random streams (relative 1e-9, every tally)."""
    inps = workloads.batch_samples(16, seed=23, budget=(1e3, 4e3))
    print("layers", [len(i.layers) for i in inps], "elements", [max(len(l.Z) for l in i.layers) for i in inps],
          flush=True)
    Zs = sorted({z for inp in inps for l in inp.layers for z in l.Z})
    tables = api.SharedTables(Zs)
    grid = _deep_grid(tables, Zs)
    preps = []
    for inp in inps:
        p = api.Prepared(inp, tables=tables)
        p.solid_angle = grid.solid_angle
        p._sa_owner = "borrowed"
        preps.append(p)
    out, _ = api.run_batch(preps, tables, seed=91, n_streams=4)
    assert api.batch_stats().histories == sum(inp.n_photons_line for inp in inps)
    for k, (p, (ch, vr)) in enumerate(zip(preps, out)):
        och, ovr, _, st = orc.run(p, seed=91, rng_mode=orc.RNG_PER_HISTORY)
        assert st.histories == p.inp.n_photons_line
        scale = max(ovr.max(), 1e-300)
        bad = np.abs(vr - ovr) > 1e-9 * np.maximum(np.abs(ovr), 1e-6 * scale)
        assert not bad.any(), (k, np.argwhere(bad)[:5])
        cs = max(och.max(), 1e-300)
        assert not (np.abs(ch - och) > 1e-9 * np.maximum(np.abs(och), 1e-6 * cs)).any(), k
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()
    tables.close()


def test_full_batch_runs_every_history():
    """The full 1024-sample batch: the call succeeds (the engine fails a
    sample whose device error counters are non-zero), runs the sum of the
    budgets, and every sample scores lines."""
    inps = workloads.batch_samples(1024, seed=7)
    Zs = sorted({z for inp in inps for l in inp.layers for z in l.Z})
    tables = api.SharedTables(Zs)
    grid = _deep_grid(tables, Zs)
    preps = []
    for inp in inps:
        p = api.Prepared(inp, tables=tables)
        p.solid_angle = grid.solid_angle
        p._sa_owner = "borrowed"
        preps.append(p)
    out, done = api.run_batch(preps, tables, seed=5, n_streams=4)
    st = api.batch_stats()
    print("histories", st.histories, "wall ms", st.total_ms, flush=True)
    assert st.histories == sum(inp.n_photons_line for inp in inps)
    assert np.all(done > 0)
    assert all(vr.sum() > 0 and ch[-1].sum() > 0 for ch, vr in out)
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()
    tables.close()
