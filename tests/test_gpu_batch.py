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
    for p in preps:
        p.solid_angle = None
        p.close()
    grid.close()
    tables.close()
