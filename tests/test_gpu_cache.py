"""GPU tests of the HDF5 caches in the harness order of bin/xmimsim.c: a
grid and escape ratios computed on the GPU are stored, found again for the
same input, and give the same results as the freshly computed ones."""
import numpy as np
import pytest

from xmimsim_amd import api, cache, workloads

pytestmark = pytest.mark.gpu


def test_cached_solid_angle_grid_reproduces_run(tmp_path):
    path = str(tmp_path / "sa.h5")
    inp = workloads.single_layer_steel(n_photons_line=5000)
    prep = api.Prepared(inp)
    grid, r, t = prep.compute_solid_angle(hits_per_single=50, seed=4)
    cache.store_solid_angle(path, inp, grid, r, t)
    a = prep.run(seed=9)
    prep2 = api.Prepared(workloads.single_layer_steel(n_photons_line=5000))
    found = cache.find_solid_angle(path, prep2.inp)
    assert found is not None
    g2, r2, t2, ptr = found
    assert np.array_equal(g2, grid) and np.array_equal(r2, r) and np.array_equal(t2, t)
    prep2.set_solid_angle(ptr, owner="engine")
    b = prep2.run(seed=9)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    prep.close()
    prep2.close()


def test_cached_escape_ratios_reproduce_response(tmp_path):
    path = str(tmp_path / "esc.h5")
    inp = workloads.single_layer_steel(n_photons_line=5000)
    Z, fluo, comp, h = api.escape_ratios(api.escape_input(inp), n_photons=20000, n_input_energies=60,
                                         input_energy_min=1.0, input_energy_delta=0.5,
                                         n_compton_output_energies=80, seed=3)
    cache.store_escape_ratios(path, inp, h)
    h2 = cache.find_escape_ratios(path, inp)
    assert h2 is not None
    e1, e2 = h.contents, h2.contents
    n = e1.n_fluo_input_energies * 109 * e1.n_elements
    assert np.array_equal(np.ctypeslib.as_array(e1.fluo_escape_ratios, shape=(n,)),
                          np.ctypeslib.as_array(e2.fluo_escape_ratios, shape=(n,)))
    prep = api.Prepared(inp)
    prep.compute_solid_angle(hits_per_single=50, seed=4)
    ch, vr, _, _ = prep.run(seed=9)
    o = api.default_options()
    o.use_escape_peaks, o.use_default_seeds = 1, 1
    r1 = api.detector_response(prep, ch, escape_ratios=e1, options=o)
    r2 = api.detector_response(prep, ch, escape_ratios=e2, options=o)
    assert np.array_equal(r1, r2) and r1[1].sum() > 0
    api.free_escape_ratios(h)
    api.free_escape_ratios(h2)
    prep.close()
