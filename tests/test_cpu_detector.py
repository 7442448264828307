"""CPU tests pinning the oracle's detector response (src/xmi_detector_f.F90):
a pure Gaussian line above 50 keV keeps its counts and has the resolution
FWHM^2 = noise^2 + 2.3548^2 * 3.85 eV * Fano * E; the window and crystal give
exp(-mu rho d) and 1 - exp(-mu rho d) at the channel energy."""
import math

import numpy as np
import pytest

from xmimsim_amd import api, workloads


def _prep(det_absorbers=(), crystal=(), gain=0.05, nch=2048, detector_type=0):
    inp = workloads.single_layer_steel(n_photons_line=10)
    inp.det_absorbers = list(det_absorbers)
    inp.crystal = list(crystal)
    inp.gain = gain
    inp.nchannels = nch
    inp.detector_type = detector_type
    return api.Prepared(inp)


def test_gaussian_line_keeps_counts_and_has_detector_fwhm(orc):
    prep = _prep()
    x = np.zeros(2048)
    i0 = 1200  # 60 keV: pure Gaussian
    x[i0] = 1e6
    y = orc.detector_convolute(prep, x)
    assert y.sum() == pytest.approx(1e6, rel=1e-12)
    E = np.arange(2048) * 0.05
    mean = (y * E).sum() / y.sum()
    var = (y * (E - mean) ** 2).sum() / y.sum()
    fwhm = 2.3548 * math.sqrt(var)
    inp = prep.inp
    expect = math.sqrt(inp.noise ** 2 + 2.3548 ** 2 * 3.85e-3 * inp.fano * 60.0)
    assert mean == pytest.approx(60.0, abs=1e-9)
    assert fwhm == pytest.approx(expect, rel=2e-3)  # float literals and the 0.05 keV sampling
    prep.close()


def test_low_energy_tail_moves_counts_down(orc):
    prep = _prep()
    x = np.zeros(2048)
    x[200] = 1e6  # 10 keV: Gaussian + tail (Si(Li))
    y = orc.detector_convolute(prep, x)
    assert y.sum() == pytest.approx(1e6, rel=1e-9)
    assert y[:180].sum() > 1e-3 * 1e6 and y[230:].sum() < 1e-6 * 1e6
    prep.close()


def test_window_and_crystal_absorption(orc):
    be = api.Layer([4], [1.0], 1.85, 0.002)
    si = api.Layer([14], [1.0], 2.33, 0.05)
    import xmimsim_amd.lines as L
    for i0 in (60, 120, 400):  # 3, 6, 20 keV
        E = i0 * 0.05
        x = np.zeros(2048)
        x[i0] = 1.0
        p1 = _prep(det_absorbers=[be])
        t = orc.detector_convolute(p1, x).sum()
        assert t == pytest.approx(math.exp(-L.cs_total(4, E) * 1.85 * 0.002), rel=1e-9)
        p2 = _prep(det_absorbers=[be], crystal=[si])
        t2 = orc.detector_convolute(p2, x).sum()
        assert t2 == pytest.approx(t * (1.0 - math.exp(-L.cs_total(14, E) * 2.33 * 0.05)), rel=1e-9)
        p1.close()
        p2.close()
