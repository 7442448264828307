"""GPU tests of the detector response (xmi_detector_convolute_all): the
deterministic part against the oracle, escape peaks against the escaped
spectrum built by hand, Poisson noise and pile-up against their statistics."""
import ctypes as C
import dataclasses
import math

import numpy as np
import pytest

from xmimsim_amd import api, workloads

pytestmark = pytest.mark.gpu


def _opts(escape=0, sum_peaks=0, poisson=0):
    o = api.default_options()
    o.use_escape_peaks, o.use_sum_peaks, o.use_poisson, o.use_default_seeds = escape, sum_peaks, poisson, 1
    return o


def test_convolution_matches_oracle_per_order(orc):
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=20000))
    prep.compute_solid_angle(hits_per_single=200, seed=3)
    ch, vr, _, _ = prep.run(seed=5)
    vr_gpu = vr.copy()
    out = api.detector_response(prep, ch, vr_gpu, options=_opts())
    for n in range(1, ch.shape[0]):
        ref = orc.detector_convolute(prep, ch[n])
        scale = np.abs(ref).max()
        assert np.abs(out[n] - ref).max() <= 1e-10 * scale, n
    assert np.all(out[0] == 0.0)
    # histories: detector efficiency at the line energy (Be window x Si crystal)
    import xmimsim_amd.lines as L
    E = L.line_energy(26, 3)
    eff = math.exp(-L.cs_total(4, E) * 1.85 * 0.002) * (1 - math.exp(-L.cs_total(14, E) * 2.33 * 0.5))
    assert vr_gpu[25, 2, 0] == pytest.approx(vr[25, 2, 0] * eff, rel=1e-12)
    prep.close()


def test_escape_peak_matches_hand_built_spectrum(orc):
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=10))
    import xmimsim_amd.lines as L
    nE = 400
    fe = np.linspace(1.0, 40.9, nE)
    fluo = np.zeros((nE, 109, 1))  # C order of the Fortran (element, line, energy)
    fluo[:, 2, 0] = 0.01           # Si KL3 escape: 1 % of every photon above the K edge
    ce = np.array([1.0, 50.0])
    co = np.array([0.1, 0.2])
    comp = np.zeros((2, 2))
    Z = (C.c_int * 1)(14)
    keep = [np.ascontiguousarray(a) for a in (fluo, fe, comp, ce, co)]
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    er = api.xmi_escape_ratios(1, nE, 2, 2, C.cast(Z, C.POINTER(C.c_int)), P(keep[0]), P(keep[1]), P(keep[2]),
                               P(keep[3]), P(keep[4]), None)
    nch = prep.inp.nchannels
    ch = np.zeros((2, nch))
    i, N = 1000, 1e6  # 20 keV
    ch[1, i] = N
    out = api.detector_response(prep, ch, escape_ratios=er, options=_opts(escape=1))
    # by hand: after the absorption, 1 % of channel i moves to channel
    # int((E_i + gain/2 - E(Si KL3) - zero) / gain); the oracle applies the
    # absorption itself, and the response of a unit line sums to the
    # detection efficiency of its channel
    g = prep.inp.gain
    j = int(((i + 0.5) * g - L.line_energy(14, 3)) / g)
    eff = lambda k: orc.detector_convolute(prep, np.eye(1, nch, k)[0]).sum()
    x = np.zeros(nch)
    x[i] = 0.99 * N
    x[j] = 0.01 * N * eff(i) / eff(j)
    ref = orc.detector_convolute(prep, x)
    assert np.abs(out[1] - ref).max() <= 1e-9 * np.abs(ref).max()
    prep.close()


def test_poisson_noise_statistics():
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=10))
    nch = prep.inp.nchannels
    for lam in (5.0, 2000.0):
        ch = np.zeros((2, nch))
        ch[1, :] = lam
        noisy = api.detector_response(prep, ch, options=_opts(poisson=1))[1]
        clean = api.detector_response(prep, ch, options=_opts(poisson=0))[1]
        sel = (clean > 3 * lam / 4) & (np.arange(nch) > 300) & (np.arange(nch) < 1700)
        r = (noisy[sel] - clean[sel]) / np.sqrt(clean[sel])
        assert np.all(noisy[sel] == np.round(noisy[sel]))
        assert abs(r.mean()) < 4.0 / math.sqrt(sel.sum())
        assert r.var() == pytest.approx(1.0, rel=0.15)
    prep.close()


def test_pileup_sum_peak_ratio():
    inp = workloads.single_layer_steel(n_photons_line=10)
    inp.pulse_width = 1e-7
    inp.det_absorbers = []
    inp.crystal = []
    prep = api.Prepared(inp)
    ch = np.zeros((2, inp.nchannels))
    ch[1, 200] = 1e6  # 1e6 counts/s at 4 keV (gain 0.02)
    # pile-up acts before the Gaussian response: compare the areas of the two peaks
    out = api.detector_response(prep, ch, options=_opts(sum_peaks=1))[1]
    main = out[180:221].sum()
    summ = out[381:422].sum()   # (2 * 201 - 1) as the reference places sums
    p = math.exp(-inp.pulse_width * 1e6)
    assert summ / main == pytest.approx((1 - p), rel=0.05)
    prep.close()


def test_pileup_train_limit_is_an_error():
    """Pulse trains longer than 100 pulses stop the reference
    (src/xmi_detector_f.F90:152-156); here the call fails with that reason
    and returns no spectra."""
    inp = workloads.single_layer_steel(n_photons_line=10)
    inp.pulse_width = 1e-4   # 100 mean gaps at 1e6 counts/s: every train overflows
    inp.det_absorbers = []
    inp.crystal = []
    prep = api.Prepared(inp)
    ch = np.zeros((2, inp.nchannels))
    ch[1, 200] = 1e6
    with pytest.raises(RuntimeError, match="pulse train maximum"):
        api.detector_response(prep, ch, options=_opts(sum_peaks=1))
    prep.close()


def test_escape_ratios_match_oracle():
    """xmi_escape_ratios_calculation on the GPU (which builds the crystal's
    escape input itself) against the oracle's restatement run on the escape
    input built as src/xmi_detector.c:98-116 does, with the same streams:
    every fluorescence and Compton ratio to 1e-9; then the ratios feed the
    detector response's escape peaks."""
    import oracle
    inp = api.Input(layers=[api.Layer([26, 28], [0.7, 0.3], 7.9, 0.5)], discrete=[api.DiscreteLine(20.0, 1e9, 0.0)],
                    crystal=[api.Layer([32], [1.0], 5.32, 0.5)], n_photons_line=20000)
    args = dict(n_photons=20000, n_input_energies=40, input_energy_min=5.0, input_energy_delta=0.75,
                n_compton_output_energies=400, compton_output_energy_min=0.1, compton_output_energy_delta=0.1)
    Z, fluo, comp, handle = api.escape_ratios(inp, seed=31, **args)
    try:
        assert Z == [32]
        o = api.default_options(use_M_lines=0, use_cascade_auger=0, use_cascade_radiative=0,
                                use_variance_reduction=0, escape_ratios_mode=1)
        prep = api.Prepared(api.escape_input(inp), options=o)
        f_ref, c_ref = oracle.escape_ratios(prep, args["n_photons"], args["n_input_energies"],
                                            args["input_energy_min"], args["input_energy_delta"],
                                            args["n_compton_output_energies"], args["compton_output_energy_min"],
                                            args["compton_output_energy_delta"], seed=31)
        prep.close()
        assert fluo.shape == f_ref.shape and comp.shape == c_ref.shape
        assert fluo.max() > 0.05 and comp.max() > 0.0
        np.testing.assert_allclose(fluo, f_ref, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(comp, c_ref, rtol=1e-9, atol=1e-15)
        # the computed ratios make escape peaks in the detector response
        prep = api.Prepared(dataclasses.replace(inp, n_photons_line=100000))
        prep.compute_solid_angle(hits_per_single=1000, seed=2)
        ch, vr, _, _ = prep.run(seed=5)
        conv_esc = api.detector_response(prep, ch, vr.copy(), escape_ratios=handle.contents)
        conv = api.detector_response(prep, ch, vr.copy())
        prep.close()
        # Fe K-alpha (6.40 keV) loses to its Ge K-alpha escape at 6.40 - 9.89 < 0: no escape;
        # the 20 keV Rayleigh/Compton peak has its Ge escape near 20 - 9.9 = 10.1 keV
        e = inp.gain * np.arange(inp.nchannels)
        win = (e > 9.6) & (e < 10.6)
        assert conv_esc[-1][win].sum() > conv[-1][win].sum()
    finally:
        api.free_escape_ratios(handle)
