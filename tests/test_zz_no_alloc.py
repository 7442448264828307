"""The allocation-free detector response (xmi_detector_convolute_all_noalloc)
against the allocating entry point xmi_detector_convolute_all: absorption,
escape peaks, pile-up, Gaussian response and Poisson noise on the same seed
give the same spectra and histories, bit for bit, on two calls in a row (the
second smaller); sizes beyond the scratch array are refused."""
import numpy as np
import pytest

from xmimsim_amd import api, workloads

pytestmark = pytest.mark.gpu


def _opts():
    o = api.default_options()
    o.use_escape_peaks, o.use_sum_peaks, o.use_poisson, o.use_default_seeds = 1, 1, 1, 1
    return o


def _spectra(nch, n_int, seed):
    rng = np.random.default_rng(seed)
    ch = np.zeros((n_int + 1, nch))
    for n in range(1, n_int + 1):
        ch[n] = rng.uniform(0.0, 50.0, nch) / n
        for c in rng.integers(50, nch - 50, 6):
            ch[n, c] += rng.uniform(1e3, 1e5)
    ch[1:] = np.cumsum(ch[1:], axis=0)
    vr = np.zeros((100, 385, n_int))
    vr[25, 2, :] = rng.uniform(1e3, 1e4, n_int)
    vr[27, 0, :] = rng.uniform(1e2, 1e3, n_int)
    return ch, vr


def test_noalloc_equals_allocating_entry_point():
    inp = workloads.single_layer_steel(n_photons_line=10)
    _, _, _, er = api.escape_ratios(inp, n_photons=5000, n_input_energies=60, input_energy_min=1.0,
                                    input_energy_delta=0.5, n_compton_output_energies=200,
                                    compton_output_energy_min=0.1, compton_output_energy_delta=0.2, seed=7)
    try:
        for nch, n_int, seed in ((2048, 4, 1), (1000, 2, 2)):
            print("noalloc case", nch, n_int, flush=True)
            inp = workloads.single_layer_steel(n_photons_line=10, n_interactions=n_int)
            inp.nchannels = nch
            inp.pulse_width = 1e-6
            prep = api.Prepared(inp)
            ch, vr = _spectra(nch, n_int, seed)
            vr_a, vr_b = vr.copy(), vr.copy()
            a = api.detector_response(prep, ch, vr_a, escape_ratios=er.contents, options=_opts())
            b = api.detector_response_noalloc(prep, ch, vr_b, escape_ratios=er.contents, options=_opts())
            assert a.shape == b.shape == (n_int + 1, nch)
            assert np.array_equal(a, b), (nch, np.abs(a - b).max())
            assert np.array_equal(vr_a, vr_b)
            assert np.all(b[0] == 0.0) and b[1:].sum() > 0.0
            prep.close()
    finally:
        api.free_escape_ratios(er)


def test_noalloc_refuses_more_channels_than_its_scratch():
    inp = workloads.single_layer_steel(n_photons_line=10, n_interactions=1)
    inp.nchannels = 65537
    prep = api.Prepared(inp)
    with pytest.raises(RuntimeError, match="at most 65536"):
        api.detector_response_noalloc(prep, np.zeros((2, 65537)), options=_opts())
    prep.close()
