"""GPU test of the surrogate-screened quantification: the ensemble trained
on Monte Carlo spectra of the engine (small: Z 20-30, 2048 samples) makes the
inverse loop need fewer Monte Carlo runs than the plain Monte Carlo loop, and
both reach the 2 % net-line tolerance."""
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_screening_saves_monte_carlo_runs():
    from xmimsim_amd import quant
    r = quant.run_pipeline(n_train=2048, photons_train=50_000, steps=1500, n_quant=6, photons_quant=300_000,
                           zrange=(20, 30), n_el_train=(2, 8), n_el_quant=(3, 5), batch=1024, seed=3)
    print({k: r[k] for k in ("loss_first", "loss_last", "holdout_err_median", "screened", "plain")}, flush=True)
    assert r["loss_last"] < 0.2 * r["loss_first"]
    # the 2 % criterion is met about as often both ways (Monte Carlo noise at
    # 3e5 photons decides the marginal cases), with fewer Monte Carlo runs
    assert r["plain"]["frac_within_2pct"] >= 0.5
    assert r["screened"]["frac_within_2pct"] >= r["plain"]["frac_within_2pct"] - 1 / 6
    assert r["screened"]["mc_runs"] < r["plain"]["mc_runs"]
