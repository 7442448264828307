"""Synthetic inputs of the benchmark workloads (seeded, no external data).

single_layer_steel   one homogeneous layer Fe .70 / Cr .18 / Ni .10 / Mn .02,
                     7.9 g/cm3, 0.5 cm; one 20 keV line, fully polarised,
                     1e9 ph/s; 45/45 geometry; 2048 channels of 0.01 keV;
                     1e6 photons; 4 interactions.
three_layer_kramers  2 cm air / 100 um eight-element layer (Z uniform 20-82,
                     Dirichlet(1) weights, 2 g/cm3) / 1 mm Si; Kramers
                     continuum 1-50 keV in 0.1 keV steps plus W L-alpha and
                     L-beta lines; 1e7 photons per interval; 4 interactions.
batch_samples        independent random samples for the batched engine.
casO4_kat            the reference's CaSO4 known-answer set-up
                     (tests/libxmimsim-test.c:182-264).
"""
from __future__ import annotations

from typing import List

import numpy as np

from .api import ContinuousPoint, DiscreteLine, Input, Layer


def single_layer_steel(n_photons_line: int = 1_000_000, n_interactions: int = 4) -> Input:
    return Input(
        layers=[Layer([26, 24, 28, 25], [0.70, 0.18, 0.10, 0.02], 7.9, 0.5)],
        discrete=[DiscreteLine(20.0, 1e9, 0.0)],
        n_photons_line=n_photons_line,
        n_interactions_trajectory=n_interactions,
        gain=0.01,
        nchannels=2048,
        fano=0.12,
        noise=0.1,
    )


def kramers_spectrum(e0: float = 1.0, e1: float = 50.0, step: float = 0.1, scale: float = 1e9):
    n = int(round((e1 - e0) / step)) + 1
    energies = e0 + step * np.arange(n)
    intens = scale * np.clip(e1 - energies, 0.0, None) / energies
    return [ContinuousPoint(float(e), float(i), 0.0) for e, i in zip(energies, intens)]


def three_layer_kramers(seed: int = 2024, n_photons_interval: int = 10_000_000, n_photons_line: int = 10_000_000,
                        n_interactions: int = 4) -> Input:
    rng = np.random.default_rng(seed)
    Zs = sorted(rng.choice(np.arange(20, 83), size=8, replace=False).tolist())
    w = rng.dirichlet(np.ones(8)).tolist()
    air = Layer([7, 8, 18], [0.755, 0.232, 0.013], 0.0012, 2.0)
    mid = Layer(Zs, w, 2.0, 0.01)
    si = Layer([14], [1.0], 2.33, 0.1)
    return Input(
        layers=[air, mid, si],
        reference_layer=2,
        continuous=kramers_spectrum(),
        discrete=[DiscreteLine(8.40, 5e10, 0.0), DiscreteLine(9.67, 3e10, 0.0)],
        n_photons_interval=n_photons_interval,
        n_photons_line=n_photons_line,
        n_interactions_trajectory=n_interactions,
        gain=0.025,
        nchannels=2048,
    )


def caso4_kat(n_photons_line: int = 1_000_000) -> Input:
    """CaSO4, 1 g/cm3, 1 cm, 10 keV, detector 10 cm away, first order only."""
    # weight fractions of CaSO4 from standard atomic weights
    aw = {20: 40.078, 16: 32.06, 8: 15.999}
    tot = aw[20] + aw[16] + 4 * aw[8]
    return Input(
        layers=[Layer([8, 16, 20], [4 * aw[8] / tot, aw[16] / tot, aw[20] / tot], 1.0, 1.0)],
        discrete=[DiscreteLine(10.0, 1e12, 1e9)],
        n_photons_line=n_photons_line,
        n_interactions_trajectory=1,
        p_detector_window=(0.0, 10.0, 100.0),
    )


def random_sample(rng: np.random.Generator, n_layers=(1, 5), n_elements=(2, 30), zrange=(11, 92), alpha=0.5,
                  density=(1.0, 20.0), thickness=(1e-4, 1.0), energy=(10.0, 40.0)) -> Input:
    nl = int(rng.integers(n_layers[0], n_layers[1] + 1))
    layers = []
    for _ in range(nl):
        ne = int(rng.integers(n_elements[0], n_elements[1] + 1))
        Z = sorted(rng.choice(np.arange(zrange[0], zrange[1] + 1), size=ne, replace=False).tolist())
        w = rng.dirichlet(np.full(ne, alpha)).tolist()
        rho = float(rng.uniform(*density))
        d = float(np.exp(rng.uniform(np.log(thickness[0]), np.log(thickness[1]))))
        layers.append(Layer(Z, w, rho, d))
    E = float(rng.uniform(*energy))
    return Input(layers=layers, discrete=[DiscreteLine(E, 1e9, 0.0)], gain=0.02, nchannels=2048)


def batch_samples(n: int = 1024, seed: int = 7, budget=(1e5, 1e7), n_layers=(1, 5), n_elements=(2, 30),
                  zrange=(11, 92), alpha=0.5, density=(1.0, 20.0), thickness=(1e-4, 1.0),
                  energy=(10.0, 40.0)) -> List[Input]:
    """Independent samples for the batched engine: 1-5 layers of 2-30 elements
    (Z 11-92, Dirichlet(0.5) weight fractions, density uniform 1-20 g/cm3,
    thickness log-uniform 1 um - 1 cm), one monochromatic line uniform in
    10-40 keV, photon budget log-uniform 1e5-1e7, 4 interactions."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        inp = random_sample(rng, n_layers, n_elements, zrange, alpha, density, thickness, energy)
        inp.n_photons_line = int(np.exp(rng.uniform(np.log(budget[0]), np.log(budget[1]))))
        inp.n_interactions_trajectory = 4
        out.append(inp)
    return out
