"""Training data, training and inverse quantification for the net-line
surrogate (xmimsim_amd.surrogate), all driven by the GPU Monte Carlo engine.

Samples are single homogeneous layers excited by one monochromatic line, in
the reference's default 45/45 degree geometry.  The Monte Carlo intensities
are the var_red_history rows of xmi_main_msim summed over interaction orders.

Quantification solves for the weight fractions of known elements from measured
net-line intensities by damped Gauss-Newton on log10 intensities:
  - screened: Gauss-Newton on the surrogate ensemble (finite-difference
    Jacobians, one batched forward pass per iteration for all samples), then
    one Monte Carlo run at the answer that confirms it; samples whose ensemble
    spread exceeds the gate, or whose confirmation misses the tolerance, fall
    back to the Monte Carlo loop;
  - plain: the same Gauss-Newton loop with every function value and Jacobian
    column from a Monte Carlo run (common random numbers across columns).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import api, surrogate as S


@dataclass
class Sample:
    Z: List[int]
    w: np.ndarray
    thickness: float
    density: float
    energy: float


def random_sample(rng: np.random.Generator, n_el=(3, 12), alpha=0.5) -> Sample:
    """Single-layer sample: n_el elements of Z 11..92 with Dirichlet(alpha)
    weight fractions, thickness log-uniform 1 um - 1 cm, density 1-20 g/cm3,
    excitation 10-40 keV."""
    n = int(rng.integers(n_el[0], n_el[1] + 1))
    Z = sorted(rng.choice(np.arange(S.Z_MIN, S.Z_MAX + 1), size=n, replace=False).tolist())
    w = rng.dirichlet(np.full(n, alpha))
    w = np.maximum(w, 1e-3)
    w /= w.sum()
    d = float(np.exp(rng.uniform(np.log(1e-4), np.log(1.0))))
    return Sample(Z, w, d, float(rng.uniform(1.0, 20.0)), float(rng.uniform(10.0, 40.0)))


def make_input(s: Sample, w: Optional[np.ndarray] = None, photons: int = 1_000_000) -> api.Input:
    w = s.w if w is None else w
    return api.Input(layers=[api.Layer(list(s.Z), [float(x) for x in w], s.density, s.thickness)],
                     discrete=[api.DiscreteLine(s.energy, 1e9, 1e9)], n_photons_line=photons,
                     n_interactions_trajectory=4)


class Engine:
    """Monte Carlo evaluations of net lines with shared tables for Z 11..92."""

    def __init__(self, grid_hits: int = 500, Zs: Sequence[int] = range(S.Z_MIN, S.Z_MAX + 1)):
        self.tables = api.SharedTables(Zs)
        self.grid_hits = grid_hits
        self.mc_runs = 0
        self.mc_histories = 0

    def close(self):
        self.tables.close()

    def lines(self, s: Sample, w: np.ndarray, photons: int, seed: int, grid=None):
        """log10 net lines [N_OUT] and mask at composition w; grid: a prepared
        input whose solid-angle grid is reused (same geometry and thickness)."""
        prep = api.Prepared(make_input(s, w, photons), tables=self.tables)
        try:
            if grid is None:
                prep.compute_solid_angle(hits_per_single=self.grid_hits, seed=seed)
            else:
                prep.solid_angle = grid.solid_angle
                prep._sa_owner = "borrowed"
            _, vr, _, st = prep.run(seed=seed)
            self.mc_runs += 1
            self.mc_histories += st.histories
            t, m = S.targets(vr, s.Z)
            return t, m, prep if grid is None else None
        finally:
            if grid is not None:
                prep.solid_angle = None
                prep.close()


def generate(engine: Engine, n: int, photons: int, seed: int, n_el=(2, 30), alpha=0.5):
    """Training set: X [n][85], T [n][256], M [n][256]."""
    rng = np.random.default_rng(seed)
    X = np.zeros((n, S.N_IN), np.float32)
    T = np.zeros((n, S.N_OUT), np.float32)
    M = np.zeros((n, S.N_OUT), np.float32)
    for k in range(n):
        s = random_sample(rng, n_el, alpha)
        t, m, prep = engine.lines(s, s.w, photons, seed=seed * 100003 + k)
        prep.close()
        X[k] = S.encode(s.Z, s.w, s.thickness, s.density, s.energy)
        T[k], M[k] = t, m
    return X, T, M


def train(sg: S.Surrogate, X, T, M, steps: int, batch: int, lr: float = 1e-3, seed: int = 0, log=None):
    """Adam on random mini-batches; returns the per-step mean loss."""
    rng = np.random.default_rng(seed)
    hist = []
    n = X.shape[0]
    for k in range(steps):
        idx = rng.integers(0, n, batch)
        loss = sg.train_step(X[idx], T[idx], M[idx], lr)
        hist.append(float(loss.mean()))
        if log and (k % max(1, steps // 10) == 0 or k == steps - 1):
            log(f"step {k}: loss {hist[-1]:.5f}")
    return np.array(hist)


def _encode_batch(s: Sample, W: np.ndarray) -> np.ndarray:
    return np.stack([S.encode(s.Z, w, s.thickness, s.density, s.energy) for w in W])


def _lines_mask(s: Sample) -> np.ndarray:
    return np.array([z in set(s.Z) for z, _ in S.OUT_LINES])


def _gn_update(w, J, r, lam):
    """Damped Gauss-Newton step on log weights, renormalised."""
    A = J.T @ J + lam * np.eye(J.shape[1])
    step = np.linalg.solve(A, -J.T @ r)
    w_new = w * np.exp(np.clip(step, -1.0, 1.0))
    return w_new / w_new.sum()


def quantify_screened(sg: S.Surrogate, engine: Engine, samples: List[Sample], meas: np.ndarray, meas_mask: np.ndarray,
                      photons: int, iters: int = 12, gate: float = 0.02, tol: float = 0.02, seed: int = 0,
                      grids=None):
    """Surrogate Gauss-Newton for all samples at once, then one MC run per
    sample to confirm; fallback to plain MC for gated or unconfirmed ones.
    Returns (w_hat list, info dict)."""
    n = len(samples)
    W = [np.full(len(s.Z), 1.0 / len(s.Z)) for s in samples]
    h = 1e-3
    for it in range(iters):
        rows, where = [], []
        for i, s in enumerate(samples):
            base = W[i]
            cand = [base]
            for j in range(len(s.Z)):
                wp = base.copy()
                wp[j] *= np.exp(h)
                cand.append(wp / wp.sum())
            where.append((len(rows), len(cand)))
            rows.extend(_encode_batch(s, np.array(cand)))
        Y = sg.predict(np.array(rows)).mean(axis=0)
        for i, s in enumerate(samples):
            o, c = where[i]
            mk = meas_mask[i] > 0
            f0 = Y[o][mk]
            J = ((Y[o + 1:o + c][:, mk] - f0) / h).T
            r = f0 - meas[i][mk]
            W[i] = _gn_update(W[i], J, r, 1e-4)
    # ensemble spread at the answer (relative, in linear intensity)
    rows = np.array([S.encode(s.Z, W[i], s.thickness, s.density, s.energy) for i, s in enumerate(samples)])
    Ye = sg.predict(rows)
    spread = np.zeros(n)
    for i in range(n):
        mk = meas_mask[i] > 0
        lin = 10.0 ** Ye[:, i, mk]
        spread[i] = float(np.max(lin.std(axis=0) / lin.mean(axis=0))) if mk.any() else 0.0
    gated = spread > gate
    confirmed = np.zeros(n, bool)
    line_err = np.zeros(n)
    fallback = 0
    for i, s in enumerate(samples):
        mk = meas_mask[i] > 0
        if not gated[i]:
            t, m, _ = engine.lines(s, W[i], photons, seed=seed + 7 * i, grid=grids[i] if grids else None)
            line_err[i] = float(np.max(np.abs(10.0 ** (t[mk] - meas[i][mk]) - 1.0))) if mk.any() else 0.0
            confirmed[i] = line_err[i] <= tol
        if gated[i] or not confirmed[i]:
            fallback += 1
            W[i], line_err[i] = _mc_gn(engine, s, meas[i], meas_mask[i], photons, seed + 7 * i, W[i], tol,
                                       grids[i] if grids else None)
    return W, {"spread": spread, "gated": int(gated.sum()), "fallback": fallback, "line_err": line_err}


def _mc_gn(engine: Engine, s: Sample, meas, meas_mask, photons, seed, w0, tol, grid, iters: int = 6):
    """Gauss-Newton with Monte Carlo values and finite-difference Jacobians
    (the same seed for all columns of one iteration)."""
    w = w0.copy()
    mk = meas_mask > 0
    h = 0.05
    err = np.inf
    for it in range(iters):
        t0, _, _ = engine.lines(s, w, photons, seed=seed + it, grid=grid)
        err = float(np.max(np.abs(10.0 ** (t0[mk] - meas[mk]) - 1.0))) if mk.any() else 0.0
        if err <= tol:
            break
        cols = []
        for j in range(len(s.Z)):
            wp = w.copy()
            wp[j] *= np.exp(h)
            tj, _, _ = engine.lines(s, wp / wp.sum(), photons, seed=seed + it, grid=grid)
            cols.append((tj[mk] - t0[mk]) / h)
        J = np.array(cols).T
        w = _gn_update(w, J, t0[mk] - meas[mk], 1e-3)
    return w, err


def quantify_plain(engine: Engine, samples: List[Sample], meas, meas_mask, photons, tol=0.02, seed=0, grids=None):
    W, errs = [], []
    for i, s in enumerate(samples):
        w, e = _mc_gn(engine, s, meas[i], meas_mask[i], photons, seed + 7 * i, np.full(len(s.Z), 1.0 / len(s.Z)),
                      tol, grids[i] if grids else None)
        W.append(w)
        errs.append(e)
    return W, {"line_err": np.array(errs)}
