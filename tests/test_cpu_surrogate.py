"""CPU tests of the surrogate's reference arithmetic (oracle/surrogate_ref.py)
and of the XRF encoding: the reference backward pass equals central finite
differences of its own loss, and bf16 rounding is round-to-nearest-even; the
first-order line estimate; the array forms of the screened quantification."""
import numpy as np
import pytest


def _tiny(rng, dims=(5, 7, 6, 3)):
    layers = []
    for l in range(len(dims) - 1):
        K = dims[l]
        layers.append((rng.normal(0, 0.6, (dims[l + 1], K)), rng.normal(0, 0.1, dims[l + 1])))
    return layers


def test_reference_gradients_equal_finite_differences(orc):
    import surrogate_ref as R
    rng = np.random.default_rng(3)
    layers = _tiny(rng)
    X = rng.normal(size=(9, 5))
    T = rng.normal(size=(9, 3))
    M = (rng.random((9, 3)) > 0.3).astype(float)
    loss, grads = R.loss_and_grads(layers, X, T, M, use_bf16=False)
    h = 1e-6
    for l, (W, b) in enumerate(layers):
        for (i, j) in [(0, 0), (2, 3), (W.shape[0] - 1, W.shape[1] - 1)]:
            Wp = [(w.copy(), c.copy()) for w, c in layers]
            Wm = [(w.copy(), c.copy()) for w, c in layers]
            Wp[l][0][i, j] += h
            Wm[l][0][i, j] -= h
            fd = (R.loss_and_grads(Wp, X, T, M, False)[0] - R.loss_and_grads(Wm, X, T, M, False)[0]) / (2 * h)
            assert grads[l][0][i, j] == pytest.approx(fd, rel=1e-5, abs=1e-9)
        for i in (0, b.shape[0] - 1):
            Wp = [(w.copy(), c.copy()) for w, c in layers]
            Wm = [(w.copy(), c.copy()) for w, c in layers]
            Wp[l][1][i] += h
            Wm[l][1][i] -= h
            fd = (R.loss_and_grads(Wp, X, T, M, False)[0] - R.loss_and_grads(Wm, X, T, M, False)[0]) / (2 * h)
            assert grads[l][1][i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_bf16_rounding_is_nearest_even(orc):
    import surrogate_ref as R
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1)
    assert R.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0
    assert R.bf16(np.float32([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert R.bf16(np.float32([1.0 + 2 ** -8 + 2 ** -20]))[0] == 1.0 + 2 ** -7
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(R.bf16(x) - x) / np.abs(x)) <= 2 ** -8  # half an ulp of an 8-bit significand


def test_xrf_encoding():
    from xmimsim_amd import surrogate as S
    assert S.N_IN == 85 and S.N_OUT == 256 and len(set(S.OUT_LINES)) == 256
    x = S.encode([26, 82], [0.7, 0.3], 0.01, 7.9, 20.0)
    assert x[26 - 11] == pytest.approx(1 + np.log10(0.7) / 4) and x[82 - 11] == pytest.approx(1 + np.log10(0.3) / 4)
    assert x[12 - 11] == 0.0
    assert x[-3] == np.float32(-2.0) and x[-1] == np.float32(0.5)
    vr = np.zeros((100, 385, 4))
    vr[25, 2, :] = [1e3, 10.0, 1.0, 0.0]   # Fe KL3
    t, m = S.targets(vr, [26, 82])
    k = S.OUT_LINES.index((26, 3))
    assert m[k] == 1.0 and t[k] == pytest.approx(np.log10(1011.0))
    assert m.sum() == 1.0


def test_reference_loop_step_limits():
    """bin/xmimsim-pymca.c:698-725 step limits; src/xmi_pymca.c:1684-1751 rescaling"""
    from xmimsim_amd import quant
    assert [quant._max_scale(w) for w in (1e-5, 5e-3, 0.05, 0.2, 0.3, 0.45, 0.55, 0.65, 0.9)] == \
        [100.0, 10.0, 2.5, 1.5, 1.2, 1.1, 1.05, 1.025, 1.01]
    # weights above 1 in total: the largest and those before it absorb the excess
    w = quant._pymca_composition(np.array([0.2, 0.7, 0.3]))
    assert w.sum() == pytest.approx(1.0) and w[2] == 0.3


def test_line_baseline_limits(lib):
    """The first-order estimate from the encoded inputs against the same
    physics evaluated directly at the sample's energy: thin layers (x << 1)
    go as w P_s omega R rho d, thick ones as w P_s omega R / chi, an absent
    element and an unexcitable line give 0."""
    import math
    from xmimsim_amd import lines as L, surrogate as S
    B = S.LineBaseline()
    comp = ([26, 29], [0.6, 0.4])
    E0 = 17.4  # a grid node: no interpolation error
    j_fe = S.OUT_LINES.index((26, S.KL3))
    j_cu = S.OUT_LINES.index((29, S.KL3))
    j_ni = S.OUT_LINES.index((28, S.KL3))

    def direct(j, w, d, rho):
        z, line = S.OUT_LINES[j]
        e_line = L.line_energy(z, line)
        mu0 = sum(wi * L.cs_total(zi, E0) for zi, wi in zip(*comp))
        mul = sum(wi * L.cs_total(zi, e_line) for zi, wi in zip(*comp))
        chi = math.sqrt(2.0) * (mu0 + mul)
        P = L.shell_cross_sections(z, E0, 4)[0]
        wr = lib.xmi_provider_line_data(z, 0, 0) * lib.xmi_provider_line_data(z, 1, line)
        return math.log10(w * P * wr * -math.expm1(-chi * rho * d) / chi), math.log10(w * P * wr * rho * d), \
            math.log10(w * P * wr / chi)

    for d, limit in ((1e-7, 1), (1.0, 2)):
        X = S.encode(comp[0], comp[1], d, 8.0, E0)[None]
        b = B(X)[0]
        for j, w in ((j_fe, 0.6), (j_cu, 0.4)):
            exact, thin, thick = direct(j, w, d, 8.0)
            assert b[j] == pytest.approx(exact, abs=2e-5)  # weights pass through float32 inputs
            assert b[j] == pytest.approx((thin, thick)[limit - 1], abs=1e-3)
        assert b[j_ni] == 0.0
    # Ag at 20 keV: below its K edge (25.5 keV) no K line, its L lines excited
    X = S.encode([47], [1.0], 1e-3, 10.5, 20.0)[None]
    b = B(X)[0]
    assert b[S.OUT_LINES.index((47, S.KL3))] == 0.0 and b[S.OUT_LINES.index((47, S.L3M5))] != 0.0


def test_screening_arrays_equal_the_per_sample_forms():
    """The screened quantification's array forms: grouped candidates and
    encodings equal surrogate.encode of each composition bit for bit, the
    grouped Gauss-Newton step equals the per-sample step, and the Jacobians
    of a linear stand-in network are its own coefficients."""
    from xmimsim_amd import quant as Q, surrogate as S
    rng = np.random.default_rng(4)
    samples = [Q.random_sample(rng, (2, 12)) for _ in range(60)]
    W = [rng.dirichlet(np.ones(len(s.Z))) for s in samples]
    groups = Q._Groups(samples)
    assert sorted(i for idx, _, _ in groups.groups for i in idx) == list(range(60))
    h = 0.2
    for idx, Z, geo in groups.groups:
        C = Q._candidates(np.array([W[i] for i in idx]), h)
        X = Q._encode_group(Z, C, geo)
        for k, i in enumerate(idx):
            s = samples[i]
            assert np.array_equal(X[k, 0], S.encode(s.Z, W[i], s.thickness, s.density, s.energy))
            for j in range(len(s.Z)):
                wp = W[i].copy()
                wp[j] *= np.exp(h)
                assert np.array_equal(X[k, 1 + j], S.encode(s.Z, wp / wp.sum(), s.thickness, s.density, s.energy))
    # Gauss-Newton: grouped against one sample at a time, over masked lines
    for n in (3, 7):
        g = 5
        Wg = rng.dirichlet(np.ones(n), size=g)
        J = rng.normal(size=(g, S.N_OUT, n))
        r = rng.normal(size=(g, S.N_OUT))
        m = rng.random((g, S.N_OUT)) < 0.1
        m[0] = False  # a sample without lines keeps its weights
        out = Q._gn_update_many(Wg, J, np.where(m, r, 0.0), m, 1e-3)
        for k in range(g):
            ref = Q._gn_update(Wg[k], J[k][m[k]], r[k][m[k]], 1e-3)
            assert np.allclose(out[k], ref, rtol=1e-10, atol=0)
        assert np.allclose(out[0], Wg[0], rtol=1e-12)

    # a linear network in the encoded inputs: d log10 I / d log w_j = A[z_j] / (4 ln 10) up to the renormalisation
    A = rng.normal(size=(S.N_IN, S.N_OUT))

    class Linear:
        n_nets = 2

        def predict(self, X):
            y = X.astype(np.float64) @ A
            return np.stack([y, y])

    jac = Q._surrogate_jacobians(Linear(), samples, W, h=h)
    for i in (0, 17, 59):
        s = samples[i]
        x0 = S.encode(s.Z, W[i], s.thickness, s.density, s.energy).astype(np.float64)
        assert np.allclose(jac[i][0], x0 @ A, rtol=1e-12, atol=1e-12) and jac[i][2].shape == (2, S.N_OUT)
        for j in range(len(s.Z)):
            wp = W[i].copy()
            wp[j] *= np.exp(h)
            x1 = S.encode(s.Z, wp / wp.sum(), s.thickness, s.density, s.energy).astype(np.float64)
            assert np.allclose(jac[i][1][:, j], (x1 - x0) @ A / h, rtol=1e-9, atol=1e-9)


def test_line_baseline_sparse_form_equals_dense(lib):
    """The estimate is evaluated only at a row's own elements and lines; a
    dense evaluation of the same expression over every (row, line) gives the
    same values, and zeros elsewhere."""
    from xmimsim_amd import quant as Q, surrogate as S
    B = S.LineBaseline()
    rng = np.random.default_rng(8)
    samples = [Q.random_sample(rng, (2, 30)) for _ in range(200)]
    X = np.stack([S.encode(s.Z, s.w, s.thickness, s.density, s.energy) for s in samples]).astype(np.float64)
    xw = X[:, :S.Z_MAX - S.Z_MIN + 1]
    w = np.where(xw > 0.0, 10.0 ** (4.0 * (xw - 1.0)), 0.0)
    d, rho, E0 = 10.0 ** X[:, -3], 20.0 * X[:, -2], 40.0 * X[:, -1]
    t = np.clip((E0 - B.E_LO) / B.DE, 0.0, len(B.E) - 1.000001)
    i0 = t.astype(np.int64)
    f = t - i0
    mu0 = np.einsum("nz,zn->n", w, B.sig[:, i0] * (1.0 - f) + B.sig[:, i0 + 1] * f)
    chi = np.sqrt(2.0) * (mu0[:, None] + w @ B.sig_line)
    P = (B.ps[:, i0] * (1.0 - f) + B.ps[:, i0 + 1] * f).T
    with np.errstate(divide="ignore", invalid="ignore"):
        A = np.where(chi > 0.0, -np.expm1(-chi * (rho * d)[:, None]) / chi, 0.0)
        v = w[:, S._OUT_Z - S.Z_MIN] * P * B.wr[None] * A
        dense = np.where((v > 0.0) & B.ok[None], np.log10(v), 0.0)
    got = B(X)
    assert np.allclose(got, dense.astype(np.float32), rtol=1e-6, atol=0)
    assert np.array_equal(got != 0.0, dense != 0.0) and (got != 0.0).sum() > 1000
