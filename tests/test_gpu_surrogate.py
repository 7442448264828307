"""GPU tests of the net-line surrogate kernels against the CPU reference
(oracle/surrogate_ref.py): matrix-core forward pass, loss and parameter
gradients, and that Adam training reduces the loss."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (85, 128, 128, 64)


def _net(nets=2, seed=1):
    from xmimsim_amd.surrogate import Surrogate
    return Surrogate(n_nets=nets, dims=DIMS, max_batch=256, seed=seed)


def _data(rng, B=128):
    X = rng.normal(0, 1, (B, DIMS[0])).astype(np.float32)
    T = rng.normal(0, 1, (B, DIMS[-1])).astype(np.float32)
    M = (rng.random((B, DIMS[-1])) > 0.2).astype(np.float32)
    return X, T, M


def test_forward_matches_reference():
    import surrogate_ref as R
    rng = np.random.default_rng(7)
    s = _net()
    X, _, _ = _data(rng)
    Y = s.predict(X)
    P = s.get_params()
    for e in range(s.n_nets):
        layers = s.unpack(P[e])
        ref, _ = R.forward(layers, X, use_bf16=True)
        exact, _ = R.forward(layers, X, use_bf16=False)
        scale = np.abs(ref).max()
        # same bf16 operands, f32 vs f64 accumulation
        assert np.abs(Y[e] - ref).max() <= 2e-3 * scale, np.abs(Y[e] - ref).max() / scale
        # against exact arithmetic: the bf16 rounding of operands only
        assert np.abs(Y[e] - exact).max() <= 5e-2 * scale
    s.close()


def test_gradients_match_reference():
    import surrogate_ref as R
    rng = np.random.default_rng(11)
    s = _net()
    X, T, M = _data(rng)
    loss, g = s.gradients(X, T, M)
    P = s.get_params()
    for e in range(s.n_nets):
        layers = s.unpack(P[e])
        rl, rg = R.loss_and_grads(layers, X, T, M, use_bf16=True)
        el, eg = R.loss_and_grads(layers, X, T, M, use_bf16=False)
        assert loss[e] == pytest.approx(rl, rel=1e-3)
        flat_ref = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in rg])
        flat_ex = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in eg])
        err = np.linalg.norm(g[e] - flat_ref) / np.linalg.norm(flat_ref)
        assert err < 1e-2, err
        assert np.linalg.norm(g[e] - flat_ex) / np.linalg.norm(flat_ex) < 5e-2
    s.close()


def test_adam_training_reduces_loss():
    rng = np.random.default_rng(5)
    s = _net(nets=2, seed=3)
    X = rng.normal(0, 1, (256, DIMS[0])).astype(np.float32)
    A = rng.normal(0, 0.2, (DIMS[0], DIMS[-1]))
    T = np.tanh(X @ A).astype(np.float32)
    M = np.ones_like(T)
    first = s.train_step(X, T, M, 1e-3)
    for _ in range(300):
        last = s.train_step(X, T, M, 1e-3)
    assert np.all(last < 0.1 * first), (first, last)
    s.close()


def test_batch_constraints_are_errors():
    s = _net()
    X = np.zeros((100, DIMS[0]), np.float32)
    with pytest.raises(RuntimeError):
        s.gradients(X, np.zeros((100, DIMS[-1]), np.float32), np.ones((100, DIMS[-1]), np.float32))
    s.close()


def test_gradients_at_training_scale():
    """The training configuration of the net-line surrogate: 85 -> 4 x 512
    (GELU) -> 256, an ensemble of 4, batch 4096.  Loss and every parameter
    gradient of each member against the CPU reference with the same bf16
    operand rounding (relative 1e-2 in norm), member 0 also against exact
    arithmetic (5e-2)."""
    import surrogate_ref as R
    from xmimsim_amd.surrogate import DEFAULT_DIMS, Surrogate
    B = 4096
    rng = np.random.default_rng(17)
    s = Surrogate(n_nets=4, dims=DEFAULT_DIMS, max_batch=B, seed=5)
    X = rng.normal(0, 1, (B, DEFAULT_DIMS[0])).astype(np.float32)
    T = rng.normal(0, 1, (B, DEFAULT_DIMS[-1])).astype(np.float32)
    M = (rng.random((B, DEFAULT_DIMS[-1])) > 0.3).astype(np.float32)
    loss, g = s.gradients(X, T, M)
    P = s.get_params()
    for e in range(4):
        print("member", e, flush=True)
        layers = s.unpack(P[e])
        rl, rg = R.loss_and_grads(layers, X, T, M, use_bf16=True)
        assert loss[e] == pytest.approx(rl, rel=1e-3)
        flat_ref = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in rg])
        assert np.linalg.norm(g[e] - flat_ref) / np.linalg.norm(flat_ref) < 1e-2
        if e == 0:
            _, eg = R.loss_and_grads(layers, X, T, M, use_bf16=False)
            flat_ex = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in eg])
            assert np.linalg.norm(g[e] - flat_ex) / np.linalg.norm(flat_ex) < 5e-2
    s.close()


def test_resident_training_equals_step_by_step():
    """Adam steps on device-resident data (rows chosen by index on the GPU,
    no host round trip per step) give the same parameters and losses, bit
    for bit, as train_step on the same rows."""
    rng = np.random.default_rng(23)
    n, B, steps = 1000, 128, 6
    X = rng.normal(0, 1, (n, DIMS[0])).astype(np.float32)
    T = rng.normal(0, 1, (n, DIMS[-1])).astype(np.float32)
    M = (rng.random((n, DIMS[-1])) > 0.25).astype(np.float32)
    idx = rng.integers(0, n, (steps, B)).astype(np.int32)
    lr = np.linspace(2e-3, 5e-4, steps).astype(np.float32)
    a, b = _net(seed=9), _net(seed=9)
    la = np.array([a.train_step(X[i], T[i], M[i], float(r)) for i, r in zip(idx, lr)])
    b.set_data(X, T, M)
    lb = b.train_resident(idx, lr)
    assert np.array_equal(la, lb)
    assert np.array_equal(a.get_params(), b.get_params())
    with pytest.raises(RuntimeError):
        b.train_resident(np.full((1, B), n, np.int32), lr[:1])
    a.close()
    b.close()
