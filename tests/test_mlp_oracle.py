"""The MLP restatement of the learned correction (oracle/mlp_oracle.py),
pinned by what the mathematics fixes: its gradient equals central
differences of its loss and torch.autograd's gradient in FP64 (no bf16
rounding), and the bf16 rounding is round-to-nearest-even."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def params(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in M.SHAPES:
        fan = shape[0] if len(shape) == 2 else 1
        d[name] = rng.standard_normal(shape) * (np.sqrt(2.0 / fan) if len(shape) == 2 else 0.1)
    return M.flatten(d)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2.0 ** -6  # tie -> even (upwards)
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = params(1)
    x = rng.standard_normal((5, M.NI))
    t = rng.standard_normal((5, M.NO))
    _, g = M.loss_and_grad(p, x, t, emulate=False)
    idx = rng.choice(M.NPARAMS, 40, replace=False)
    idx = np.concatenate([idx, [M.NPARAMS - 1, M.NI * M.NH]])  # a b3 and a b1 entry
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t, False)[0] - M.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    p = params(2)
    x = rng.standard_normal((7, M.NI))
    t = rng.standard_normal((7, M.NO))
    loss, g = M.loss_and_grad(p, x, t, emulate=False)
    P = {k: torch.tensor(v, requires_grad=True) for k, v in M.unflatten(p).items()}
    X, T = torch.tensor(x), torch.tensor(t)
    gel = torch.nn.functional.gelu
    Y = gel(gel(X @ P["W1"] + P["b1"]) @ P["W2"] + P["b2"]) @ P["W3"] + P["b3"]
    L = ((Y - T) ** 2).sum() / (2 * X.shape[0])
    L.backward()
    assert abs(L.item() - loss) <= 1e-12 * abs(loss)
    gt = M.flatten({k: v.grad.numpy() for k, v in P.items()})
    assert np.max(np.abs(gt - g)) <= 1e-12 * np.max(np.abs(g))


def test_planted_error_is_caught():
    """dropping GELU's density term from the derivative fails the check"""
    rng = np.random.default_rng(5)
    p = params(3)
    x = rng.standard_normal((4, M.NI))
    t = rng.standard_normal((4, M.NO))
    good = M.gelu_grad
    try:
        M.gelu_grad = lambda z: 0.5 * (1.0 + M.erf(z / np.sqrt(2.0)))
        _, g_bad = M.loss_and_grad(p, x, t, emulate=False)
    finally:
        M.gelu_grad = good
    i = 17  # a W1 entry
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (M.loss_and_grad(pp, x, t, False)[0] - M.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
    _, g = M.loss_and_grad(p, x, t, emulate=False)
    assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))
    assert abs(fd - g_bad[i]) > 1e-4 * max(1.0, abs(g[i]))
