"""Pins of the MLP oracle (CPU): bf16 rounding, forward, gradient and Adam against
independent implementations (torch's bfloat16 cast, torch.nn.functional,
torch.autograd, torch.optim.Adam) and finite differences."""
import numpy as np
import pytest

import mlp_oracle as mo

torch = pytest.importorskip("torch")


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in mo.SHAPES:
        fan = shape[1] if len(shape) > 1 else 8
        d[name] = rng.normal(0.0, scale / np.sqrt(fan), size=shape) + (0.1 if name.startswith("b") else 0.0)
    return mo.join(d)


def _data(seed, n=257):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.5, size=(n, mo.IN))
    T = rng.normal(0.0, 1.0, size=(n, mo.OUT))
    T[rng.random(n) < 0.3, 1] = np.nan
    return X, T


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 500), rng.normal(0, 1e30, 500),
                        [0.0, -0.0, np.inf, -np.inf, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1e-45, 3.4e38]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mo.bf16(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(mo.bf16(np.array([np.nan], np.float32)))[0]


def test_forward_matches_torch():
    p = _params(1)
    X, _ = _data(2)
    y, _ = mo.forward(p, X, rounding=True)
    d = {k: torch.from_numpy(v) for k, v in mo.split(p).items()}

    def b(t):  # f32 -> bf16 -> f64
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)

    x = b(torch.from_numpy(X))
    z1 = torch.nn.functional.linear(x, b(d["W1"]), d["b1"].to(torch.float32).to(torch.float64))
    h1 = b(torch.relu(z1))
    z2 = torch.nn.functional.linear(h1, b(d["W2"]), d["b2"].to(torch.float32).to(torch.float64))
    h2 = b(torch.relu(z2))
    yt = torch.nn.functional.linear(h2, b(d["W3"]), d["b3"].to(torch.float32).to(torch.float64)).numpy()
    # the oracle rounds z to f32 before bf16, torch rounds f64 -> bf16 via f32 here too
    np.testing.assert_allclose(y, yt, rtol=1e-6, atol=1e-6)


def test_gradient_matches_finite_differences():
    p = _params(3)
    X, T = _data(4, n=64)
    loss, g = mo.loss_grad(p, X, T, rounding=False)
    rng = np.random.default_rng(5)
    idx = rng.choice(mo.NPARAMS, size=60, replace=False)
    # every parameter block is probed
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
    idx = np.concatenate([idx, offs[:-1], offs[1:] - 1])
    eps = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += eps
        lp, _ = mo.loss_grad(q, X, T, rounding=False)
        q[k] -= 2 * eps
        lm, _ = mo.loss_grad(q, X, T, rounding=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gradient_matches_autograd():
    p = _params(6)
    X, T = _data(7)
    loss, g = mo.loss_grad(p, X, T, rounding=False)
    d = {k: torch.from_numpy(v.copy()).requires_grad_(True) for k, v in mo.split(p).items()}
    x = torch.from_numpy(X)
    h1 = torch.relu(x @ d["W1"].T + d["b1"])
    h2 = torch.relu(h1 @ d["W2"].T + d["b2"])
    y = h2 @ d["W3"].T + d["b3"]
    t = torch.from_numpy(T)
    m = ~torch.isnan(t)
    e = torch.where(m, y - torch.nan_to_num(t), torch.zeros_like(y))
    L = (e * e).sum() / X.shape[0]
    L.backward()
    gt = np.concatenate([d[name].grad.numpy().ravel() for name, _ in mo.SHAPES])
    assert abs(L.item() - loss) <= 1e-12 * max(1.0, loss)
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-12)


def test_weighted_gradient_matches_finite_differences():
    p = _params(12)
    X, T = _data(13, n=48)
    W = np.random.default_rng(14).exponential(1.0, 48)
    loss, g = mo.loss_grad(p, X, T, rounding=False, W=W)
    for k in np.random.default_rng(15).choice(mo.NPARAMS, size=40, replace=False):
        q = p.copy()
        q[k] += 1e-6
        lp, _ = mo.loss_grad(q, X, T, rounding=False, W=W)
        q[k] -= 2e-6
        lm, _ = mo.loss_grad(q, X, T, rounding=False, W=W)
        assert abs((lp - lm) / 2e-6 - g[k]) <= 1e-6 * max(1.0, abs(g[k])), k


def test_rounded_gradient_close_to_exact():
    """bf16 operands perturb the gradient by well under a percent (norm-wise)."""
    p = _params(8)
    X, T = _data(9, n=1024)
    _, g = mo.loss_grad(p, X, T, rounding=False)
    _, gr = mo.loss_grad(p, X, T, rounding=True)
    assert np.linalg.norm(gr - g) <= 2e-2 * np.linalg.norm(g)


def test_adam_matches_torch():
    p = _params(10)
    rng = np.random.default_rng(11)
    w = torch.from_numpy(p.copy()).requires_grad_(True)
    opt = torch.optim.Adam([w], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    q, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    for step in range(1, 5):
        g = rng.normal(0, 1e-2, size=p.shape)
        w.grad = torch.from_numpy(g.copy())
        opt.step()
        q, m, v = mo.adam(q, m, v, g, step, 3e-3)
    np.testing.assert_allclose(q, w.detach().numpy(), rtol=1e-12, atol=1e-15)
