"""The warm-start MLP kernels (BF16 MFMA forward, backward, Adam) against the CPU oracle.

Bar: forward outputs and loss/gradients equal the oracle's up to f32 accumulation
order -- an activation whose f32 pre-image lies on a bf16 rounding boundary may
round the other way, so elementwise checks allow a few outliers and norm-wise
checks are tight; Adam (pure f32 elementwise) to 1e-6 relative.
"""
import numpy as np
import pytest

import mlp_oracle as mo
from xpbi.mlp import Mlp, init_params

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = init_params(seed).astype(np.float64)
    d = mo.split(p)
    d["b1"] = rng.normal(0, 0.3, 64)
    d["b2"] = rng.normal(0, 0.3, 64)
    d["b3"] = rng.normal(0, 0.3, 2)
    return mo.join(d).astype(np.float32)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1.5, size=(n, mo.IN)).astype(np.float32)
    T = rng.normal(0, 1.0, size=(n, mo.OUT)).astype(np.float32)
    T[rng.random(n) < 0.3, 1] = np.nan
    return X, T


@pytest.mark.parametrize("n", [1, 31, 1000, 4133])
def test_predict_matches_oracle(n):
    p = _params(1)
    X, _ = _data(2, n)
    net = Mlp(0, p)
    y = net.predict(X)
    yo, _ = mo.forward(p, X)
    err = np.abs(y - yo)
    scale = max(float(np.abs(yo).max()), 1e-3)
    print(f"n={n}: max |y - oracle| {err.max():.3e} (scale {scale:.3e}), "
          f"rows > 1e-5 rel: {(err.max(1) > 1e-5 * scale).sum()}", flush=True)
    assert err.max() <= 2e-2 * scale
    assert (err.max(1) > 1e-5 * scale).mean() <= 0.02
    net.close()


@pytest.mark.parametrize("n,weighted", [(100, False), (4096, False), (20000, False), (4096, True)])
def test_loss_grad_matches_oracle(n, weighted):
    p = _params(3)
    X, T = _data(4, n)
    W = np.random.default_rng(5).exponential(1.0, n).astype(np.float32) if weighted else None
    net = Mlp(0, p)
    net.set_data(X, T, W)
    loss, g = net.loss_grad()
    lo, go = mo.loss_grad(p, X, T, W=W)
    blocks = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
    print(f"n={n}: loss {loss:.6e} oracle {lo:.6e}", flush=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for (name, _), a, b in zip(mo.SHAPES, blocks[:-1], blocks[1:]):
        e = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        print(f"  {name}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-2, name
    # a sub-range of the data
    loss2, g2 = net.loss_grad(first=n // 3, count=n // 2)
    sl = slice(n // 3, n // 3 + n // 2)
    lo2, go2 = mo.loss_grad(p, X[sl], T[sl], W=None if W is None else W[sl])
    assert abs(loss2 - lo2) <= 1e-4 * lo2
    assert np.linalg.norm(g2 - go2) <= 1e-2 * np.linalg.norm(go2)
    net.close()


def test_adam_step_matches_oracle():
    p = _params(5)
    net = Mlp(0, p)
    rng = np.random.default_rng(6)
    q, m, v = p.astype(np.float64), np.zeros(p.size), np.zeros(p.size)
    for step in range(1, 4):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        net.adam_step(1e-3, g)
        q, m, v = mo.adam(q, m, v, g, step, 1e-3)
    np.testing.assert_allclose(net.params(), q, rtol=1e-5, atol=1e-7)
    net.close()


def test_training_fits_a_smooth_map():
    """GPU training (minibatch Adam through the library's kernels) fits y = f(x)."""
    rng = np.random.default_rng(7)
    n = 16384
    X = rng.uniform(-1, 1, size=(n, mo.IN)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2], 0.5 * np.tanh(3 * X[:, 3]) - X[:, 4] ** 2], 1)
    T = T.astype(np.float32)
    net = Mlp(0, init_params(1))
    net.set_data(X, T)
    l0, _ = net.loss_grad()
    losses = net.train(epochs=30, batch=1024, lr=3e-3, seed=2)
    l1, _ = net.loss_grad()
    print(f"loss {l0:.4f} -> {l1:.5f}; epochs {np.array2string(losses[::6], precision=4)}", flush=True)
    assert l1 < 0.05 * l0
    # the trained weights evaluate the same on the GPU and in the oracle
    p = net.params()
    yo, _ = mo.forward(p, X[:2000])
    y = net.predict(X[:2000])
    assert np.abs(y - yo).max() <= 2e-2 * np.abs(yo).max()
    net.close()
