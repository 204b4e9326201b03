"""CPU oracle of the learned warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/ (and bench.py's cpu_baseline leg), as the checker of the
library's MLP kernels; the product never imports it.  The reference has no
learned component (it starts every substep at lambda = 0, simkernels.js:146),
so this restates the model the library defines:

    x = bf16(features), h1 = bf16(relu(W1 x + b1)), h2 = bf16(relu(W2 h1 + b2)),
    y = W3 h2 + b3,   weights bf16-rounded from f32 masters,

its loss L = (1/B) sum_n w_n sum_k m_nk (y_nk - t_nk)^2 (w: sample weights, 1
by default; m = 0 where t is NaN), the
gradient with every operand of a weight-gradient product rounded to bf16, and
Adam (Kingma & Ba 2015, beta 0.9 / 0.999, eps 1e-8).  Sums run in float64.
With rounding=False the same functions compute the exact float64 network and
gradient, which the tests pin with finite differences and torch.autograd.
"""
import numpy as np

IN, H, OUT = 12, 64, 2
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 5122


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32; NaN stays NaN."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def split(p):
    p = np.asarray(p, dtype=np.float64).ravel()
    assert p.size == NPARAMS
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[k:k + n].reshape(shape)
        k += n
    return out


def join(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def _r(a, rounding):
    """bf16 rounding of a float64 array through float32 (the kernels round f32 values)."""
    return bf16(np.asarray(a, np.float64).astype(np.float32)).astype(np.float64) if rounding else np.asarray(a)


def forward(p, X, rounding=True):
    """y (n x 2) and the cache of the backward pass."""
    d = split(p)
    X = np.asarray(X, np.float64)
    W1, W2, W3 = (_r(d[k].astype(np.float32), rounding) if rounding else d[k] for k in ("W1", "W2", "W3"))
    b1, b2, b3 = (np.asarray(d[k], np.float32).astype(np.float64) if rounding else d[k] for k in ("b1", "b2", "b3"))
    x = _r(X.astype(np.float32), rounding) if rounding else X
    z1 = x @ W1.T + b1
    if rounding:
        z1 = z1.astype(np.float32).astype(np.float64)
    h1 = _r(np.maximum(z1, 0.0), rounding)
    z2 = h1 @ W2.T + b2
    if rounding:
        z2 = z2.astype(np.float32).astype(np.float64)
    h2 = _r(np.maximum(z2, 0.0), rounding)
    y = h2 @ W3.T + b3
    if rounding:
        y = y.astype(np.float32).astype(np.float64)
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, W2=W2, W3=W3)


def loss_grad(p, X, T, rounding=True, W=None):
    """(loss, gradient in the canonical order) over all rows of X, T; W: sample weights."""
    y, c = forward(p, X, rounding)
    T = np.asarray(T, np.float64)
    m = ~np.isnan(T)
    B = X.shape[0]
    w = np.ones((B, 1)) if W is None else np.asarray(W, np.float32).astype(np.float64).reshape(B, 1)
    e = np.where(m, y - np.where(m, T, 0.0), 0.0)
    loss = float((w * e * e).sum() / B)
    dy = 2.0 / B * w * e
    if rounding:
        dy = dy.astype(np.float32).astype(np.float64)
    dyr = _r(dy, rounding)
    dz2 = (dy @ c["W3"]) * (c["z2"] > 0)
    dz2r = _r(dz2, rounding)
    dz1 = (dz2r @ c["W2"]) * (c["z1"] > 0)
    dz1r = _r(dz1, rounding)
    g = dict(W3=dyr.T @ c["h2"], b3=dyr.sum(0), W2=dz2r.T @ c["h1"], b2=dz2r.sum(0),
             W1=dz1r.T @ c["x"], b1=dz1r.sum(0))
    return loss, join(g)


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new (p, m, v)."""
    p, m, v, g = (np.asarray(a, np.float64) for a in (p, m, v, g))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
