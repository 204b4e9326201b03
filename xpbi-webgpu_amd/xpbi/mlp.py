"""The warm-start network's trainer (xpbi_mlp_* in include/xpbi.h).

The network is 12 -> 64 -> 64 -> 2 with BF16 weights and activations and f32
accumulation; parameters travel as one float32 vector in the canonical order
W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].  Forward, backward and
Adam run on the GPU in the library's own kernels.
"""
import ctypes as C

import numpy as np

from ._abi import MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS, XpbiError, lib

SHAPES = [("W1", (64, MLP_INPUTS)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (MLP_OUTPUTS, 64)),
          ("b3", (MLP_OUTPUTS,))]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed=0):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[1])
            out.append(rng.uniform(-lim, lim, size=shape))
        else:
            out.append(np.zeros(shape))
    return np.concatenate([o.ravel() for o in out]).astype(np.float32)


def split_params(p):
    """Canonical vector -> dict of arrays."""
    p = np.asarray(p)
    d, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        d[name] = p[k:k + n].reshape(shape)
        k += n
    assert k == MLP_PARAMS
    return d


class Mlp:
    def __init__(self, device=0, params=None):
        h = C.c_void_p()
        rc = lib().xpbi_mlp_create(int(device), C.byref(h))
        if rc != 0:
            raise XpbiError(f"xpbi_mlp_create failed ({rc})")
        self._h = h
        self.set_params(init_params() if params is None else params)

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().xpbi_mlp_last_error(self._h)
            raise XpbiError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(lib().xpbi_mlp_set_params(self._h, _f(p), p.size), "set_params")

    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self._check(lib().xpbi_mlp_get_params(self._h, _f(p), p.size), "get_params")
        return p

    def set_data(self, x, t, w=None):
        """Training rows: features x (n x 12), targets t (n x 2, NaN = none), weights w (n, optional)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == MLP_INPUTS and t.shape == (x.shape[0], MLP_OUTPUTS)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float32)
            assert w.shape == (x.shape[0],)
            wp = _f(w)
        self._check(lib().xpbi_mlp_set_data(self._h, _f(x), _f(t), wp, x.shape[0]), "set_data")
        self.rows = x.shape[0]

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], MLP_OUTPUTS), np.float32)
        self._check(lib().xpbi_mlp_predict(self._h, _f(x), x.shape[0], _f(y)), "predict")
        return y

    def loss_grad(self, first=0, count=None):
        count = self.rows - first if count is None else count
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._check(lib().xpbi_mlp_loss_grad(self._h, first, count, _f(g), C.byref(loss)), "loss_grad")
        return float(loss.value), g

    def adam_step(self, lr, grad=None):
        g = None if grad is None else np.ascontiguousarray(grad, dtype=np.float32)
        self._check(lib().xpbi_mlp_adam_step(self._h, None if g is None else _f(g), float(lr)), "adam_step")

    def train(self, epochs, batch=4096, lr=1e-3, seed=0):
        losses = np.zeros(max(epochs, 1), np.float32)
        self._check(lib().xpbi_mlp_train(self._h, int(epochs), int(batch), float(lr), int(seed), _f(losses)),
                    "train")
        return losses[:epochs]

    def close(self):
        if getattr(self, "_h", None):
            lib().xpbi_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
