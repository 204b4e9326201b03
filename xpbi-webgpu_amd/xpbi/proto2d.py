"""2D fp64 engine: the prototype's sand column (prototype/xpbi2d.py) on the GPU.

The prototype is a script whose state is x, v, F (xpbi2d.py:52-58), whose constants
are module globals (xpbi2d.py:29-49) and whose main loop (xpbi2d.py:123-241) advances
the state by substeps of dt.  ``Prototype2D`` holds that state on the device
(``xpbi2d_create``), ``step(k)`` runs k substeps of the loop in one launch, and the
attributes ``x``, ``v``, ``F`` and ``lam`` read it back under the script's names.

``prototype_params()`` / ``prototype_positions()`` are the script's own constants
and initial column (666 particles); ``column40x80_params()`` / ``column40x80_positions()`` the 40 x 80
column at r = 0.01 m (E = 3.5e5 Pa, rho = 2000, phi = 30 deg, dt = 1/3000 s,
20 iterations, Delta-lambda under-relaxed by 0.5).
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from ._abi import FIELD2D, XpbiError, lib


def params(h, E, nu, rho, phi_deg, dt, iterations, gravity, xsph_c, domain, omega=1.0):
    return dict(h=h, E=E, nu=nu, rho=rho, phi_deg=phi_deg, dt=dt, iterations=iterations, gravity=list(gravity),
                xsph_c=xsph_c, domain=list(domain), omega=omega)


def prototype_params():
    """xpbi2d.py:29-49: dx = 0.008, E = 1e4, nu = 0.3, rho = 1000, phi = 35 deg,
    dt = 2e-4, 8 iterations, g = -9.8, XSPH 0.01, box 1.0 x 0.6."""
    return params(0.008, 1e4, 0.3, 1000.0, 35.0, 2e-4, 8, [0.0, -9.8], 0.01, [1.0, 0.6])


def column40x80_params():
    return params(0.01, 3.5e5, 0.3, 2000.0, 30.0, 1.0 / 3000.0, 20, [0.0, -9.81], 0.01, [2.0, 1.0], omega=0.5)


def column(nx, ny, dx, x0, y0, seed):
    """Lattice rows of nx at spacing dx from (x0, y0), jittered by U(-0.05 dx, 0.05 dx)
    from numpy's default_rng(seed) (xpbi2d.py:52-56)."""
    xs, ys = np.meshgrid(np.arange(nx) * dx + x0, np.arange(ny) * dx + y0)
    x = np.stack([xs.ravel(), ys.ravel()], axis=1)
    return x + np.random.default_rng(seed).uniform(-0.05 * dx, 0.05 * dx, x.shape)


def prototype_positions():
    dx = 0.008
    return column(int(0.15 / dx), int(0.30 / dx), dx, 0.425, dx, 0)


def column40x80_positions(seed=0):
    r = 0.01
    return column(40, 80, r, 1.0 - 20 * r + 0.5 * r, r, seed)


def _params_record(p):
    r = _abi.Params2D()
    for k in ("h", "E", "nu", "rho", "phi_deg", "dt", "xsph_c", "omega"):
        setattr(r, k, float(p[k]))
    r.iterations = int(p["iterations"])
    for k in range(2):
        r.gravity[k] = float(p["gravity"][k])
        r.domain[k] = float(p["domain"][k])
    return r


def _check(rc, h, what):
    if rc != _abi.OK:
        msg = lib().xpbi2d_last_error(h) if h else b""
        raise XpbiError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


class Prototype2D:
    """The 2D column on one GPU (xpbi2d.py's x, v, F and main loop)."""

    _WIDTH = dict(X=2, V=2, F=4, LAMBDA=1)

    def __init__(self, positions, prm, device=0):
        x = np.ascontiguousarray(positions, dtype=np.float64)
        if x.ndim != 2 or x.shape[1] != 2:
            raise ValueError("positions must be (N, 2)")
        self.params = dict(prm)
        self.N = len(x)
        h = C.c_void_p()
        rec = _params_record(prm)
        _check(lib().xpbi2d_create(self.N, x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rec), device,
                                   C.byref(h)), None, "xpbi2d_create")
        self._h = h

    def step(self, substeps=1):
        _check(lib().xpbi2d_step(self._h, int(substeps)), self._h, "xpbi2d_step")

    def set_workgroups(self, single):
        """single=True: one workgroup runs the scene; False: a cooperative grid, one
        thread per particle (the default).  Results are identical."""
        _check(lib().xpbi2d_set_workgroups(self._h, int(bool(single))), self._h, "xpbi2d_set_workgroups")

    def time(self, substeps):
        """Milliseconds of `substeps` substeps (one launch, HIP events)."""
        ms = C.c_float()
        _check(lib().xpbi2d_time(self._h, int(substeps), C.byref(ms)), self._h, "xpbi2d_time")
        return ms.value

    def read(self, field):
        w = self._WIDTH[field]
        out = np.empty((self.N, w) if w > 1 else self.N, np.float64)
        _check(lib().xpbi2d_read(self._h, FIELD2D[field], out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
               self._h, f"read {field}")
        return out.reshape(self.N, 2, 2) if field == "F" else out

    def write(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(lib().xpbi2d_write(self._h, FIELD2D[field], a.ctypes.data_as(C.POINTER(C.c_double)), a.size),
               self._h, f"write {field}")

    x = property(lambda self: self.read("X"))
    v = property(lambda self: self.read("V"))
    F = property(lambda self: self.read("F"))
    lam = property(lambda self: self.read("LAMBDA"))

    def close(self):
        if self._h:
            lib().xpbi2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
