"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker; the product library and host never import it.

State arrays use the reference's buffer layout (web/js/sim.js:32-50): pos, vel
and aux are vec4 per particle; F, L and Sp are three vec4-padded columns.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = [os.path.join(HERE, "xpbi_oracle.cpp"), os.path.join(HERE, "xpbi_oracle.hpp"), os.path.abspath(__file__)]
STAGE = dict(SORT=1, REORDER=2, COMPUTE_L=3, SOLVE=4, XSPH=5, UPDATE_F=6, FULL=7)

GXX_CMD = ["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
           os.path.join(HERE, "xpbi_oracle.cpp"), "-o", LIB]


def build(force=False):
    """Compile liboracle.so with g++ (no FMA contraction, OpenMP)."""
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(p) for p in SRC)
        if os.path.getmtime(LIB) >= newest:
            return LIB
    subprocess.run(GXX_CMD, check=True)
    return LIB


class _St(C.Structure):
    pass


def _mk_state_struct(ctype):
    P = C.POINTER(ctype)
    U = C.POINTER(C.c_uint32)

    class S(C.Structure):
        _fields_ = [("N", C.c_uint32), ("C", C.c_uint32), ("pos", P), ("velA", P), ("velB", P), ("F", P),
                    ("L", P), ("aux", P), ("Sp", P), ("cellTab", U), ("sortedIdx", U), ("id", U)]
    return S


class _Opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("wsMode", C.c_int), ("wsParams", C.c_void_p), ("feat", C.c_void_p),
                ("target", C.c_void_p), ("itersDone", C.c_int), ("lam0ById", C.c_void_p)]


S32 = _mk_state_struct(C.c_float)
S64 = _mk_state_struct(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.xo_substep_f32.argtypes = [C.POINTER(S32), vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.xo_substep_f64.argtypes = [C.POINTER(S64), vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.xo_substep_f32.restype = L.xo_substep_f64.restype = C.c_int
        L.xo_substep_ex_f32.argtypes = [C.POINTER(S32), vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                        C.POINTER(_Opts)]
        L.xo_substep_ex_f64.argtypes = [C.POINTER(S64), vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                        C.POINTER(_Opts)]
        L.xo_substep_ex_f32.restype = L.xo_substep_ex_f64.restype = C.c_int
        L.xo_mlp_forward.argtypes = [vp, C.c_int, vp, vp]
        L.xo_hb_return_map_f32.argtypes = [C.c_int, vp, C.c_float] + [vp] * 4
        L.xo_hb_return_map_f64.argtypes = [C.c_int, vp, C.c_double] + [vp] * 4
        L.xo_kernel_f32.argtypes = [C.c_float, C.c_int] + [vp] * 6
        L.xo_kernel_f64.argtypes = [C.c_double, C.c_int] + [vp] * 6
        L.xo_svd3_f32.argtypes = [C.c_int, vp, vp, vp, vp]
        L.xo_svd3_f64.argtypes = [C.c_int, vp, vp, vp, vp]
        L.xo_return_map_f32.argtypes = [C.c_int, vp, C.c_float, C.c_int] + [vp] * 6
        L.xo_return_map_f64.argtypes = [C.c_int, vp, C.c_double, C.c_int] + [vp] * 6
        L.xo_cell_index.argtypes = [vp, C.c_int, vp, vp]
        _lib = L
    return _lib


class use_library:
    """Context manager: the functions of this module call the oracle library at
    `path` (a build of a modified copy of the sources) instead of liboracle.so."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib, LIB
        self.saved = (_lib, LIB)
        _lib, LIB = None, self.path
        lib()
        return self

    def __exit__(self, *exc):
        global _lib, LIB
        _lib, LIB = self.saved
        return False


def _p(a):
    return a.ctypes.data


class State:
    """Particle state in the reference layout (dtype float32 or float64)."""
    FIELDS = ("pos", "velA", "velB", "F", "L", "aux", "Sp")

    def __init__(self, N, numCells, dtype=np.float32):
        self.N, self.C, self.dtype = int(N), int(numCells), np.dtype(dtype)
        self.pos = np.zeros((N, 4), dtype)
        self.velA = np.zeros((N, 4), dtype)
        self.velB = np.zeros((N, 4), dtype)
        self.F = np.zeros((N, 12), dtype)
        self.L = np.zeros((N, 12), dtype)
        self.aux = np.zeros((N, 4), dtype)
        self.Sp = np.zeros((N, 12), dtype)
        self.cellTab = np.zeros(numCells, np.uint32)
        self.sortedIdx = np.arange(N, dtype=np.uint32)
        self.id = np.arange(N, dtype=np.uint32)

    @classmethod
    def from_arrays(cls, numCells, pos, vel, F, aux, ids=None, dtype=np.float32):
        s = cls(pos.shape[0], numCells, dtype)
        s.pos[:] = pos
        s.velA[:] = vel
        s.F[:] = F
        s.aux[:] = aux
        if ids is not None:
            s.id[:] = ids
        return s

    def copy(self, dtype=None):
        s = State(self.N, self.C, dtype or self.dtype)
        for f in self.FIELDS:
            getattr(s, f)[:] = getattr(self, f)
        s.cellTab[:] = self.cellTab
        s.sortedIdx[:] = self.sortedIdx
        s.id[:] = self.id
        return s

    def _struct(self):
        T = S32 if self.dtype == np.float32 else S64
        ct = C.c_float if self.dtype == np.float32 else C.c_double
        P = lambda a: a.ctypes.data_as(C.POINTER(ct))
        U = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        return T(self.N, self.C, P(self.pos), P(self.velA), P(self.velB), P(self.F), P(self.L), P(self.aux),
                 P(self.Sp), U(self.cellTab), U(self.sortedIdx), U(self.id))


def substep(state, params, mats, iterations, has_water, stop="FULL", nthreads=None, resid=None, tol=0.0,
            warm=None, collect=False, feat=None, target=None, lam0_by_id=None):
    """Run one (partial) substep in place.  params/mats: ctypes records (bytes-compatible).

    tol > 0: `iterations` is a cap; the Jacobi loop stops after the first
    iteration whose residual ||num|| / ||C at lambda = 0|| is <= tol.
    warm: canonical MLP parameters -> learned warm start; collect: training
    pairs; lam0_by_id: a given lambda0 per persistent particle id, applied as the
    first iteration's Delta-lambda where the constraint is active (the ceiling
    experiments of the warm start).  feat (N x 16) and target (N x 2) receive the feature rows and targets
    (allocated when None).  Returns the number of Jacobi iterations applied."""
    st = state._struct()
    nt = nthreads or os.cpu_count() or 1
    rp = resid.ctypes.data if resid is not None else None
    mode = 1 if warm is not None else (2 if collect else (3 if lam0_by_id is not None else 0))
    lp = None
    if lam0_by_id is not None:
        lp = np.ascontiguousarray(lam0_by_id, state.dtype)
        assert lp.shape == (state.N,)
    wp = None
    if warm is not None:
        wp = np.ascontiguousarray(warm, np.float32)
        assert wp.size == 5122
    if mode and feat is None:
        feat = np.zeros((state.N, 16), state.dtype)
    if mode == 2 and target is None:
        target = np.zeros((state.N, 2), state.dtype)
    o = _Opts(float(tol), mode, wp.ctypes.data if wp is not None else None,
              feat.ctypes.data if feat is not None else None, target.ctypes.data if target is not None else None, 0,
              lp.ctypes.data if lp is not None else None)
    f = lib().xo_substep_ex_f32 if state.dtype == np.float32 else lib().xo_substep_ex_f64
    rc = f(C.byref(st), C.addressof(params), C.addressof(mats), int(iterations), int(bool(has_water)),
           STAGE[stop] if isinstance(stop, str) else int(stop), int(nt), rp, C.byref(o))
    if rc != 0:
        raise RuntimeError("oracle substep rejected the state (N/numCells mismatch)")
    return o.itersDone


def mlp_forward(params, x):
    """The C++ restatement of the warm-start MLP forward (n x 12 -> n x 2)."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    y = np.zeros((x.shape[0], 2), np.float64)
    lib().xo_mlp_forward(p.ctypes.data, x.shape[0], x.ctypes.data, y.ctypes.data)
    return y


def initial_state(scene, numCells, dtype=np.float32):
    """The state xpbi_create / createBuffers uploads (sim.js:52-81)."""
    pos3 = np.asarray(scene["positions"], np.float32).reshape(-1, 3)
    N = pos3.shape[0]
    h = float(scene["h"])
    V0 = np.float32(h * h * h)
    pos = np.zeros((N, 4), np.float32)
    pos[:, :3] = pos3
    pos[:, 3] = scene["tags"] if scene.get("tags") is not None else scene["matIds"].astype(np.float32)
    vel = np.zeros((N, 4), np.float32)
    if scene.get("velocities") is not None:
        vel[:, :3] = np.asarray(scene["velocities"], np.float32).reshape(-1, 3)
    F = np.zeros((N, 12), np.float32)
    F[:, 0] = F[:, 5] = F[:, 10] = 1.0
    aux = np.zeros((N, 4), np.float32)
    aux[:, 0] = V0
    if scene.get("states") is not None:
        aux[:, 2] = scene["states"]
    return State.from_arrays(numCells, pos, vel, F, aux, dtype=dtype)


def kernels(h, r, dxv, dtype=np.float64):
    r = np.ascontiguousarray(r, dtype)
    dxv = np.ascontiguousarray(dxv, dtype).reshape(-1, 3)
    n = r.size
    W, P6 = np.zeros(n, dtype), np.zeros(n, dtype)
    G, S = np.zeros((n, 3), dtype), np.zeros((n, 3), dtype)
    f = lib().xo_kernel_f32 if dtype == np.float32 else lib().xo_kernel_f64
    f(h, n, _p(r), _p(dxv), _p(W), _p(G), _p(P6), _p(S))
    return W, G, P6, S


def svd3(F, dtype=np.float64):
    """F: (n,3,3) ordinary row-major matrices.  Returns U, sig, V as (n,3,3), (n,3), (n,3,3)."""
    F = np.asarray(F, dtype).reshape(-1, 3, 3)
    n = F.shape[0]
    cm = np.ascontiguousarray(np.transpose(F, (0, 2, 1)).reshape(n, 9))  # column-major flatten
    U, V, sig = np.zeros((n, 9), dtype), np.zeros((n, 9), dtype), np.zeros((n, 3), dtype)
    f = lib().xo_svd3_f32 if dtype == np.float32 else lib().xo_svd3_f64
    f(n, _p(cm), _p(U), _p(sig), _p(V))
    U = np.transpose(U.reshape(n, 3, 3), (0, 2, 1))
    V = np.transpose(V.reshape(n, 3, 3), (0, 2, 1))
    return U, sig, V


def return_map(mat, sig, state, dt, harden=False, dtype=np.float64):
    """Return map + StVK psi, dpsi for n singular-value triples with one material record."""
    sig = np.ascontiguousarray(sig, dtype).reshape(-1, 3)
    n = sig.shape[0]
    state = np.ascontiguousarray(np.broadcast_to(np.asarray(state, dtype), (n,)))
    so, sto = np.zeros((n, 3), dtype), np.zeros(n, dtype)
    psi, dpsi = np.zeros(n, dtype), np.zeros((n, 3), dtype)
    f = lib().xo_return_map_f32 if dtype == np.float32 else lib().xo_return_map_f64
    f(n, C.addressof(mat), dt, int(harden), _p(sig), _p(state), _p(so), _p(sto), _p(psi), _p(dpsi))
    return so, sto, psi, dpsi


def hb_return_map(mat, sig, hint, dt, dtype=np.float64):
    """HB return map with bracket hints (-1: none): (sigma out, root position or NaN)."""
    sig = np.ascontiguousarray(sig, dtype).reshape(-1, 3)
    n = sig.shape[0]
    hint = np.ascontiguousarray(np.broadcast_to(np.asarray(hint, dtype), (n,)))
    so, t = np.zeros((n, 3), dtype), np.zeros(n, dtype)
    f = lib().xo_hb_return_map_f32 if dtype == np.float32 else lib().xo_hb_return_map_f64
    f(n, C.addressof(mat), dt, _p(sig), _p(hint), _p(so), _p(t))
    return so, t


def cell_index(params, p):
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.uint32)
    lib().xo_cell_index(C.addressof(params), p.shape[0], _p(p), _p(out))
    return out


SAN_DIR = os.path.join(HERE, "_san")


def build_sanitized():
    """The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the driver
    oracle/sanitize_main.cpp linked with xpbi_oracle.cpp into oracle/_san/oracle_san
    (a finding aborts the run).  Returns the executable's path."""
    os.makedirs(SAN_DIR, exist_ok=True)
    exe = os.path.join(SAN_DIR, "oracle_san")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", os.path.join(HERE, "xpbi_oracle.cpp"),
           os.path.join(HERE, "sanitize_main.cpp"), "-o", exe]
    subprocess.run(cmd, check=True)
    return exe
