"""ctypes binding of libxpbi.so (the C ABI declared in include/xpbi.h).

The library is loaded from the package directory (``xpbi-webgpu_amd/libxpbi.so``,
built in-tree by ``__graft_entry__.build()``).  There is no fallback: if the
library is missing or fails to load, every call raises ``XpbiError``.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# XPBI_LIB points at another build of the library (e.g. an instrumented one)
LIB_PATH = os.environ.get("XPBI_LIB") or os.path.join(PKG_DIR, "libxpbi.so")

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM = 0, -1, -2, -3, -4

# material kinds (web/js/sim.js:5-7)
MAT = dict(ELASTIC=0, DP=1, NACC=2, VM=3, HB=4, SNOW=5, WATER=6)

FIELD = dict(POS=0, VEL=1, F=2, L=3, AUX=4, SP=5, CELLSTART=6, PERM=7, ID=8, MATS=9)
STAGE = dict(SORT=1, REORDER=2, COMPUTE_L=3, SOLVE=4, XSPH=5, UPDATE_F=6, FULL=7)


class XpbiError(RuntimeError):
    pass


class Material(C.Structure):
    """64-B record, web/js/shaders/common.js:21-33."""
    _fields_ = [("mu", C.c_float), ("lam", C.c_float), ("density", C.c_float), ("kind", C.c_uint32),
                ("p0", C.c_float), ("p1", C.c_float), ("p2", C.c_float), ("p3", C.c_float),
                ("color", C.c_float * 4), ("boundFriction", C.c_float), ("mass", C.c_float),
                ("pad0", C.c_float), ("pad1", C.c_float)]


class Params(C.Structure):
    """256-B record, web/js/shaders/common.js:35-60."""
    _fields_ = [("gridOrigin", C.c_float * 3), ("cellSize", C.c_float), ("gridDims", C.c_int32 * 3),
                ("numCells", C.c_uint32), ("domainMin", C.c_float * 3), ("dt", C.c_float),
                ("domainMax", C.c_float * 3), ("h", C.c_float), ("gravity", C.c_float * 3),
                ("support", C.c_float), ("colliderPos", C.c_float * 3), ("colliderR", C.c_float),
                ("colliderVel", C.c_float * 3), ("colliderActive", C.c_float), ("N", C.c_uint32),
                ("V0", C.c_float), ("xsphC", C.c_float), ("omega", C.c_float), ("pbfEps", C.c_float),
                ("sCorrK", C.c_float), ("time", C.c_float), ("contactOmega", C.c_float),
                ("kDamp", C.c_float), ("pad", C.c_float * 27)]


class SceneDesc(C.Structure):
    _fields_ = [("N", C.c_uint32), ("positions", C.POINTER(C.c_float)), ("tags", C.POINTER(C.c_float)),
                ("matIds", C.POINTER(C.c_uint32)), ("states", C.POINTER(C.c_float)),
                ("velocities", C.POINTER(C.c_float)), ("h", C.c_double), ("domainMin", C.c_double * 3),
                ("domainMax", C.c_double * 3), ("device", C.c_int)]


class RenderView(C.Structure):
    """xpbi_render_view: the device buffers a renderer binds (main.js:189-199)."""
    _fields_ = [("pos", C.c_void_p), ("vel", C.c_void_p), ("mats", C.c_void_p), ("stream", C.c_void_p),
                ("N", C.c_uint32), ("scenes", C.c_uint32), ("device", C.c_int)]


class Params2D(C.Structure):
    """xpbi2d_params: the 2D engine's constants (prototype/xpbi2d.py:29-49)."""
    _fields_ = [("h", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
                ("phi_deg", C.c_double), ("dt", C.c_double), ("iterations", C.c_int), ("gravity", C.c_double * 2),
                ("xsph_c", C.c_double), ("domain", C.c_double * 2), ("omega", C.c_double)]


FIELD2D = dict(X=0, V=1, F=2, LAMBDA=3)

assert C.sizeof(Material) == 64 and C.sizeof(Params) == 256

# every symbol include/xpbi.h declares, with (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "xpbi_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(_P)]),
    "xpbi_destroy": (None, [_P]),
    "xpbi_last_error": (C.c_char_p, [_P]),
    "xpbi_grid_info": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "xpbi_write_materials": (C.c_int, [_P, C.POINTER(Material), C.c_int]),
    "xpbi_write_params": (C.c_int, [_P, C.POINTER(Params)]),
    "xpbi_substep": (C.c_int, [_P, C.c_int, C.c_int]),
    "xpbi_step": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "xpbi_substep_partial": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "xpbi_read": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "xpbi_write": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "xpbi_sync": (C.c_int, [_P]),
    "xpbi_get_render_view": (C.c_int, [_P, C.POINTER(RenderView)]),
    "xpbi_profile_substep": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int,
                                       C.POINTER(C.c_int)]),
    "xpbi_profile_label": (C.c_char_p, [_P, C.c_int]),
    "xpbi_time_substeps": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "xpbi_time_substeps_detail": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_int,
                                            C.POINTER(C.c_int)]),
    "xpbi_copy_bandwidth": (C.c_int, [_P, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "xpbi_set_residual_monitor": (C.c_int, [_P, C.c_int]),
    "xpbi_read_residuals": (C.c_int, [_P, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)]),
    "xpbi_set_tolerance": (C.c_int, [_P, C.c_float, C.c_int]),
    "xpbi_create_batch": (C.c_int, [C.POINTER(SceneDesc), C.c_int, C.POINTER(_P)]),
    "xpbi_scene_count": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "xpbi_scene_info": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
    "xpbi_write_scene_materials": (C.c_int, [_P, C.c_int, C.POINTER(Material), C.c_int]),
    "xpbi_write_scene_params": (C.c_int, [_P, C.c_int, C.POINTER(Params)]),
    "xpbi_set_launch": (C.c_int, [_P, C.c_int]),
    "xpbi_set_neighbour_slots": (C.c_int, [_P, C.c_int]),
    "xpbi_get_launch": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "xpbi_graph_step": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "xpbi_autotune": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "xpbi_last_iterations": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "xpbi_set_warmstart": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "xpbi_set_lambda0": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "xpbi_collect_warmstart_pairs": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                               C.POINTER(C.c_float), C.c_size_t]),
    "xpbi_substep_noalloc": (C.c_int, [C.POINTER(Params), C.POINTER(Material), C.c_int, _P, _P, _P, _P, C.c_int,
                                       C.c_int, _P]),
    "xpbi_noalloc_scratch_bytes": (C.c_size_t, []),
    "xpbi_mlp_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "xpbi_mlp_destroy": (None, [_P]),
    "xpbi_mlp_last_error": (C.c_char_p, [_P]),
    "xpbi_mlp_set_params": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "xpbi_mlp_get_params": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "xpbi_mlp_set_data": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.c_size_t]),
    "xpbi_mlp_predict": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float)]),
    "xpbi_mlp_loss_grad": (C.c_int, [_P, C.c_size_t, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "xpbi_mlp_adam_step": (C.c_int, [_P, C.POINTER(C.c_float), C.c_float]),
    "xpbi_mlp_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "xpbi2d_create": (C.c_int, [C.c_uint32, C.POINTER(C.c_double), C.POINTER(Params2D), C.c_int, C.POINTER(_P)]),
    "xpbi2d_destroy": (None, [_P]),
    "xpbi2d_last_error": (C.c_char_p, [_P]),
    "xpbi2d_step": (C.c_int, [_P, C.c_int]),
    "xpbi2d_time": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float)]),
    "xpbi2d_set_workgroups": (C.c_int, [_P, C.c_int]),
    "xpbi2d_read": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
    "xpbi2d_write": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
}

# warm-start network (include/xpbi.h XPBI_MLP_*)
MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 12, 2, 5122

_lib = None


def lib():
    """Load libxpbi.so once; raise XpbiError (never fall back) if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise XpbiError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise XpbiError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, handle=None, what="xpbi call"):
    if rc != OK:
        msg = lib().xpbi_last_error(handle)
        raise XpbiError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
