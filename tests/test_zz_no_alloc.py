"""The allocation-free substep (xpbi_substep_noalloc) through the C ABI.

The caller owns the state (device buffers from hipMalloc, made here through the
HIP runtime, as such a caller would) and the library allocates nothing; every
intermediate lives in the library's static device array.  Bar: bit-exact against
xpbi_substep from the same state (the same kernels on the same data; the
allocation-free path has a shorter neighbour list, so more particles take the
row walk, which gives the same bits).  Three calls in a row share the array: a
landed sand column, then a smaller scene, then the water + sand dam breach.
"""
import ctypes as C

import numpy as np
import pytest

import xpbi
from xpbi import scenes
from helpers import mats_record

pytestmark = pytest.mark.gpu
FIELDS = ("POS", "VEL", "F", "AUX")


class Hip:
    """The few HIP runtime calls a caller owning its device buffers makes."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _lib():
    return xpbi.lib()


def _one(hip, L, scene, settle):
    """A settled snapshot of `scene`, one substep through xpbi_substep and one
    through xpbi_substep_noalloc from it; returns both results."""
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    I, water = scene["iterations"], scene["hasWater"]
    sim.step(settle, I, water)
    snap = {k: sim.read(k) for k in FIELDS}
    sim.substep(I, water)
    ref = {k: sim.read(k) for k in FIELDS}
    mats = mats_record(scene)
    bufs = {k: hip.upload(snap[k]) for k in FIELDS}
    rc = L.xpbi_substep_noalloc(C.byref(sim.params), mats, len(scene["materials"]),
                                bufs["POS"], bufs["VEL"], bufs["F"], bufs["AUX"], I, int(water), None)
    assert rc == 0, L.xpbi_last_error(None)
    assert hip.L.hipDeviceSynchronize() == 0
    got = {k: hip.download(bufs[k], snap[k]) for k in FIELDS}
    for b in bufs.values():
        hip.free(b)
    sim.close()
    return snap, ref, got


def test_noalloc_equals_substep_three_calls():
    hip, L = Hip(), _lib()
    print(f"scratch: {L.xpbi_noalloc_scratch_bytes() / 2**20:.1f} MiB per device", flush=True)
    cases = [("sand column 16x32x16", scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=5, iterations=20), 230),
             ("sand column 8x16x8", scenes.sand_column(8, 16, 8, 0.01, [0.4, 0.3, 0.4], seed=6, iterations=20), 200),
             ("dam breach", scenes.SCENES["Dam breach: water + sand"](seed=1), 120)]
    for label, scene, settle in cases:
        snap, ref, got = _one(hip, L, scene, settle)
        print(f"{label}: N={snap['POS'].shape[0]}", flush=True)
        assert not np.array_equal(ref["POS"], snap["POS"]), f"{label}: the substep moved nothing"
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k]), f"{label}: {k} differs from xpbi_substep"


def test_noalloc_rejects_oversized_input():
    L = _lib()
    scene = scenes.sand_column(4, 4, 4, 0.01, [0.2, 0.2, 0.2], seed=1)
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    p = xpbi._abi.Params.from_buffer_copy(sim.params)
    p.N = 65537
    mats = mats_record(scene)
    dummy = C.c_void_p(256)  # never dereferenced: the size check comes first
    assert L.xpbi_substep_noalloc(C.byref(p), mats, 1, dummy, dummy, dummy, dummy, 3, 0, None) == xpbi._abi.ERR_ARG
    assert b"more particles" in L.xpbi_last_error(None)
    p.N = sim.N
    p.numCells = 1048577
    assert L.xpbi_substep_noalloc(C.byref(p), mats, 1, dummy, dummy, dummy, dummy, 3, 0, None) == xpbi._abi.ERR_ARG
    assert b"more cells" in L.xpbi_last_error(None)
    sim.close()
