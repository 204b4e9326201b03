"""C ABI and host surface, without a GPU: the library loads and exports every
symbol the header declares, argument errors are reported before any HIP call,
the JavaScript host generates scenes bit-identical to the Python one, and the
N-API addon loads."""
import ctypes
import hashlib
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import xpbi
from xpbi import scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "xpbi.h")
JS = os.path.join(REPO, "xpbi-webgpu_amd", "js")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xpbi(?:2d)?_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = xpbi.lib()
    names = header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), f"libxpbi.so does not export {n}"
    assert sorted(xpbi._abi.SIGNATURES) == names, "ctypes binding and header disagree"


def test_record_layouts():
    assert ctypes.sizeof(xpbi._abi.Material) == 64
    assert ctypes.sizeof(xpbi._abi.Params) == 256
    assert xpbi._abi.Params.N.offset == 28 * 4      # word 28 (sim.js:182)
    assert xpbi._abi.Params.kDamp.offset == 36 * 4  # word 36 (sim.js:184)
    assert xpbi._abi.Material.mass.offset == 13 * 4


def test_argument_errors_without_device():
    lib = xpbi.lib()
    h = ctypes.c_void_p()
    assert lib.xpbi_create(None, ctypes.byref(h)) == xpbi._abi.ERR_ARG
    assert b"desc is null" in lib.xpbi_last_error(None)
    d = xpbi._abi.SceneDesc()
    d.N = 10
    assert lib.xpbi_create(ctypes.byref(d), ctypes.byref(h)) == xpbi._abi.ERR_ARG  # positions null
    pos = np.zeros((10, 3), np.float32)
    d.positions = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    mids = np.zeros(10, np.uint32)
    d.matIds = mids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    d.h = 0.0
    assert lib.xpbi_create(ctypes.byref(d), ctypes.byref(h)) == xpbi._abi.ERR_ARG
    assert b"h must be" in lib.xpbi_last_error(None)
    d.h = 0.01
    d.domainMax[0] = -1.0
    assert lib.xpbi_create(ctypes.byref(d), ctypes.byref(h)) == xpbi._abi.ERR_ARG
    assert h.value is None
    for f in ("xpbi_substep", "xpbi_sync", "xpbi_grid_info", "xpbi_write_params", "xpbi_read",
              "xpbi_set_neighbour_slots"):
        args = {"xpbi_substep": (None, 1, 0), "xpbi_sync": (None,), "xpbi_write_params": (None, None),
                "xpbi_grid_info": (None, None, None, None, None, None), "xpbi_read": (None, 0, None, 0),
                "xpbi_set_neighbour_slots": (None, 16)}[f]
        assert getattr(lib, f)(*args) == xpbi._abi.ERR_ARG


def _scene_hash(s):
    m = hashlib.sha256()
    for k in ("positions", "tags", "states"):
        m.update(np.ascontiguousarray(s[k], np.float32).tobytes())
    if s.get("velocities") is not None:
        m.update(np.ascontiguousarray(s["velocities"], np.float32).tobytes())
    return m.hexdigest()


# particle counts of the eight reference scenes (web/js/scenes.js:94-176, counted from the
# lattice formulas, independent of the random jitter)
REF_COUNTS = {"Sand: column collapse": 12312, "Sand: two blocks collide": 15552, "Snow: snowball smash": 14510,
              "Plasticine: drop & poke": 7799, "Goo: Herschel-Bulkley": 9357, "Snow: classic (Stomakhin)": 3810,
              "Dam breach: water + sand": 27965, "Elastic: StVK block": 4913}


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
@pytest.mark.parametrize("name", list(scenes.SCENES) + ["column128k"])
def test_js_scenes_bit_identical(name):
    f = scenes.sand_column_128k if name == "column128k" else scenes.SCENES[name]
    for seed in (0, 3):
        s = f(seed=seed)
        out = subprocess.check_output(["node", os.path.join(JS, "run.js"), "hash", name, str(seed)], text=True)
        js = json.loads(out)
        assert js["N"] == s["positions"].shape[0]
        assert js["sha256"] == _scene_hash(s)
    if name in REF_COUNTS:
        assert s["positions"].shape[0] == REF_COUNTS[name]


def test_scene_generation_properties():
    s = scenes.sand_column_128k()
    p = s["positions"].astype(np.float64)
    assert p.shape == (128000, 3)
    # lattice at r = 0.01 with jitter within +-0.05 r
    r = 0.01
    lattice = np.round((p - [0.8, 0.0, 0.8]) / r - 0.5)
    off = (p - [0.8, 0.0, 0.8]) / r - 0.5 - lattice
    assert np.abs(off).max() <= 0.05 + 2e-5  # + f32 rounding of positions near 1.2 m
    assert (np.diff(np.sort(np.unique(lattice[:, 0]))) == 1).all()
    tags = s["tags"]
    assert ((tags >= 0) & (tags < 0.49)).all()
    # seeded: same seed same scene, different seed different jitter
    assert np.array_equal(scenes.sand_column_128k(2)["positions"], s["positions"])
    assert not np.array_equal(scenes.sand_column_128k(5)["positions"], s["positions"])
    u = scenes.rand_unit(1, 0, np.arange(200000))
    assert abs(u.mean() - 0.5) < 0.005 and u.min() >= 0 and u.max() < 1


def test_dam_breach_scales():
    """The dam breach at scale 1 is the 1M layout (600k water beside 400k sand); at
    scale 0.5 the blocks keep spacing r = 0.005 with half the lattice counts per axis
    (30x50x50 water, 25x40x50 sand) in a domain of half the extent, side by side."""
    full = scenes.dam_breach_1m()
    assert full["positions"].shape[0] == 1_000_000
    assert int((full["matIds"] == 1).sum()) == 600_000 and int((full["matIds"] == 0).sum()) == 400_000
    half = scenes.dam_breach(0.5, seed=11)
    p, m = half["positions"].astype(np.float64), half["matIds"]
    assert int((m == 1).sum()) == 30 * 50 * 50 and int((m == 0).sum()) == 25 * 40 * 50
    assert half["domainMax"] == [0.75, 0.375, 0.25]
    assert (p >= 0).all() and (p <= np.array(half["domainMax"]) + 1e-6).all()
    water, sand = p[m == 1], p[m == 0]
    assert water[:, 0].max() < sand[:, 0].min() + 0.005  # water left of the sand wall
    assert abs(np.ptp(sand[:, 1]) - 40 * 0.005) < 0.01 and abs(np.ptp(water[:, 1]) - 50 * 0.005) < 0.01
    assert half["sCorrK"] == full["sCorrK"] == 0.1 and half["hasWater"]


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_napi_addon_loads():
    out = subprocess.check_output(["node", os.path.join(JS, "run.js"), "exports"], text=True)
    names = json.loads(out)
    for n in ("create", "destroy", "writeMaterials", "writeParams", "step", "read", "write", "profileSubstep",
              "timeSubsteps", "gridInfo", "sync", "createBatch", "sceneInfo", "writeSceneMaterials",
              "writeSceneParams", "setTolerance", "lastIterations", "setWarmstart", "graphStep", "setLaunch",
              "getLaunch", "autotune", "setNeighbourSlots"):
        assert n in names
