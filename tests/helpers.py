"""Shared helpers of the test-suite: grid/param records for the oracle, error metrics."""
import math

import numpy as np

import xpbi


def grid_of(scene):
    """Grid of web/js/sim.js:20-26 in double arithmetic."""
    h = float(scene["h"])
    cs = 2 * h
    dims = [int(math.ceil((scene["domainMax"][k] - scene["domainMin"][k]) / cs)) + 4 for k in range(3)]
    origin = [scene["domainMin"][k] - 2 * cs for k in range(3)]
    return origin, cs, dims, dims[0] * dims[1] * dims[2]


def params_record(scene, dt=None, N=None, collider=None):
    """The 256-B Params record the host writes (sim.js:159-186), without a device."""
    p = xpbi.params_from_scene(scene, dt=dt, collider=collider)
    origin, cs, dims, C = grid_of(scene)
    r = xpbi._abi.Params()
    for k in range(3):
        r.gridOrigin[k] = origin[k]
        r.gridDims[k] = dims[k]
        r.domainMin[k] = p["domainMin"][k]
        r.domainMax[k] = p["domainMax"][k]
        r.gravity[k] = p["gravity"][k]
        r.colliderPos[k] = p.get("colliderPos", [0, 0, 0])[k]
        r.colliderVel[k] = p.get("colliderVel", [0, 0, 0])[k]
    h = float(scene["h"])
    r.cellSize = cs
    r.numCells = C
    r.dt = p["dt"]
    r.h = h
    r.support = 2 * h
    r.colliderR = p.get("colliderR", 0.0)
    r.colliderActive = 1.0 if p.get("colliderActive") else 0.0
    r.N = N if N is not None else np.asarray(scene["positions"]).reshape(-1, 3).shape[0]
    r.V0 = h * h * h
    r.xsphC, r.omega, r.pbfEps, r.sCorrK = p["xsphC"], p["omega"], p["pbfEps"], p["sCorrK"]
    r.contactOmega, r.kDamp = p["contactOmega"], p["kDamp"]
    return r, C


def mats_record(scene):
    h = float(scene["h"])
    return xpbi.pack_materials(scene["materials"], h * h * h)


def rel_err_rows(a, b, floor=1e-6):
    """max over particles of ||a_i - b_i|| / max(||b_i||, floor)."""
    d = np.linalg.norm(a - b, axis=1)
    n = np.maximum(np.linalg.norm(b, axis=1), floor)
    return float(np.max(d / n)) if d.size else 0.0


def F3(F12):
    """reference 12-float padded columns -> (n, 9) column-major entries."""
    F12 = np.asarray(F12).reshape(-1, 3, 4)
    return F12[:, :, :3].reshape(-1, 9)
