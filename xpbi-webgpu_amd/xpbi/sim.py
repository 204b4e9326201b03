"""Python host surface of the substep engine, mirroring web/js/sim.js.

``Simulation(scene)`` keeps the reference's method names and argument meaning:
``writeMaterials`` (sim.js:84-99), ``writeParams`` (sim.js:159-186),
``encodeSubstep``/``substep`` (sim.js:220-272), ``profileSubstep``
(sim.js:199-215) and the readback hooks ``readPos``/``readVel``
(main.js:291-312).  Everything runs through libxpbi.so; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import FIELD, STAGE, RenderView, XpbiError, check, lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def pack_materials(materials, V0):
    """writeMaterials packing (sim.js:84-99): mu, lambda and mass from E, nu, density (float64)."""
    arr = (_abi.Material * 8)()
    for i, m in enumerate(materials):
        E, nu = float(m["E"]), float(m["nu"])
        r = arr[i]
        r.mu = E / (2 * (1 + nu))
        r.lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        r.density = float(m["density"])
        r.kind = int(m["kind"])
        r.p0, r.p1 = float(m.get("p0", 0.0)), float(m.get("p1", 0.0))
        r.p2, r.p3 = float(m.get("p2", 0.0)), float(m.get("p3", 0.0))
        col = m.get("color", [0.0, 0.0, 0.0])
        r.color[0], r.color[1], r.color[2], r.color[3] = col[0], col[1], col[2], 1.0
        r.boundFriction = float(m.get("boundFriction", 0.4))
        r.mass = float(m["density"]) * V0
    return arr


def _scene_desc(scene, device, keep):
    """SceneDesc of one scene dict; arrays it points to are appended to `keep`."""
    pos = np.ascontiguousarray(scene["positions"], dtype=np.float32).reshape(-1, 3)
    d = _abi.SceneDesc()
    d.N = pos.shape[0]
    d.positions = _fptr(pos)
    keep.append(pos)
    tags = scene.get("tags")
    if tags is not None:
        tags = np.ascontiguousarray(tags, dtype=np.float32)
        d.tags = _fptr(tags)
        keep.append(tags)
    mids = np.ascontiguousarray(scene.get("matIds", np.zeros(d.N)), dtype=np.uint32)
    d.matIds = mids.ctypes.data_as(C.POINTER(C.c_uint32))
    keep.append(mids)
    if scene.get("states") is not None:
        st = np.ascontiguousarray(scene["states"], dtype=np.float32)
        d.states = _fptr(st)
        keep.append(st)
    if scene.get("velocities") is not None:
        v = np.ascontiguousarray(scene["velocities"], dtype=np.float32).reshape(-1, 3)
        d.velocities = _fptr(v)
        keep.append(v)
    d.h = float(scene["h"])
    for k in range(3):
        d.domainMin[k] = float(scene["domainMin"][k])
        d.domainMax[k] = float(scene["domainMax"][k])
    d.device = device
    return d


class SceneInfo:
    """Grid and particle range of one scene of a simulation."""

    def __init__(self, sim_handle, q, scene):
        o = (C.c_float * 3)()
        cs = C.c_float()
        dims = (C.c_int32 * 3)()
        nc, n, fp, fc = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().xpbi_scene_info(sim_handle, q, o, C.byref(cs), dims, C.byref(nc), C.byref(n), C.byref(fp),
                                    C.byref(fc)), sim_handle)
        self.scene = scene
        self.gridOrigin, self.cellSize, self.gridDims = list(o), cs.value, list(dims)
        self.numCells, self.N, self.first, self.firstCell = nc.value, n.value, fp.value, fc.value
        self.h = float(scene["h"])
        self.support = 2 * self.h
        self.V0 = self.h * self.h * self.h

    @property
    def rows(self):
        return slice(self.first, self.first + self.N)


class Simulation:
    """One scene resident on one GPU (the reference's ``new Simulation(device, scene)``), or,
    given a list of scenes, a batch: independent scenes concatenated in one simulation whose
    every kernel is one launch for all of them (fields are read and written whole, scene
    after scene; ``scenes_info[q].rows`` is scene q's slice)."""

    def __init__(self, scene, device=0):
        scenes = list(scene) if isinstance(scene, (list, tuple)) else [scene]
        self.scene = scenes[0]
        keep = []
        descs = (_abi.SceneDesc * len(scenes))()
        for q, sc in enumerate(scenes):
            descs[q] = _scene_desc(sc, device, keep)
        h = C.c_void_p()
        check(lib().xpbi_create_batch(descs, len(scenes), C.byref(h)), None, "xpbi_create_batch")
        self._h = h
        self.scenes_info = [SceneInfo(h, q, sc) for q, sc in enumerate(scenes)]
        self.N = sum(si.N for si in self.scenes_info)
        s0 = self.scenes_info[0]
        self.h, self.support, self.V0 = s0.h, s0.support, s0.V0
        self.gridOrigin, self.cellSize, self.gridDims, self.numCells = (s0.gridOrigin, s0.cellSize, s0.gridDims,
                                                                        s0.numCells)
        self.totalCells = sum(si.numCells for si in self.scenes_info)
        self.num_scenes = len(scenes)
        for q, sc in enumerate(scenes):
            self.writeMaterials(sc["materials"], scene=q)

    # -- reference surface -------------------------------------------------
    def writeMaterials(self, materials, scene=0):
        arr = pack_materials(materials, self.scenes_info[scene].V0)
        check(lib().xpbi_write_scene_materials(self._h, scene, arr, len(materials)), self._h, "writeMaterials")

    def writeParams(self, p, scene=0):
        """sim.js:159-186.  p: dt, domainMin/Max, gravity, collider*, xsphC, omega, contactOmega,
        pbfEps, sCorrK, time, kDamp (default 0.4)."""
        si = self.scenes_info[scene]
        r = _abi.Params()
        for k in range(3):
            r.gridOrigin[k] = si.gridOrigin[k]
            r.gridDims[k] = si.gridDims[k]
            r.domainMin[k] = p["domainMin"][k]
            r.domainMax[k] = p["domainMax"][k]
            r.gravity[k] = p["gravity"][k]
            r.colliderPos[k] = p.get("colliderPos", [0.0, 0.0, 0.0])[k]
            r.colliderVel[k] = p.get("colliderVel", [0.0, 0.0, 0.0])[k]
        r.cellSize = si.cellSize
        r.numCells = si.numCells
        r.dt = p["dt"]
        r.h = si.h
        r.support = si.support
        r.colliderR = p.get("colliderR", 0.0)
        r.colliderActive = 1.0 if p.get("colliderActive", False) else 0.0
        r.N = si.N
        r.V0 = si.V0
        r.xsphC, r.omega = p["xsphC"], p["omega"]
        r.pbfEps, r.sCorrK = p["pbfEps"], p["sCorrK"]
        r.time = p.get("time", 0.0)
        r.contactOmega = p["contactOmega"]
        r.kDamp = p.get("kDamp", 0.4)
        if scene == 0:
            self.params = r
        si.params = r
        check(lib().xpbi_write_scene_params(self._h, scene, C.byref(r)), self._h, "writeParams")

    def encodeSubstep(self, iterations, hasWater):
        check(lib().xpbi_substep(self._h, int(iterations), int(bool(hasWater))), self._h, "substep")

    substep = encodeSubstep

    def step(self, substeps, iterations, hasWater):
        check(lib().xpbi_step(self._h, int(substeps), int(iterations), int(bool(hasWater))), self._h, "step")

    def substep_partial(self, iterations, hasWater, stage):
        st = STAGE[stage] if isinstance(stage, str) else int(stage)
        check(lib().xpbi_substep_partial(self._h, int(iterations), int(bool(hasWater)), st), self._h,
              "substep_partial")

    def profileSubstep(self, iterations, hasWater, cap=4096):
        ms = (C.c_float * cap)()
        n = C.c_int()
        check(lib().xpbi_profile_substep(self._h, int(iterations), int(bool(hasWater)), ms, cap, C.byref(n)),
              self._h, "profileSubstep")
        out = []
        for k in range(min(n.value, cap)):
            out.append((lib().xpbi_profile_label(self._h, k).decode(), ms[k]))
        return out

    def time_substeps(self, substeps, iterations, hasWater):
        ms = C.c_float()
        check(lib().xpbi_time_substeps(self._h, int(substeps), int(iterations), int(bool(hasWater)),
                                       C.byref(ms)), self._h, "time_substeps")
        return ms.value

    def time_substeps_detail(self, substeps, iterations, hasWater, cap=64):
        """Timed substeps with a HIP event after every launch: (total_ms, {label: (ms_sum, launches)})."""
        tot = C.c_float()
        ms = (C.c_float * cap)()
        cnt = (C.c_int * cap)()
        n = C.c_int()
        check(lib().xpbi_time_substeps_detail(self._h, int(substeps), int(iterations), int(bool(hasWater)),
                                              C.byref(tot), ms, cnt, cap, C.byref(n)), self._h,
              "time_substeps_detail")
        out = {}
        for k in range(min(n.value, cap)):
            out[lib().xpbi_profile_label(self._h, k).decode()] = (ms[k], cnt[k])
        return tot.value, out

    def copy_bandwidth(self, nbytes, reps=3):
        """Milliseconds of one device-to-device float4 copy of nbytes (the stream-copy ceiling)."""
        ms = C.c_float()
        check(lib().xpbi_copy_bandwidth(self._h, int(nbytes), int(reps), C.byref(ms)), self._h, "copy_bandwidth")
        return ms.value

    def sync(self):
        check(lib().xpbi_sync(self._h), self._h, "sync")

    # -- readback (main.js:289-313) ----------------------------------------
    _SHAPES = {"POS": (4, np.float32), "VEL": (4, np.float32), "AUX": (4, np.float32), "F": (12, np.float32),
               "L": (12, np.float32), "SP": (12, np.float32), "PERM": (1, np.uint32), "ID": (1, np.uint32)}

    def read(self, field):
        if field == "CELLSTART":
            out = np.empty(self.totalCells, dtype=np.uint32)
        elif field == "MATS":
            out = np.empty((self.num_scenes, 8, 16), dtype=np.float32)
        else:
            w, dt = self._SHAPES[field]
            out = np.empty((self.N, w) if w > 1 else self.N, dtype=dt)
        check(lib().xpbi_read(self._h, FIELD[field], out.ctypes.data, out.nbytes), self._h, f"read {field}")
        return out

    def write(self, field, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().xpbi_write(self._h, FIELD[field], arr.ctypes.data, arr.nbytes), self._h, f"write {field}")

    def render_view(self):
        """Device pointers of pos, velA and the material table, and the sim's stream
        (xpbi_get_render_view); valid until the next step, write or close."""
        v = RenderView()
        check(lib().xpbi_get_render_view(self._h, C.byref(v)), self._h, "render_view")
        return dict(pos=v.pos, vel=v.vel, mats=v.mats, stream=v.stream, N=v.N, scenes=v.scenes, device=v.device)

    def readPos(self):
        return self.read("POS")

    def readVel(self):
        return self.read("VEL")

    def set_residual_monitor(self, on):
        check(lib().xpbi_set_residual_monitor(self._h, int(bool(on))), self._h)

    def read_residuals(self, cap=4096):
        out = np.zeros(2 * cap, dtype=np.float64)
        n = C.c_int()
        check(lib().xpbi_read_residuals(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(n)),
              self._h)
        return out[: 2 * min(n.value, cap)].reshape(-1, 2)

    def set_tolerance(self, tol, check_every=4):
        """Residual-driven Jacobi: with tol > 0 the iteration argument of substep/step is
        a cap and each substep stops once ||C + alpha~ lambda + gamma dt dC.v|| /
        ||C at iteration 0|| <= tol (tested on the device)."""
        check(lib().xpbi_set_tolerance(self._h, float(tol), int(check_every)), self._h)

    def last_iterations(self):
        """(Jacobi iterations applied by the last substep, its final residual or -1)."""
        n, r = C.c_int(), C.c_double()
        check(lib().xpbi_last_iterations(self._h, C.byref(n), C.byref(r)), self._h)
        return n.value, r.value

    def set_launch(self, blocks):
        """0: one block per 256-particle run; > 0: persistent grid of that many blocks."""
        check(lib().xpbi_set_launch(self._h, int(blocks)), self._h, "set_launch")

    def set_neighbour_slots(self, slots):
        """Slots of the support-neighbour list (default 256; 0 walks the cell rows in
        every sweep, like the reference): same results, other speed and memory."""
        check(lib().xpbi_set_neighbour_slots(self._h, int(slots)), self._h, "set_neighbour_slots")

    def get_launch(self):
        b, n = C.c_int(), C.c_int()
        check(lib().xpbi_get_launch(self._h, C.byref(b), C.byref(n)), self._h)
        return b.value, n.value

    def graph_step(self, substeps, iterations, hasWater):
        """step() replayed from a captured HIP graph of two substeps."""
        check(lib().xpbi_graph_step(self._h, int(substeps), int(iterations), int(bool(hasWater))), self._h,
              "graph_step")

    def autotune(self, iterations, hasWater, substeps=4):
        """Pick the fastest launch shape on this simulation; returns (blocks, ms per substep)."""
        b, ms = C.c_int(), C.c_float()
        check(lib().xpbi_autotune(self._h, int(iterations), int(bool(hasWater)), int(substeps), C.byref(b),
                                  C.byref(ms)), self._h, "autotune")
        return b.value, ms.value

    def set_warmstart(self, params):
        """Learned warm start with the given network (canonical float vector of
        MLP_PARAMS), or plain Jacobi from lambda = 0 with params = None."""
        if params is None:
            check(lib().xpbi_set_warmstart(self._h, None, 0), self._h)
            return
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        check(lib().xpbi_set_warmstart(self._h, _fptr(p), p.size), self._h, "set_warmstart")

    def set_lambda0(self, lam0_by_id):
        """Start each substep's Jacobi from lambda0[id] (N floats indexed by persistent
        particle id), or turn that off with None (xpbi_set_lambda0)."""
        if lam0_by_id is None:
            check(lib().xpbi_set_lambda0(self._h, None, 0), self._h, "set_lambda0")
            return
        a = np.ascontiguousarray(lam0_by_id, dtype=np.float32).ravel()
        check(lib().xpbi_set_lambda0(self._h, _fptr(a), a.size), self._h, "set_lambda0")

    def collect_warmstart_pairs(self, iterations, hasWater):
        """One plain substep; returns (features N x 12, targets N x 2 with NaN = no target,
        residual weights N)."""
        from ._abi import MLP_INPUTS
        x = np.zeros((self.N, MLP_INPUTS), np.float32)
        t = np.zeros((self.N, 2), np.float32)
        w = np.zeros(self.N, np.float32)
        check(lib().xpbi_collect_warmstart_pairs(self._h, int(iterations), int(bool(hasWater)), _fptr(x), _fptr(t),
                                                 _fptr(w), self.N), self._h, "collect_warmstart_pairs")
        return x, t, w

    def close(self):
        if getattr(self, "_h", None):
            lib().xpbi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def params_from_scene(scene, dt=None, collider=None):
    """The per-frame Params of main.js:386-402 for a scene: dt = (1/60) timeScale / substeps
    unless the scene fixes dt."""
    if dt is None:
        dt = scene.get("dt") or (1.0 / 60.0) * scene["timeScale"] / scene["substeps"]
    p = dict(dt=dt, domainMin=scene["domainMin"], domainMax=scene["domainMax"], gravity=scene["gravity"],
             xsphC=scene["xsphC"], omega=scene["omega"], contactOmega=scene["contactOmega"],
             pbfEps=scene["pbfEps"], sCorrK=scene["sCorrK"], time=0.0, kDamp=scene.get("kDamp", 0.4))
    col = collider or scene.get("collider")
    if col:
        p.update(colliderPos=col["pos"], colliderR=col["R"], colliderVel=col.get("vel", [0, 0, 0]),
                 colliderActive=True)
    return p
