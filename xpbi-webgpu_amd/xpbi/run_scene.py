"""Step a named workload through the C ABI, with checkpoints.

    python -m xpbi.run_scene WORKLOAD [--substeps K] [--load PATH] [--save PATH]
                             [--s-corr K] [--every N] [--copy-mib M]
                             [--tol R [--cap I]] [--warmstart NET.npy]

WORKLOAD is one of the benchmark workloads (sand_column_1m, sand_column_128k,
dam_breach_1m, vm_poke_1m) or the name of a reference scene ("Sand: column
collapse", ...).  The run takes the workload's own dt and iteration count,
optionally continues from a checkpoint written by an earlier run (--load), runs
K substeps and optionally writes a checkpoint (--save).  A collider with a
velocity sits at pos + vel * t, t = substep count * dt, set before each substep.
Prints one progress line every N substeps (default 50) and a final line with the
substep count and the wall time per substep.  --copy-mib M ends the run with one
M MiB device copy through the library's copy kernel: the calibration of the byte
counters in xpbi.counters.  --tol R runs each substep's Jacobi loop to residual R
(at most --cap iterations, default 200) instead of the workload's fixed count;
--warmstart starts it from the learned warm start of a saved network
(python -m xpbi.warmstart --out NET.npy).

A profiler run of a few substeps of a late state of a scene is then
    python -m xpbi.run_scene dam_breach_1m --substeps 100 --save state.npz
    python -m xpbi.counters --copy-bytes 268435456 -- \
        python3 -m xpbi.run_scene dam_breach_1m --load state.npz --substeps 2 --copy-mib 256
"""
import argparse
import sys
import time

from . import checkpoint, scenes
from .sim import Simulation, params_from_scene

WORKLOADS = {
    "sand_column_1m": scenes.sand_column_1m,
    "sand_column_128k": scenes.sand_column_128k,
    "dam_breach_1m": scenes.dam_breach_1m,
    "vm_poke_1m": scenes.vm_poke_1m,
}


def make_scene(name):
    if name in WORKLOADS:
        return WORKLOADS[name]()
    if name in scenes.SCENES:
        return scenes.SCENES[name]()
    raise KeyError(f"unknown workload {name!r}; known: {sorted(WORKLOADS) + sorted(scenes.SCENES)}")


def run(name, substeps, load=None, save=None, s_corr=None, every=50, out=sys.stdout, copy_mib=0, tol=None,
        cap=200, warmstart=None):
    """name: a workload or reference scene name, or a scene dict."""
    sc = make_scene(name) if isinstance(name, str) else dict(name)
    name = name if isinstance(name, str) else sc.get("info", "scene")
    if s_corr is not None:
        sc["sCorrK"] = s_corr
    sim = Simulation(sc, device=0)
    dt = sc.get("dt") or (1.0 / 60.0) * sc["timeScale"] / sc["substeps"]
    I, water = sc["iterations"], bool(sc["hasWater"])
    col = sc.get("collider")
    k0 = checkpoint.load(sim, load) if load else 0
    sim.writeParams(params_from_scene(sc, dt=dt))
    if warmstart:
        from . import warmstart as ws
        sim.set_warmstart(ws.load(warmstart))
    if tol is not None:
        sim.set_tolerance(tol, 1)
        I = cap
    t0 = time.perf_counter()
    for k in range(k0, k0 + substeps):
        if col is not None:
            t = k * dt
            pos = [col["pos"][j] + col.get("vel", [0.0, 0.0, 0.0])[j] * t for j in range(3)]
            sim.writeParams(params_from_scene(sc, dt=dt, collider=dict(pos=pos, R=col["R"], vel=col.get("vel"))))
        sim.step(1, I, water)
        if every and (k + 1 - k0) % every == 0:
            sim.sync()
            print(f"{name}: substep {k + 1}", file=out, flush=True)
    sim.sync()
    ms = (time.perf_counter() - t0) * 1e3 / max(substeps, 1)
    if save:
        checkpoint.save(sim, save, k0 + substeps)
    if copy_mib:
        sim.copy_bandwidth(copy_mib << 20, reps=1)
    print(f"{name}: {sim.N} particles, substeps {k0} -> {k0 + substeps}, {ms:.3f} ms per substep (wall)",
          file=out, flush=True)
    sim.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("workload")
    ap.add_argument("--substeps", type=int, default=10)
    ap.add_argument("--load")
    ap.add_argument("--save")
    ap.add_argument("--s-corr", type=float, default=None)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--tol", type=float, default=None)
    ap.add_argument("--cap", type=int, default=200)
    ap.add_argument("--warmstart")
    a = ap.parse_args(argv)
    run(a.workload, a.substeps, a.load, a.save, a.s_corr, a.every, copy_mib=a.copy_mib, tol=a.tol, cap=a.cap,
        warmstart=a.warmstart)


if __name__ == "__main__":
    main()
