#!/usr/bin/env python3
"""Benchmark of the XPBI substep on one MI355X.

Workload: 1M-particle Drucker-Prager sand -- an 80x160x80 seeded-jitter column
at r = 0.005 m (1,024,000 particles; E 3.5e5 Pa, nu 0.3, rho 2000, phi 30 deg),
dt = 1/6000 s.  Before warm-up the column is advanced `--settle` substeps
(untimed, 20 iterations each) so that it has landed and deforms; a free-falling
column would skip the constraint sweeps.  A step is one substep of all
particles, with every array resident in HBM.

`value` = particles x substeps / wall seconds of `--steps` substeps whose Jacobi
iterations run to residual 1e-4 (cap `--residual-cap`) from the learned warm
start; `plain_value` is the same from lambda = 0, from the same snapshot, to the
same residual.  The warm-start network is trained in the run on the GPU (pairs
from DP columns of 128,000 particles at r = 0.01 and 0.005 and a von Mises poke,
never the 1M column itself).  With --no-residual, `value` is the fixed
20-iteration throughput instead.

Also reported: the fixed 20-iteration throughput replayed from a HIP graph under
the autotuned launch shape; per-kernel HIP-event times of an eager run; the HBM
roofline of the kernel with the largest share of the substep (solveB) and of
solveA beside it, with HBM traffic, VALU and address-unit busy fractions from
rocprofv3 counter passes over one substep of the settled column (loaded from a
checkpoint written after settling) run at the end; the 64-scene batched ensemble; the mixed water + sand scene;
the von Mises poke to residual; the prototype's 2D fp64 40x80 column
beside its numpy oracle; and the f32 C++ oracle's substep on the host cores (a
reported baseline).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "xpbi-webgpu_amd")]

METRIC = "particle-substeps/s @1M DP sand, Jacobi iters to 1e-4: learned vs plain"
HBM_PEAK_GBS = 8000.0
# compulsory bytes per particle per launch (every per-particle array read or
# written once; neighbour re-reads not counted)
ALG_BYTES = {"solveA": 164, "solveB": 96, "computeL": 68, "xsph": 48, "updateF": 148, "integrate": 64,
             "reorder": 172, "pbfA": 36, "pbfB": 52}
GATHER = ("computeL", "solveA", "solveB", "xsph", "updateF")


def to_residual(sim, snap, params, tol, cap, substeps, before=None):
    """From the snapshot, `substeps` substeps with Jacobi run to residual `tol`
    (cap `cap`), plain (params None) or with the learned warm start.  The solve is
    deterministic: a first pass reads each substep's iteration count, a second from
    the same snapshot is timed without host reads between substeps.  Returns the
    iterations of each substep, wall ms per substep and the final positions
    indexed by particle id."""
    import numpy as np
    sim.set_warmstart(params)
    sim.set_tolerance(tol, 1)
    its = []
    for k, v in snap.items():
        sim.write(k, v)
    for k in range(substeps):
        if before:
            before(k)
        sim.substep(cap, False)
        its.append(sim.last_iterations()[0])
    first = sim.read("POS")
    for k, v in snap.items():
        sim.write(k, v)
    sim.sync()
    t0 = time.perf_counter()
    for k in range(substeps):
        if before:
            before(k)
        sim.substep(cap, False)
    sim.sync()
    ms = (time.perf_counter() - t0) * 1e3 / substeps
    pos, ids = sim.read("POS"), sim.read("ID")
    assert np.array_equal(pos, first), "the timed pass differs from the counted one"
    out = np.zeros_like(pos)
    out[ids] = pos
    sim.set_tolerance(0.0, 4)
    sim.set_warmstart(None)
    return its, ms, out


def warm_start_ceiling(sim, start, tol, cap, substeps, before=None):
    """What any predictor of the first iteration's multipliers could gain, on the
    same substeps: `substeps` consecutive substeps to residual `tol` from the
    state `start` (the learned-vs-plain snapshot), each run three times from the
    same state -- plain (lambda = 0),
    exact (the first iteration applies lambda* of that substep's plain solve,
    per particle id: a perfect predictor of lambda0) and previous (lambda of the
    previous substep's solve, free at run time) -- through xpbi_set_lambda0.
    Iterations count the given first iteration; before(k) places a moving
    collider for substep k.  Returns mean iterations, ratios and wall ms per
    substep (synchronised around each substep)."""
    import numpy as np
    keys = ("POS", "VEL", "F", "AUX", "ID")
    sim.set_warmstart(None)
    sim.set_tolerance(tol, 1)

    def by_id(vals, ids):
        out = np.zeros(sim.N, np.float32)
        out[ids] = vals
        return out

    def one(state, lam0, k):
        for key, v in state.items():
            sim.write(key, v)
        if before:
            before(k)
        sim.set_lambda0(lam0)
        sim.sync()
        t0 = time.perf_counter()
        sim.substep(cap, False)
        sim.sync()
        ms = (time.perf_counter() - t0) * 1e3
        return sim.last_iterations()[0], ms
    states, stars, prevs, plain = [], [], [], []
    for k, v in start.items():
        sim.write(k, v)
    for j in range(substeps):
        st = {k: sim.read(k) for k in keys}
        states.append(st)
        prevs.append(by_id(st["AUX"][:, 1], st["ID"]))
        n, ms = one(st, None, j)
        plain.append((n, ms))
        stars.append(by_id(sim.read("AUX")[:, 1], sim.read("ID")))
    after = {k: sim.read(k) for k in keys}
    exact = [one(states[k], stars[k], k) for k in range(substeps)]
    prev = [one(states[k], prevs[k], k) for k in range(substeps)]
    sim.set_lambda0(None)
    for k, v in after.items():
        sim.write(k, v)
    sim.set_tolerance(0.0, 4)
    mean = lambda rows, j: float(np.mean([r[j] for r in rows]))
    out = dict(tol=tol, cap=cap, substeps=substeps, plain_iterations=mean(plain, 0),
               exact_lambda0_iterations=mean(exact, 0), previous_lambda0_iterations=mean(prev, 0),
               plain_ms=round(mean(plain, 1), 3), exact_lambda0_ms=round(mean(exact, 1), 3),
               previous_lambda0_ms=round(mean(prev, 1), 3), per_substep=dict(
                   plain=[r[0] for r in plain], exact=[r[0] for r in exact], previous=[r[0] for r in prev]))
    out["exact_iteration_ratio"] = round(out["plain_iterations"] / out["exact_lambda0_iterations"], 3)
    out["previous_iteration_ratio"] = round(out["plain_iterations"] / out["previous_lambda0_iterations"], 3)
    out["exact_time_ratio"] = round(out["plain_ms"] / out["exact_lambda0_ms"], 3)
    out["note"] = ("exact: the first iteration applies each particle's converged lambda of the plain solve of the "
                   "same substep (a perfect lambda0 predictor); previous: the last substep's converged lambda")
    return out


def trained_network(weights=None, log=None):
    """The warm-start network: trained here on the GPU (pairs of the training columns),
    or loaded from `weights`.  Returns (params, training record)."""
    from xpbi import warmstart
    t0 = time.perf_counter()
    if weights:
        return warmstart.load(weights), dict(weights=os.path.basename(weights))
    X, T, W = warmstart.training_set(log=log)
    t1 = time.perf_counter()
    params, losses = warmstart.train(X, T, W, epochs=30)
    return params, dict(scenes=[c[0] for c in warmstart.training_columns()], pairs=int(len(X)), epochs=len(losses),
                        schedule="30 epochs at lr 1e-3, 15 at 3e-4, 15 at 1e-4",
                        loss_first_epoch=float(losses[0]), loss_last_epoch=float(losses[-1]),
                        collect_s=round(t1 - t0, 2), train_s=round(time.perf_counter() - t1, 2))


def learned_vs_plain(sim, tol, cap, substeps, params, train, before=None, warmup=0):
    import numpy as np
    snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    if warmup:
        to_residual(sim, snap, None, tol, cap, warmup, before)
        to_residual(sim, snap, params, tol, cap, warmup, before)
    pit, pms, ppos = to_residual(sim, snap, None, tol, cap, substeps, before)
    lit, lms, lpos = to_residual(sim, snap, params, tol, cap, substeps, before)
    pi, li = float(np.mean(pit)), float(np.mean(lit))
    ceiling = warm_start_ceiling(sim, snap, tol, cap, substeps, before)
    return dict(tol=tol, cap=cap, substeps=substeps, plain_iterations=pi, learned_iterations=li,
                iteration_ratio=round(pi / li, 3), plain_ms_per_substep=round(pms, 4),
                learned_ms_per_substep=round(lms, 4), time_ratio=round(pms / lms, 3),
                plain_particle_substeps_per_s=round(sim.N / (pms / 1e3), 1),
                learned_particle_substeps_per_s=round(sim.N / (lms / 1e3), 1),
                final_position_max_diff_m=float(np.abs(ppos[:, :3] - lpos[:, :3]).max()),
                note="learned iterations count the learned first iteration", training=train, ceiling=ceiling)


def along_collapse(scene, params, points=(120, 200, 337, 500, 1000), substeps=10, tol=1e-4, cap=200):
    """Learned against plain Jacobi to `tol` from snapshots of the column's
    collapse from rest (a fresh simulation, fixed-iteration substeps between the
    snapshots): `substeps` substeps from each, iterations counted as in
    learned_vs_plain.  The ratio depends on the state of the collapse."""
    import numpy as np
    import xpbi
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    I = scene["iterations"]
    done, rows = 0, []
    for at in points:
        sim.graph_step(at - done, I, False)
        sim.sync()
        done = at
        snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
        pit, pms, _ = to_residual(sim, snap, None, tol, cap, substeps)
        lit, lms, _ = to_residual(sim, snap, params, tol, cap, substeps)
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_warmstart(None)
        sim.set_tolerance(0.0, 4)
        pi, li = float(np.mean(pit)), float(np.mean(lit))
        rows.append(dict(substep=at, plain_iterations=pi, learned_iterations=li, iteration_ratio=round(pi / li, 3),
                         time_ratio=round(pms / lms, 3)))
    sim.close()
    return dict(tol=tol, substeps_each=substeps, snapshots=rows,
                note="a fresh column from rest, fixed-iteration substeps between the snapshots; the headline's "
                     "snapshot comes from the bench's own run of the same column, about 330 substeps in")


def host_cpu():
    """The host's CPU model and core counts (a reported baseline's provenance)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(model=model, logical_cpus=os.cpu_count(), usable_cpus=usable)


def counter_passes(N, checkpoint_path):
    """HBM traffic (FETCH_SIZE, WRITE_SIZE, calibrated on the library's copy kernel of
    a known size), VALU issue and texture-address-unit busy fractions and the cache
    lines per wave-wide load of the gather kernels, from rocprofv3 --pmc passes of
    one substep of the settled column: each pass loads the checkpoint written after
    settling (xpbi.run_scene --load) and runs one counted substep.  Passes of a few
    substeps profile cleanly; runs of hundreds of substeps under --pmc have ended in
    a segmentation fault inside the profiler's dispatch interception."""
    import shutil
    import subprocess
    from xpbi import counters
    if shutil.which("rocprofv3") is None:
        return None
    prog = [sys.executable, "-m", "xpbi.run_scene", "sand_column_1m", "--load", checkpoint_path, "--substeps", "1",
            "--copy-mib", "256", "--every", "0"]
    try:
        d, _ = counters.collect(prog, "k_solve|k_computeL|k_xsph|k_updateF", copy_bytes=256 << 20)
    except (RuntimeError, OSError, IndexError, KeyError, ZeroDivisionError, subprocess.TimeoutExpired) as e:
        print(f"[bench] counter passes failed: {str(e)[:300]}", file=sys.stderr, flush=True)
        return None
    if d["fetch_scale"] is None:
        return None
    for k, e in d["kernels"].items():
        if "traffic_bytes" in e:
            e["traffic_bytes_per_particle"] = round(e["traffic_bytes"] / N, 1)
    return d


def cpu_baseline(sim, scene, iterations, threads):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py
    from xpbi import pack_materials
    oracle_py.build()
    st = oracle_py.State(sim.N, sim.numCells)
    st.pos[:] = sim.read("POS")
    st.velA[:] = sim.read("VEL")
    st.F[:] = sim.read("F")
    st.aux[:] = sim.read("AUX")
    mats = pack_materials(scene["materials"], sim.V0)
    t0 = time.perf_counter()
    oracle_py.substep(st, sim.params, mats, iterations, scene["hasWater"], nthreads=threads)
    dt = time.perf_counter() - t0
    cpu = host_cpu()
    return dict(value=sim.N / dt, unit="particle-substeps/s", cores=threads, kind="port",
                sample=f"1 substep ({iterations} Jacobi iterations) of the settled {sim.N}-particle column, "
                       f"f32 C++ oracle, OpenMP {threads} threads, {dt:.1f} s",
                cpu_model=cpu["model"], host_logical_cpus=cpu["logical_cpus"], host_usable_cpus=cpu["usable_cpus"])


def column_2d(with_cpu):
    """The prototype's 2D fp64 column (40x80 at r = 0.01 m): its 300
    substeps in one launch of the 2D engine; beside it the numpy oracle's substep."""
    from xpbi import proto2d
    p, x = proto2d.column40x80_params(), proto2d.column40x80_positions()
    warm = proto2d.Prototype2D(x, p)
    warm.step(5)
    warm.close()
    sim = proto2d.Prototype2D(x, p)
    ms = sim.time(300)
    N = sim.N
    sim.close()
    out = dict(workload="2D DP column 40x80 r=0.01 (3200 particles), 20 iterations, dt 1/3000, fp64",
               substeps=300, ms_per_substep=round(ms / 300, 4), particle_substeps_per_s=round(N * 300 / (ms / 1e3), 1),
               launch="cooperative grid, one thread per particle, all 300 substeps in one launch", cpu_baseline=None)
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import proto2d as o2
        prm, st = o2.column40x80_params(), o2.column40x80_state()
        t0 = time.perf_counter()
        o2.run(prm, st, 20)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=round(N * 20 / dt, 1), unit="particle-substeps/s", cores=1, kind="port",
                                   sample=f"the first 20 of the 300 substeps, numpy float64 oracle, {dt:.1f} s")
    # the prototype script's own scene (666 particles, its constants) on the engine,
    # beside the script itself, timed when its golden vectors were recorded
    pp, px = proto2d.prototype_params(), proto2d.prototype_positions()
    sim = proto2d.Prototype2D(px, pp)
    pms = sim.time(150)
    sim.close()
    proto = dict(workload="the prototype script's own 666-particle column and constants (dx 0.008, 8 iterations, "
                          "dt 2e-4)", engine_ms_per_step=round(pms / 150, 4),
                 engine_particle_steps_per_s=round(len(px) * 150 / (pms / 1e3), 1))
    golden = os.path.join(REPO, "tests", "golden", "proto2d_golden.npz")
    if os.path.exists(golden):
        import numpy as np
        ms_ref = float(np.load(golden)["ms_per_step"])
        proto["reference_script"] = dict(
            ms_per_step=round(ms_ref, 2), particle_steps_per_s=round(len(px) / (ms_ref / 1e3), 1), cores=1,
            kind="reference", where="the unmodified prototype/xpbi2d.py (numpy + scipy) on the build container's CPU "
                                    "(8 vCPU Xeon 2.7 GHz), steps 3-20, recorded with its golden vectors")
    out["prototype_scene"] = proto
    return out


def mixed_scene(settle, steps, s_corr=None):
    """Dam breach: 600k PBF water beside a 400k DP sand wall (phi 35), 10 iterations;
    s_corr: override of the configured k = 0.1."""
    import xpbi
    from xpbi import scenes
    sc = scenes.dam_breach_1m()
    if s_corr is not None:
        sc["sCorrK"] = s_corr
    I = sc["iterations"]
    sim = xpbi.Simulation(sc, device=0)
    sim.writeParams(xpbi.params_from_scene(sc))
    sim.step(settle, I, True)
    sim.graph_step(2, I, True)
    sim.sync()
    t0 = time.perf_counter()
    sim.graph_step(steps, I, True)
    sim.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    total, per = sim.time_substeps_detail(4, I, True)
    N = sim.N
    kern = {k: round(v[0] / max(v[1], 1) * 1e3, 1) for k, v in per.items()}
    g = ("computeL", "solveA", "solveB", "pbfA", "pbfB", "xsph", "updateF")
    g_ms = sum(per[k][0] for k in g if k in per) / 4
    g_bytes = (264 + 348 * I) * N  # compulsory bytes of the gather kernels with water (upper bound)
    vel = sim.read("VEL")[:, :3]
    import numpy as np
    speed = np.linalg.norm(vel, axis=1)
    vmax = 0.45 * 2 * sc["h"] / sc["dt"]  # integrate's velocity clamp (simkernels.js:535)
    at_clamp = float((speed > 0.99 * vmax).mean())
    state = (f"burst: {100 * at_clamp:.0f} % of particles at the {vmax:.0f} m/s velocity clamp (s_corr k = "
             f"{sc['sCorrK']:g}); PBF cost here is that of clumped cells, not of a stable dam breach"
             if at_clamp > 0.05 else f"no burst: {100 * at_clamp:.1f} % at the velocity clamp (s_corr k = "
             f"{sc['sCorrK']:g}), median speed {float(np.median(speed)):.2f} m/s")
    out = dict(particles=N, water=int((sc["matIds"] == 1).sum()), sand=int((sc["matIds"] == 0).sum()),
               s_corr_k=sc["sCorrK"], state=state,
               iterations=I, settle_substeps=settle, ms_per_substep_graph=round(ms, 4),
               particle_substeps_per_s=round(N / (ms / 1e3), 1), kernels_us=kern,
               gather_kernels_GBs=round(g_bytes / (g_ms / 1e3) / 1e9, 1),
               gather_kernels_frac=round(g_bytes / (g_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4))
    sim.close()
    return out


def poke_scene(params, train, settle, substeps, cap):
    """A von Mises block 100^3 (sigmaY 5e3, E 2e5) poked by a 0.05 m sphere moving
    down at 1 m/s (its position updated every substep): learned vs plain Jacobi to 1e-4."""
    import xpbi
    from xpbi import scenes
    sc = scenes.vm_poke_1m()
    sim = xpbi.Simulation(sc, device=0)
    col = sc["collider"]
    dt = sc["dt"]
    state = dict(t=0.0)

    def place(k, t0=None):
        t = state["t"] if t0 is None else t0
        pos = [col["pos"][j] + col["vel"][j] * t for j in range(3)]
        sim.writeParams(xpbi.params_from_scene(sc, collider=dict(pos=pos, R=col["R"], vel=col["vel"])))

    for k in range(settle):
        place(k)
        sim.substep(sc["iterations"], False)
        state["t"] += dt
    t_start = state["t"]

    def before(k):
        place(k, t_start + k * dt)

    out = learned_vs_plain(sim, 1e-4, cap, substeps, params, train, before)
    out.update(particles=sim.N, settle_substeps=settle, collider=dict(R=col["R"], speed_m_s=1.0),
               note="collider moved every substep; learned iterations count the learned first iteration; "
                    "network trained on the two DP sand columns and a 125k von Mises poke at r = 0.01 (another "
                    "resolution, size and seed than this block)")
    out.pop("training", None)
    sim.close()
    return out


def ensemble(I, settle, steps):
    """The 64-scene batched ensemble (2k-60k particles each, five material models)."""
    import numpy as np
    import xpbi
    from xpbi import scenes
    sl = scenes.ensemble_64()
    sim = xpbi.Simulation(sl, device=0)
    for q, sc in enumerate(sl):
        sim.writeParams(xpbi.params_from_scene(sc), scene=q)
    sim.step(settle, I, False)
    blocks, tune_ms = sim.autotune(I, False, substeps=2)
    eager_ms = sim.time_substeps(4, I, False) / 4
    sim.graph_step(2, I, False)
    sim.sync()
    t0 = time.perf_counter()
    sim.graph_step(steps, I, False)
    sim.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    sizes = [si.N for si in sim.scenes_info]
    kinds = np.bincount([sc["materials"][0]["kind"] for sc in sl], minlength=7)
    out = dict(scenes=len(sl), particles=sim.N, smallest=min(sizes), largest=max(sizes),
               materials=dict(DP=int(kinds[1]), NACC=int(kinds[2]), VM=int(kinds[3]), HB=int(kinds[4]),
                              SNOW=int(kinds[5])),
               iterations=I, settle_substeps=settle, steps=steps, launch_blocks=blocks,
               launch="one block per 256-particle run" if blocks == 0 else f"persistent grid of {blocks} blocks",
               work_entries=sim.get_launch()[1], ms_per_substep_graph=round(ms, 4),
               ms_per_substep_eager=round(eager_ms, 4), particle_substeps_per_s=round(sim.N / (ms / 1e3), 1))
    sim.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--settle", type=int, default=300)
    ap.add_argument("--residual-cap", type=int, default=200)
    ap.add_argument("--weights", default=None, help="warm-start network (.npy) instead of training in the run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-residual", action="store_true")
    ap.add_argument("--config", default="column1m", choices=["column1m", "column128k"])
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the mixed (c3), poke (c5) and 2D (c1) scenes")
    ap.add_argument("--only-2d", action="store_true", help="run only the 2D 40x80 fp64 column; print its object")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--eager", action="store_true",
                    help="time the fixed-iteration substeps as eager launches (no autotune, no HIP graph)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: a simulation runs on one device and the engine has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.only_2d:
        print(json.dumps(column_2d(not args.no_cpu)))
        return

    import xpbi
    from xpbi import scenes
    scene = scenes.sand_column_1m() if args.config == "column1m" else scenes.sand_column_128k()
    I = args.iterations
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    N = sim.N
    print(f"[bench] {N} particles, {sim.numCells} cells; settling {args.settle} substeps", file=sys.stderr,
          flush=True)
    sim.step(args.settle, I, False)
    ckpt = None
    if not args.no_counters and args.config == "column1m":
        import tempfile
        from xpbi import checkpoint
        ckpt = os.path.join(tempfile.mkdtemp(prefix="xpbi_bench_"), "settled.npz")
        checkpoint.save(sim, ckpt, args.settle)
    if args.eager:
        blocks, advance = 0, sim.step
    else:
        blocks, tune_ms = sim.autotune(I, False, substeps=2)
        print(f"[bench] autotuned launch: {blocks} blocks (0 = one per run), {tune_ms:.3f} ms/substep",
              file=sys.stderr, flush=True)
        advance = sim.graph_step
    advance(args.warmup, I, False)
    sim.sync()
    print("[bench] timed region", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    advance(args.steps, I, False)
    sim.sync()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    value = N / (ms_per_step / 1e3)
    # per-kernel HIP-event times of an eager run of the same substep (a separate pass)
    detail_steps = 10
    total_ms, per = sim.time_substeps_detail(detail_steps, I, False)
    kern = {}
    for name, (ms, cnt) in per.items():
        avg = ms / max(cnt, 1)
        ent = dict(avg_ms=round(avg, 5), launches=cnt, share=round(ms / total_ms, 4))
        if name in ALG_BYTES:
            ent["alg_GBs"] = round(ALG_BYTES[name] * N / (avg / 1e3) / 1e9, 1)
        kern[name] = ent
    # the roofline kernel is the one with the largest share of the substep (solveB
    # on this workload); solveA, the other per-iteration sweep, beside it
    shares = {k: v[0] for k, v in per.items() if k in ALG_BYTES}
    top = max(shares, key=shares.get)

    def roof(k):
        k_ms = per[k][0] / per[k][1]
        ach = ALG_BYTES[k] * N / (k_ms / 1e3) / 1e9
        return dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=k,
                    share=round(per[k][0] / total_ms, 4), alg_bytes_per_particle=ALG_BYTES[k],
                    alg_bytes_per_launch=ALG_BYTES[k] * N, avg_us=round(k_ms * 1e3, 2))
    roofline = roof(top)
    roofline_other = roof("solveA" if top != "solveA" else "solveB")
    g_ms = sum(per[k][0] for k in GATHER if k in per)
    g_bytes = (264 + 260 * I) * N * detail_steps
    gather = dict(achieved_GBs=round(g_bytes / (g_ms / 1e3) / 1e9, 1),
                  frac=round(g_bytes / (g_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4))

    fixed_timing = ("eager substeps, one block per 256-particle run" if args.eager
                    else "HIP-graph replay of two captured substeps, autotuned launch shape")
    fixed = dict(value=round(value, 1), ms_per_step=round(ms_per_step, 4), steps=args.steps, iterations=I,
                 timing=fixed_timing, launch_blocks=blocks)
    resid, net, train = None, None, None
    if not args.no_residual:
        print("[bench] Jacobi to residual 1e-4: plain vs learned warm start", file=sys.stderr, flush=True)
        net, train = trained_network(args.weights, log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
        print("[bench] timed region: plain, then learned, from one snapshot", file=sys.stderr, flush=True)
        resid = learned_vs_plain(sim, 1e-4, args.residual_cap, args.steps, net, train, warmup=args.warmup)
        ms_per_step = resid["learned_ms_per_substep"]
        value = N / (ms_per_step / 1e3)

    copy_bytes = 1 << 30
    copy_ms = sim.copy_bandwidth(copy_bytes, reps=5)
    copy_GBs = 2 * copy_bytes / (copy_ms / 1e3) / 1e9  # read + write

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        print(f"[bench] CPU oracle baseline on {threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sim, scene, I, threads)
    sim.close()
    if resid is not None and args.config == "column1m":
        print("[bench] learned vs plain along the collapse (a fresh column)", file=sys.stderr, flush=True)
        resid["along_collapse"] = along_collapse(scene, net)

    counters = None
    if ckpt:
        print("[bench] counter passes (rocprofv3 --pmc), one substep each from the settled column",
              file=sys.stderr, flush=True)
        counters = counter_passes(N, ckpt)
        import shutil
        shutil.rmtree(os.path.dirname(ckpt), ignore_errors=True)
        if counters:
            for r in (roofline, roofline_other):
                c = counters["kernels"].get("k_" + r["kernel"])
                if c:
                    r["traffic"] = c["traffic_bytes"]
                    r["traffic_source"] = ("rocprofv3 FETCH_SIZE + WRITE_SIZE of one substep of the same settled "
                                           "column (xpbi.run_scene from a checkpoint of this run), calibrated on "
                                           "a 256 MiB copy")
                    for key in ("valu_busy", "ta_busy", "ta_cycles_per_load", "lines_per_load"):
                        if key in c:
                            r[key] = c[key]

    ens = None
    if not args.no_ensemble:
        print("[bench] 64-scene batched ensemble", file=sys.stderr, flush=True)
        ens = ensemble(I, 300, 20)
    mixed = mixed_calm = poke = c1 = None
    if not args.no_configs:
        print("[bench] 2D 40x80 column on the fp64 engine", file=sys.stderr, flush=True)
        c1 = column_2d(not args.no_cpu)
        print("[bench] mixed water + sand scene", file=sys.stderr, flush=True)
        mixed = mixed_scene(60, 10)
        print("[bench] mixed water + sand scene without s_corr", file=sys.stderr, flush=True)
        mixed_calm = mixed_scene(60, 10, s_corr=0.0)
        if net is not None:
            print("[bench] von Mises block + poke: plain vs learned", file=sys.stderr, flush=True)
            poke = poke_scene(net, train, 200, 20, args.residual_cap)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"DP sand column 80x160x80 r=0.005 ({N} particles)" if args.config == "column1m"
                                else f"DP sand column 40x80x40 r=0.01 ({N} particles)")
                   + (", Jacobi to residual 1e-4 from the learned warm start" if resid else
                      f", {I} Jacobi iterations (no residual target: --no-residual)"),
                   "particles": N, "dt": scene["dt"], "settle_substeps": args.settle,
                   "residual": 1e-4 if resid else None, "iteration_cap": args.residual_cap if resid else I,
                   "timing": ("eager substeps; iterations stop on the device-side residual test, which the "
                              "host reads every iteration from pinned memory without stalling the stream, staying at most 2 "
                              "iterations ahead; no host read between "
                              "timed substeps (the iteration counts come from an identical untimed pass); plain "
                              "and learned from the same snapshot" if resid else fixed_timing),
                   "launch_blocks": blocks},
        "plain_value": resid["plain_particle_substeps_per_s"] if resid else None,
        "learned_vs_plain_time_ratio": resid["time_ratio"] if resid else None,
        "fixed_iterations": fixed,
        "eager_ms_per_step": round(total_ms / detail_steps, 4),
        "roofline": roofline, "roofline_solveA" if roofline_other["kernel"] == "solveA" else "roofline_solveB":
        roofline_other, "gather_kernels_roofline": gather, "kernels": kern, "counters": counters,
        "stream_copy_GBs": round(copy_GBs, 1),
        "jacobi_to_residual": resid, "batched_ensemble": ens, "mixed_water_sand": mixed,
        "mixed_water_sand_no_s_corr": mixed_calm,
        "vm_block_poke_to_residual": poke, "column_2d": c1, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
