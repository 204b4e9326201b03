"""HIP substep vs the CPU oracle, through the C ABI.

Bar: the counting sort is bit-exact (cell starts, permutation, ids); continuous
state after one substep from a shared snapshot matches the f32 oracle within
rel 1e-4 on positions and F (per particle, relative to the particle's norm).
Near-zero-strain particles make the solve ill-conditioned in f32 (a 1-ulp change
of the inputs moves any f32 result past 1e-4 for ~0.1% of particles on soft
scenes), so every check is made against the f32 oracle's own distance to the f64
oracle on the same snapshot (see check_field / check_vel).
"""
import numpy as np
import pytest

import xpbi
from xpbi import scenes
from helpers import F3, mats_record, rel_err_rows

pytestmark = pytest.mark.gpu

POS_TOL = 1e-4
F_TOL = 1e-4


def make_sim(scene, dt=None):
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene, dt=dt))
    return sim


def snapshot(sim, oracle, dtype=np.float32):
    st = oracle.State(sim.N, sim.numCells, dtype)
    st.pos[:] = sim.read("POS")
    st.velA[:] = sim.read("VEL")
    st.F[:] = sim.read("F")
    st.aux[:] = sim.read("AUX")
    st.id[:] = sim.read("ID")
    return st


def _rows(a, b, floor):
    d = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=1)
    return d / np.maximum(np.linalg.norm(np.asarray(b, np.float64), axis=1), floor)


def check_field(name, gpu, o32s, o64, floor, tol, label):
    """GPU vs the f32 oracle (rel per particle) with the f32 conditioning floor taken into account.

    o32s holds the f32 oracle on the snapshot and on 1-ulp perturbations of it; per
    particle, the f32 spread is the largest distance of any of them to the f64 oracle.

    * the GPU is no farther from the f64 oracle than 3x the f32 spread (or tol);
    * when the f32 spread is within tol everywhere, every particle of the GPU is within
      tol of the f32 oracle;
    * otherwise (ill-conditioned particles, where 1 ulp of input moves any f32 result past
      tol) the GPU has at most 3x as many particles past tol as the f32 spread has.
    """
    e = _rows(gpu, o32s[0], floor)
    e64 = _rows(gpu, o64, floor)
    n64 = np.max([_rows(o, o64, floor) for o in o32s], axis=0)
    bad, bad_o = int((e > tol).sum()), int((n64 > tol).sum())
    print(f"  {name}: gpu-f32 max {e.max(initial=0):.2e} (>tol: {bad}) | gpu-f64 max {e64.max(initial=0):.2e} "
          f"| f32-f64 max {n64.max(initial=0):.2e} (>tol: {bad_o})", flush=True)
    assert e64.max(initial=0) <= max(tol, 3.0 * n64.max(initial=0)), f"{label}: {name} farther from f64 than f32"
    if bad_o == 0:
        assert bad == 0, f"{label}: {name} rel err {e.max():.3e}"
    else:
        assert bad <= 3 * bad_o, f"{label}: {name} {bad} particles past tol (f32 oracle: {bad_o})"
    return float(e.max(initial=0))


def check_vel(vel, o32s, o64, vs, label):
    """Velocities relative to the largest speed: near-zero-strain particles make single
    velocities ill-conditioned in f32, so the bar is statistical: the GPU's 99.9th
    percentile distance to the f64 oracle is within 3x the f32 spread's (or 1e-4), and its
    largest distance within 10x the f32 spread's largest (or 1e-3).  o32s as in check_field."""
    e64 = np.abs(vel - o64).max(axis=1) / vs
    n64 = np.max([np.abs(o - o64).max(axis=1) / vs for o in o32s], axis=0)
    e = np.abs(vel - o32s[0]).max(axis=1) / vs
    q, qn = float(np.quantile(e64, 0.999)), float(np.quantile(n64, 0.999))
    print(f"  vel: gpu-f32 max {e.max():.2e} | gpu-f64 q999 {q:.2e} max {e64.max():.2e} | "
          f"f32-f64 q999 {qn:.2e} max {n64.max():.2e}", flush=True)
    assert q <= max(1e-4, 3.0 * qn), f"{label}: velocity q999 {q:.3e} vs f32 oracle {qn:.3e}"
    assert e64.max() <= max(1e-3, 10.0 * n64.max()), f"{label}: velocity max {e64.max():.3e}"
    return float(e.max())


def compare(sim, refs, ref64, label, vel_scale=None):
    """refs: the f32 oracle on the snapshot, then on its 1-ulp perturbations."""
    ref = refs[0]
    pos, vel, F = sim.read("POS"), sim.read("VEL"), sim.read("F")
    assert np.array_equal(sim.read("ID"), ref.id), f"{label}: id channel differs"
    print(f"{label}: N={sim.N}", flush=True)
    ep = check_field("pos", pos[:, :3], [r.pos[:, :3] for r in refs], ref64.pos[:, :3], 1e-6, POS_TOL, label)
    eF = check_field("F", F3(F), [F3(r.F) for r in refs], F3(ref64.F), 1.0, F_TOL, label)
    vs = vel_scale or max(float(np.abs(ref.velA[:, :3]).max()), 1e-3)
    ev = check_vel(vel[:, :3].astype(np.float64), [r.velA[:, :3].astype(np.float64) for r in refs],
                   ref64.velA[:, :3], vs, label)
    return ep, eF, ev


def perturbed(st, seed):
    """The snapshot with velocity and F moved by one ulp (random sign per entry); positions,
    and so the sort, are unchanged."""
    rng = np.random.default_rng(seed)
    p = st.copy()
    for a in (p.velA, p.F):
        step = rng.choice(np.array([-1, 1], np.int8), size=a.shape)
        a[:] = np.where(step > 0, np.nextafter(a, np.float32(np.inf)), np.nextafter(a, np.float32(-np.inf)))
    return p


def deformed(F12, aux, tags, scene):
    """True when some solid particle has F != I (the snapshot exercises the constraint path)."""
    kinds = np.array([m["kind"] for m in scene["materials"]])
    solid = kinds[tags.astype(np.int64) & 7] != 6
    dev = np.abs(F3(F12) - np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)).max(axis=1)
    return bool(solid.any() and dev[solid].max() > 1e-4)


def one_substep_parity(sim, scene, oracle, label, iters=None, water=None, require_deformed=True):
    iters = scene["iterations"] if iters is None else iters
    water = scene["hasWater"] if water is None else water
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    variants = [st32] + [perturbed(st32, k) for k in (1, 2)]
    mats = mats_record(scene)
    sim.substep(iters, water)
    for st in variants + [st64]:
        oracle.substep(st, sim.params, mats, iters, water)
    assert np.array_equal(sim.read("PERM"), st32.sortedIdx), f"{label}: permutation differs"
    assert np.array_equal(sim.read("CELLSTART"), st32.cellTab), f"{label}: cell starts differ"
    if require_deformed:
        assert deformed(st32.F, st32.aux, st32.pos[:, 3], scene), f"{label}: snapshot is undeformed"
    return compare(sim, variants, st64, label)


def test_sort_bit_exact_128k_column(oracle):
    """128k-particle DP column: cell starts and permutation equal the oracle and numpy's stable sort."""
    scene = scenes.sand_column_128k()
    sim = make_sim(scene)
    ref = oracle.initial_state(scene, sim.numCells)
    print("sort, 128k column: N", sim.N, "cells", sim.numCells, flush=True)
    sim.substep_partial(0, False, "SORT")
    oracle.substep(ref, sim.params, mats_record(scene), 0, False, stop="SORT")
    perm, cells = sim.read("PERM"), sim.read("CELLSTART")
    assert np.array_equal(cells, ref.cellTab)
    assert np.array_equal(perm, ref.sortedIdx)
    keys = oracle.cell_index(sim.params, scene["positions"])
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))
    counts = np.bincount(keys, minlength=sim.numCells)
    assert np.array_equal(cells, (np.cumsum(counts) - counts).astype(np.uint32))
    sim.close()


def test_sort_after_motion_and_reorder(oracle):
    """Second and later substeps sort an already cell-ordered array: still exact, ids follow."""
    scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=11, iterations=5)
    sim = make_sim(scene)
    sim.step(40, 5, False)
    st = snapshot(sim, oracle)
    sim.substep_partial(5, False, "REORDER")
    oracle.substep(st, sim.params, mats_record(scene), 5, False, stop="REORDER")
    assert np.array_equal(sim.read("PERM"), st.sortedIdx)
    assert np.array_equal(sim.read("CELLSTART"), st.cellTab)
    assert np.array_equal(sim.read("ID"), st.id)
    assert np.array_equal(sim.read("POS"), st.pos)
    np.testing.assert_array_equal(sim.read("F"), st.F)
    # scatter work: V^n = V0 clamp(det F), v += dt g, lambda = 0
    np.testing.assert_allclose(sim.read("AUX"), st.aux, rtol=2e-7, atol=0)
    np.testing.assert_allclose(sim.read("VEL"), st.velA, rtol=1e-6, atol=1e-9)
    sim.close()


def test_sort_crowded_cells_in_random_order(oracle):
    """Cells holding thousands of particles, the array in no cell order at all: the
    count's slots and the in-cell rank still give numpy's stable sort."""
    scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=4, iterations=5)
    rng = np.random.default_rng(4)
    n = scene["positions"].shape[0]
    lo, hi = np.array([0.1, 0.1, 0.1]), np.array(scene["domainMax"], np.float64) - 0.1
    pos = lo + rng.random((n, 3)) * (hi - lo)
    clump = rng.random(n) < 0.4  # 40 % in one tiny box, within a single cell
    pos[clump] = np.array([0.305, 0.205, 0.305]) + rng.random((int(clump.sum()), 3)) * 0.002
    second = (~clump) & (rng.random(n) < 0.2)  # and 20 % of the rest in another
    pos[second] = np.array([0.155, 0.405, 0.455]) + rng.random((int(second.sum()), 3)) * 0.002
    scene["positions"] = pos.astype(np.float32)
    sim = make_sim(scene)
    keys = oracle.cell_index(sim.params, scene["positions"])
    counts = np.bincount(keys, minlength=sim.numCells)
    print("sort, crowded cells: N", n, "largest cells", np.sort(counts)[-2:], flush=True)
    assert counts.max() >= 0.4 * n - 200
    sim.substep_partial(0, False, "SORT")
    assert np.array_equal(sim.read("PERM"), np.argsort(keys, kind="stable").astype(np.uint32))
    assert np.array_equal(sim.read("CELLSTART"), (np.cumsum(counts) - counts).astype(np.uint32))
    sim.close()


def test_staged_parity_sand(oracle):
    """computeL, solve, xsph, updateF stages of a landed sand column against the f32 oracle."""
    scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=3, iterations=20)
    sim = make_sim(scene)
    sim.step(260, 20, False)  # land on the floor: nonzero strain and plastic flow
    base = snapshot(sim, oracle)
    mats = mats_record(scene)
    for stage in ("COMPUTE_L", "SOLVE", "XSPH", "UPDATE_F"):
        st = base.copy()
        sim.write("POS", base.pos)
        sim.write("VEL", base.velA)
        sim.write("F", base.F)
        sim.write("AUX", base.aux)
        sim.write("ID", base.id)
        sim.substep_partial(20, False, stage)
        oracle.substep(st, sim.params, mats, 20, False, stop=stage)
        if stage == "COMPUTE_L":
            eL = rel_err_rows(F3(sim.read("L")), F3(st.L), floor=1.0)
            print(f"computeL: L rel err {eL:.2e}", flush=True)
            assert eL < 1e-4
        else:
            st64 = base.copy(np.float64)
            oracle.substep(st64, sim.params, mats, 20, False, stop=stage)
            vs = max(float(np.abs(st.velA[:, :3]).max()), 1e-3)
            eF = rel_err_rows(F3(sim.read("F")), F3(st.F), floor=1.0)
            print(f"{stage}: F {eF:.2e}", flush=True)
            check_vel(sim.read("VEL")[:, :3].astype(np.float64), [st.velA[:, :3].astype(np.float64)],
                      st64.velA[:, :3], vs, stage)
            assert eF < F_TOL
    sim.close()


def test_substep_parity_sand_column(oracle):
    """Full substeps of a DP column at several points of its collapse."""
    scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=5, iterations=20)
    sim = make_sim(scene)
    for k, n in enumerate((210, 150, 150)):
        sim.step(n, 20, False)
        one_substep_parity(sim, scene, oracle, f"sand column after {n} more substeps")
    sim.close()


def test_substep_parity_headline_column(oracle):
    """The benchmark's 1,024,000-particle DP column at full size, landed (300 substeps):
    the sort bit-exact and one substep (20 iterations) within the bar, against the f32
    oracle, its 1-ulp perturbations and the f64 oracle (OpenMP, a few seconds each)."""
    scene = scenes.sand_column_1m()
    sim = make_sim(scene)
    print("headline column: settling 300 substeps", flush=True)
    sim.step(300, 20, False)
    one_substep_parity(sim, scene, oracle, "1M DP column after 300 substeps")
    sim.close()


def test_substep_parity_vm_poke_1m(oracle):
    """The 1M von Mises block (100^3 at r = 0.005) at full size. The 0.05 m sphere
    moves down at 1 m/s and its position is rewritten every substep; it starts 4
    spacings above the falling block and after 400 substeps presses into it (960
    particles within 2h of its surface and 190 inside it before the substep, against
    418 and none after 200 substeps). One substep there must meet the bar, with the
    sphere's push in the boundary pass."""
    scene = scenes.vm_poke_1m()
    sim = xpbi.Simulation(scene, device=0)
    col, dt = scene["collider"], scene["dt"]

    def place(k):
        pos = [col["pos"][j] + col["vel"][j] * k * dt for j in range(3)]
        sim.writeParams(xpbi.params_from_scene(scene, collider=dict(pos=pos, R=col["R"], vel=col["vel"])))

    steps = 400
    print(f"von Mises poke 1M: {steps} substeps with the sphere moving", flush=True)
    for k in range(steps):
        place(k)
        sim.substep(scene["iterations"], False)
    place(steps)
    x = sim.read("POS")[:, :3]
    c = np.asarray(col["pos"]) + np.asarray(col["vel"]) * steps * dt
    d = np.linalg.norm(x - c, axis=1)
    near = int((d < col["R"] + 2 * scene["h"]).sum())
    print(f"particles within 2h of the sphere before the checked substep: {near}", flush=True)
    inside = int((d < col["R"]).sum())
    print(f"particles inside the sphere before the checked substep: {inside}", flush=True)
    assert near > 800 and inside > 100, "the sphere does not press into the block"
    one_substep_parity(sim, scene, oracle, f"1M von Mises poke after {steps} substeps", water=False)
    sim.close()


@pytest.mark.parametrize("name", list(scenes.SCENES))
def test_substep_parity_reference_scenes(oracle, name):
    """Every material model of the reference (DP, NACC, VM, HB, snow, water+sand, elastic)."""
    scene = scenes.SCENES[name](seed=1)
    sim = make_sim(scene)
    sim.step(40 * scene["substeps"], scene["iterations"], scene["hasWater"])  # 40 frames: contact and flow
    one_substep_parity(sim, scene, oracle, name)
    sim.close()


def test_residual_monitor_matches_oracle(oracle):
    scene = scenes.sand_column(12, 24, 12, 0.01, [0.5, 0.4, 0.5], seed=9, iterations=20)
    sim = make_sim(scene)
    sim.step(220, 20, False)
    st = snapshot(sim, oracle)
    sim.set_residual_monitor(True)
    sim.substep(20, False)
    g = sim.read_residuals()
    r = np.zeros(40)
    oracle.substep(st, sim.params, mats_record(scene), 20, False, resid=r)
    r = r.reshape(-1, 2)
    assert g.shape == (20, 2)
    np.testing.assert_allclose(g[:, 1], r[:, 1], rtol=1e-3)
    rg = np.sqrt(g[:, 0] / g[0, 1])
    ro = np.sqrt(r[:, 0] / r[0, 1])
    print("residual gpu   ", np.array2string(rg, precision=3), flush=True)
    print("residual oracle", np.array2string(ro, precision=3), flush=True)
    # the f32 residual has a noise floor near 3e-5 (C is evaluated from F ~ I, whose f32
    # entries carry ~6e-8 absolute error against strains ~1e-3) and single-iteration
    # spikes where a particle crosses a branch (contact on/off, the C < 1e-7 cut-off) in
    # either implementation; the curves are compared to 2% while r >= 1e-3, and over the
    # last five iterations the GPU's median is below 1e-4 or within 3x the oracle's
    hi = ro >= 1e-3
    np.testing.assert_allclose(rg[hi], ro[hi], rtol=2e-2)
    assert np.median(rg[-5:]) <= max(1e-4, 3.0 * np.median(ro[-5:]))
    assert rg[-5:].max() < 1e-3
    sim.close()


def test_error_paths():
    scene = scenes.sand_column(4, 4, 4, 0.01, [0.2, 0.2, 0.2], seed=1)
    sim = xpbi.Simulation(scene, device=0)
    with pytest.raises(xpbi.XpbiError, match="before xpbi_write_params"):
        sim.substep(1, False)
    p = xpbi.params_from_scene(scene)
    sim.writeParams(p)
    bad = xpbi._abi.Params.from_buffer_copy(sim.params)
    bad.numCells += 1
    import ctypes
    assert xpbi.lib().xpbi_write_params(sim._h, ctypes.byref(bad)) == xpbi._abi.ERR_ARG
    buf = np.zeros(3, np.float32)
    assert xpbi.lib().xpbi_read(sim._h, 0, buf.ctypes.data, buf.nbytes) == xpbi._abi.ERR_ARG
    assert xpbi.lib().xpbi_write(sim._h, xpbi.FIELD["L"], buf.ctypes.data, 48 * sim.N) == xpbi._abi.ERR_ARG
    assert xpbi.lib().xpbi_substep(sim._h, -1, 0) == xpbi._abi.ERR_ARG
    sim.substep(2, False)
    assert np.isfinite(sim.read("POS")).all()
    sim.close()


def test_empty_scene():
    scene = scenes.sand_column(4, 4, 4, 0.01, [0.2, 0.2, 0.2], seed=1)
    scene = dict(scene, positions=np.zeros((0, 3), np.float32), tags=np.zeros(0, np.float32),
                 matIds=np.zeros(0, np.uint32), states=np.zeros(0, np.float32))
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    sim.substep(3, False)
    assert sim.read("POS").shape == (0, 4)
    sim.close()
