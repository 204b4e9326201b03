"""Residual-driven Jacobi loop: the device-side stopping test and its bookkeeping.

A substep stopped at residual tol after n iterations must leave exactly the state
of the same substep run with a fixed count of n iterations (bit-exact: the same
kernels run on the same data; the stopped launches touch nothing and the
velocity ping-pong is restored), and its residual must be the first one <= tol.
"""
import numpy as np
import pytest

import xpbi
from xpbi import scenes

pytestmark = pytest.mark.gpu


def _sim(scene, settle):
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    sim.step(settle, scene["iterations"], scene["hasWater"])
    return sim


def _state(sim):
    return {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}


@pytest.mark.parametrize("check_every", [0, 1, 4])
def test_stop_equals_fixed_count(check_every):
    scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=5, iterations=20)
    a = _sim(scene, 220)
    b = _sim(scene, 220)
    snap = _state(a)
    for k, v in snap.items():
        b.write(k, v)
    seen = set()
    for tol in (3e-3, 1e-3, 5e-4, 3e-4, 2e-4):
        for k, v in snap.items():
            a.write(k, v)
            b.write(k, v)
        a.set_tolerance(tol, check_every)
        a.set_residual_monitor(True)
        a.substep(200, False)
        n, r = a.last_iterations()
        res = a.read_residuals()
        rel = np.sqrt(res[:, 0] / res[0, 1])
        print(f"tol {tol:g} check_every {check_every}: {n} iterations, residual {r:.3e}", flush=True)
        assert 0 < n < 200
        assert res.shape[0] == n
        assert rel[n - 1] <= tol and np.all(rel[: n - 1] > tol)
        assert abs(r - rel[n - 1]) <= 1e-6 * max(r, 1e-30)
        b.substep(n, False)
        for k in ("POS", "VEL", "F", "AUX", "ID"):
            assert np.array_equal(a.read(k), b.read(k)), f"{k} differs from {n} fixed iterations"
        seen.add(n % 2)
    if check_every == 0:
        assert seen == {0, 1}, "both ping-pong parities exercised"
    a.close()
    b.close()


def test_stop_with_water_and_cap():
    scene = scenes.SCENES["Dam breach: water + sand"](seed=1)
    a = _sim(scene, 20 * scene["substeps"])
    snap = _state(a)
    b = _sim(scene, 0)
    a.set_tolerance(1e-3, 0)
    a.substep(60, True)
    n, r = a.last_iterations()
    print(f"water scene: {n} iterations, residual {r:.3e}", flush=True)
    for k, v in snap.items():
        b.write(k, v)
    b.substep(n, True)
    for k in ("POS", "VEL", "F", "AUX"):
        assert np.array_equal(a.read(k), b.read(k)), k
    # a cap below convergence applies every iteration
    for k, v in snap.items():
        a.write(k, v)
    a.set_tolerance(1e-9, 2)
    a.substep(3, True)
    assert a.last_iterations()[0] == 3
    a.set_tolerance(0.0, 0)
    a.substep(5, True)
    assert a.last_iterations() == (5, -1.0)
    a.close()
    b.close()


def test_learned_residual_loop_on_a_persistent_grid():
    """The learned warm start presets the residual's denominator (sum C^2 at lambda = 0,
    reduced per block by the features pass).  On a persistent grid each thread walks
    several 256-particle runs, so its partial must accumulate over them: the loop then
    stops after the same iterations at the same residual (up to the order of the block
    sums) as with one block per run."""
    from test_gpu_warmstart import _net
    scene = scenes.sand_column(24, 48, 24, 0.01, [0.8, 0.6, 0.8], seed=11, iterations=20)
    a = _sim(scene, 260)
    snap = _state(a)
    b = _sim(scene, 0)
    entries = a.get_launch()[1]
    b.set_launch(8)
    assert b.get_launch() == (8, entries) and entries >= 32, "the persistent grid walks several runs per block"
    net = _net(3)
    for sim in (a, b):
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_warmstart(net)
        sim.set_tolerance(1e-4, 4)
    for step in range(3):
        a.substep(200, False)
        b.substep(200, False)
        (na, ra), (nb, rb) = a.last_iterations(), b.last_iterations()
        print(f"substep {step}: one block per run {na} iterations r {ra:.6e} | persistent {nb} r {rb:.6e}",
              flush=True)
        assert na == nb and 0 < na < 200
        assert abs(ra - rb) <= 1e-5 * ra
        assert ra <= 1e-4
    a.close()
    b.close()
