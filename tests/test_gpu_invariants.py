"""Whole-substep invariants on the device, through the C ABI (the same ones
tests/test_oracle_substep.py holds the oracle to): a rest lattice stays exactly at
rest; free fall and a rigid translation move every particle alike with F = I and no
constraint; a substep away from walls keeps the total momentum."""
import numpy as np
import pytest

import xpbi
from xpbi import scenes
from helpers import F3

pytestmark = pytest.mark.gpu

ID9 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)


def _blob(jitter, lift=0.0, gravity=0.0, velocities=None):
    scene = scenes.sand_column(10, 10, 10, 0.01, [0.6, 0.6, 0.6], seed=3, iterations=20, jitter=jitter)
    scene = dict(scene, positions=scene["positions"] + np.array([0, 0.2 + lift, 0], np.float32),
                 gravity=[0.0, gravity, 0.0])
    if velocities is not None:
        scene["velocities"] = velocities.astype(np.float32)
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    return scene, sim


def _by_id(sim, a):
    out = np.empty_like(a)
    out[sim.read("ID")] = a
    return out


def test_rest_lattice_stays_at_rest():
    scene, sim = _blob(0.0)
    x0 = sim.read("POS")
    sim.step(3, 20, False)
    assert np.array_equal(_by_id(sim, sim.read("POS")), x0)
    assert not sim.read("VEL")[:, :3].any()
    assert np.array_equal(F3(sim.read("F")), np.broadcast_to(ID9, (sim.N, 9)))
    assert not sim.read("AUX")[:, 1].any()
    sim.close()


def test_free_fall_and_translation_are_rigid():
    scene, sim = _blob(0.1, lift=0.1, gravity=-9.81)
    x0 = sim.read("POS")[:, :3].astype(np.float64)
    sim.step(6, 20, False)
    vel = sim.read("VEL")[:, :3]
    assert np.array_equal(vel, np.broadcast_to(vel[0], vel.shape))
    dt, g = np.float32(scene["dt"]), np.float32(-9.81)
    v = np.float32(0)
    for _ in range(6):
        v = v + dt * g
    assert vel[0, 0] == 0 and vel[0, 2] == 0 and abs(vel[0, 1] - v) <= 4 * np.finfo(np.float32).eps * abs(v)
    assert np.array_equal(F3(sim.read("F")), np.broadcast_to(ID9, (sim.N, 9)))
    d = _by_id(sim, sim.read("POS"))[:, :3] - x0
    assert np.abs(d - d[0]).max() <= 1e-6 * np.abs(d).max()
    sim.close()
    v0 = np.array([0.3, -0.1, 0.2], np.float32)
    scene, sim = _blob(0.1, velocities=np.tile(v0, (1000, 1)))
    sim.step(2, 20, False)
    assert np.array_equal(sim.read("VEL")[:, :3], np.broadcast_to(v0, (sim.N, 3)))
    assert np.array_equal(F3(sim.read("F")), np.broadcast_to(ID9, (sim.N, 9)))
    assert not sim.read("AUX")[:, 1].any()
    sim.close()


def test_substep_conserves_momentum():
    vel = np.random.default_rng(5).normal(0, 0.05, (1000, 3))
    scene, sim = _blob(0.1, velocities=vel)
    v = sim.read("VEL")[:, :3].astype(np.float64)
    p0, scale = v.sum(0), np.abs(v).sum()
    sim.step(2, 20, False)
    assert np.abs(sim.read("AUX")[:, 1]).max() > 0, "no constraint acted"
    p1 = sim.read("VEL")[:, :3].astype(np.float64).sum(0)
    print(f"momentum drift {np.abs(p1 - p0).max() / scale:.2e} of sum |m v|", flush=True)
    assert np.abs(p1 - p0).max() <= 1e-5 * scale
    sim.close()
