"""Whole-substep invariants of the CPU oracle (no device).

The oracle's solveA/solveB, XSPH, PBF and integrate restate web/js/shaders/
simkernels.js:248-542; no reference test vectors exist for them, so they are
pinned by what the mathematics fixes whatever the implementation:

* a rest lattice (F = I, v = 0, no gravity) has C = 0 everywhere and stays
  exactly at rest (simkernels.js:264-270: C < 1e-7 writes a zero Sp);
* a uniform velocity field (free fall, or a rigid translation) produces no
  velocity gradient, so F stays I, no constraint or XSPH impulse appears and every
  particle keeps the same velocity;
* the constraint gather is momentum-neutral (selfDv = -sum of the neighbours'
  impulses, simkernels.js:294-299), XSPH is antisymmetric for equal volumes and
  masses, and the contact correction is pairwise with weights m_b / (m_a + m_b):
  away from walls a substep keeps the total momentum;
* PBF water below its rest density has C = max(rho / rho0 - 1, 0) = 0 and so a
  zero multiplier (simkernels.js:418-425).

A plausible mistake (a dropped selfDv, a sign or index slip in the gather, a
wrong XSPH weight, gravity applied twice) breaks at least one of these.
"""
import numpy as np
import pytest

from xpbi import scenes
from helpers import F3, mats_record, params_record

ID9 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)


def _blob(oracle, jitter, lift=0.0, gravity=0.0, velocities=None, seed=3):
    """A 10x10x10 DP block at r = 0.01 in the middle of a 0.6 m box, away from every wall."""
    scene = scenes.sand_column(10, 10, 10, 0.01, [0.6, 0.6, 0.6], seed=seed, iterations=20, jitter=jitter)
    scene = dict(scene, positions=scene["positions"] + np.array([0, 0.2 + lift, 0], np.float32),
                 gravity=[0.0, gravity, 0.0])
    if velocities is not None:
        scene["velocities"] = velocities.astype(np.float32)
    P, C = params_record(scene)
    st = oracle.initial_state(scene, C)
    return scene, P, mats_record(scene), st


def _by_id(st, a):
    out = np.empty_like(a)
    out[st.id] = a
    return out


def test_rest_lattice_stays_at_rest(oracle):
    scene, P, mats, st = _blob(oracle, jitter=0.0)
    x0 = st.pos.copy()
    for _ in range(3):
        oracle.substep(st, P, mats, 20, False)
    assert np.array_equal(_by_id(st, st.pos), x0)
    assert not st.velA[:, :3].any()
    assert np.array_equal(F3(st.F), np.broadcast_to(ID9, (st.N, 9)))
    assert not st.aux[:, 1].any(), "lambda moved on a rest lattice"
    assert not st.Sp.any()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_free_fall_is_rigid(oracle, dtype):
    """Gravity only, a jittered block far from the floor: after 6 substeps every
    particle has the same velocity, the sum of 6 gravity steps dt g, F = I and
    lambda = 0; positions moved by the same displacement."""
    scene, P, mats, st = _blob(oracle, jitter=0.1, lift=0.1, gravity=-9.81)
    st = st.copy(dtype)
    x0 = st.pos[:, :3].astype(np.float64).copy()
    dt = np.dtype(dtype).type(P.dt)
    g = np.dtype(dtype).type(P.gravity[1])
    v = np.dtype(dtype).type(0)
    for _ in range(6):
        oracle.substep(st, P, mats, 20, False)
        v = v + dt * g
    vel = st.velA[:, :3]
    assert np.array_equal(vel, np.broadcast_to(vel[0], vel.shape)), "velocities differ between particles"
    assert vel[0, 0] == 0 and vel[0, 2] == 0
    assert abs(vel[0, 1] - v) <= 4 * np.finfo(dtype).eps * abs(v)
    assert np.array_equal(F3(st.F), np.broadcast_to(ID9, (st.N, 9)).astype(dtype))
    assert not st.aux[:, 1].any()
    d = _by_id(st, st.pos)[:, :3] - x0
    assert np.abs(d - d[0]).max() <= 1e-6 * np.abs(d).max() + 1e-12
    assert d[0, 1] < 0


def test_rigid_translation_has_no_impulse(oracle):
    v0 = np.array([0.3, -0.1, 0.2], np.float32)
    scene, P, mats, st = _blob(oracle, jitter=0.1, velocities=np.tile(v0, (1000, 1)))
    oracle.substep(st, P, mats, 20, False)
    assert np.array_equal(st.velA[:, :3], np.broadcast_to(v0, (st.N, 3)))
    assert np.array_equal(F3(st.F), np.broadcast_to(ID9, (st.N, 9)))
    assert not st.aux[:, 1].any()


def test_substep_conserves_momentum(oracle):
    """Random velocities (0.05 m/s) deform the block, so constraints, XSPH and possibly
    contacts act; no gravity and no wall: the total momentum is kept to f32 rounding."""
    rng = np.random.default_rng(5)
    vel = rng.normal(0, 0.05, (1000, 3))
    scene, P, mats, st = _blob(oracle, jitter=0.1, velocities=vel)
    m = float(scene["materials"][0]["density"]) * float(scene["h"]) ** 3
    p0 = m * st.velA[:, :3].astype(np.float64).sum(0)
    scale = m * np.abs(st.velA[:, :3].astype(np.float64)).sum()
    for _ in range(2):
        oracle.substep(st, P, mats, 20, False)
    assert np.abs(st.aux[:, 1]).max() > 0, "no constraint acted"
    dv = np.abs(st.velA[:, :3].astype(np.float64) - _by_id(st, np.zeros_like(st.velA[:, :3]))).max()
    assert dv > 0
    p1 = m * st.velA[:, :3].astype(np.float64).sum(0)
    assert np.abs(p1 - p0).max() <= 1e-5 * scale, (p0, p1)


def test_water_below_rest_density_has_no_pressure(oracle):
    """A water lattice at spacing 1.02 h (at spacing h its poly6 density is 1.0098 rho0):
    every particle's density is below rho0 (checked here with the oracle's own kernel),
    so PBF leaves lambda_pbf = 0.  With s_corr off (it acts at any density,
    simkernels.js:463-468) and no gravity, the block then stays exactly where it is."""
    scene, P, mats, st = _blob(oracle, jitter=0.0)
    scene = dict(scene, materials=[scenes.WATER(1000.0)], hasWater=True, sCorrK=0.0)
    P, _ = params_record(scene)
    mats = mats_record(scene)
    st.pos[:, 3] = 0.0  # material 0: water
    c = st.pos[:, :3].mean(0)
    st.pos[:, :3] = c + (st.pos[:, :3] - c) * np.float32(1.02)
    h = float(scene["h"])
    pts = st.pos[:, :3].astype(np.float64)
    i = int(np.argmin(np.linalg.norm(pts - pts.mean(0), axis=1)))  # an interior particle
    d = pts[i] - pts
    r = np.linalg.norm(d, axis=1)
    _, _, P6, _ = oracle.kernels(h, r, d)
    rho = 1000.0 * h ** 3 * P6.sum()  # includes the particle itself (simkernels.js:416)
    assert rho < 1000.0, rho
    x0 = st.pos.copy()
    oracle.substep(st, P, mats, 10, True)
    assert not st.aux[:, 3].any(), "PBF multiplier below rest density"
    assert np.array_equal(_by_id(st, st.pos), x0)
    assert not st.velA[:, :3].any()
