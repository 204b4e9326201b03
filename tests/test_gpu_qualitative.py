"""Qualitative behaviour the reference validates visually (docs/IMPLEMENTATION.md:111-121):
sand piling at varying friction angle.  A 16x40x16 Drucker-Prager column at
r = 0.01 m collapses for 3,500 substeps (0.58 s); the pile's surface angle, read off
the particles by xpbi.readback.pile_angle, rises with the friction angle and stays
below it (a dynamic collapse runs out below the static angle of repose), and the
pile has come to rest (99 % of the particles slower than 5 cm/s)."""
import numpy as np
import pytest

import xpbi
from xpbi import readback, scenes

pytestmark = pytest.mark.gpu


def test_pile_angle_rises_with_friction_angle():
    angles = []
    for phi in (20.0, 30.0, 40.0):
        sc = scenes.sand_column(16, 40, 16, 0.01, [1.4, 0.6, 1.4], seed=3, phi=phi)
        sim = xpbi.Simulation(sc, device=0)
        sim.writeParams(xpbi.params_from_scene(sc))
        print(f"phi {phi}: collapsing", flush=True)
        sim.step(3500, 20, False)
        pos, vel = sim.read("POS"), sim.read("VEL")
        sim.close()
        ang = readback.pile_angle(pos[:, :3], floor=0.0, bins=30)
        speed = np.linalg.norm(vel[:, :3], axis=1)
        q99 = float(np.quantile(speed, 0.99))
        print(f"phi {phi}: pile angle {ang:.1f} deg, top {pos[:, 1].max():.3f} m, speed q99 {q99:.4f} "
              f"max {speed.max():.3f} m/s", flush=True)
        assert q99 < 0.05, "the pile is still moving"
        assert 0.25 * phi < ang < phi, (phi, ang)
        angles.append(ang)
    assert angles[0] < angles[1] < angles[2], angles
