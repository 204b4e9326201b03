"""Checkpoint and resume (xpbi/checkpoint.py, xpbi/run_scene.py): a run split by a
save into a fresh simulation and a load continues bit for bit like the run that
was never interrupted, on the reference dam breach (water + sand: PBF lambda,
plastic state and the persistent ids all carry over) and on a von Mises block
poked by a moving sphere (placed from the restored substep count)."""
import io

import numpy as np
import pytest

import xpbi
from xpbi import checkpoint, run_scene, scenes

pytestmark = pytest.mark.gpu


def _poke():
    sc = scenes.vm_poke(16, 0.01, seed=5)
    top = float(sc["positions"][:, 1].max())
    sc["collider"] = dict(sc["collider"], pos=[0.5, top + 10.2 * 0.01, 0.5])  # touches within a few substeps
    return sc


@pytest.mark.parametrize("name", ["Dam breach: water + sand", "poke"])
def test_resume_is_bit_identical(tmp_path, name):
    name = _poke() if name == "poke" else name
    a, b = tmp_path / "a.npz", tmp_path / "b.npz"
    log = io.StringIO()
    run_scene.run(name, 24, save=str(a), every=0, out=log)
    run_scene.run(name, 12, save=str(b), every=0, out=log)
    c = tmp_path / "c.npz"
    run_scene.run(name, 12, load=str(b), save=str(c), every=0, out=log)
    with np.load(a) as za, np.load(c) as zc:
        assert int(za["substeps"]) == int(zc["substeps"]) == 24
        for f in checkpoint.FIELDS:
            assert np.array_equal(za[f], zc[f]), f
        # the state moved: the checkpoint is not the initial one
        sc = run_scene.make_scene(name) if isinstance(name, str) else name
        assert not np.array_equal(za["POS"][:, :3][np.argsort(za["ID"])], sc["positions"])


def test_load_rejects_another_size(tmp_path):
    p = tmp_path / "s.npz"
    sc = scenes.sand_column(6, 6, 6, 0.01, [0.3, 0.3, 0.3], seed=1)
    sim = xpbi.Simulation(sc, device=0)
    checkpoint.save(sim, str(p), 0)
    sim.close()
    sc2 = scenes.sand_column(6, 7, 6, 0.01, [0.3, 0.3, 0.3], seed=1)
    sim2 = xpbi.Simulation(sc2, device=0)
    with pytest.raises(ValueError):
        checkpoint.load(sim2, str(p))
    sim2.close()
