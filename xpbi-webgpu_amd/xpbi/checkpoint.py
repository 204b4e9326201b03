"""Checkpoints of a running simulation: the per-particle state that carries from
one substep to the next, written to and read from a NumPy .npz file.

The fields are POS (x, y, z, material tag), VEL (velA), F, AUX (V^n, lambda,
plastic state, PBF lambda) and ID (the persistent particle id; the within-cell
order of the sort is ascending id, so restoring it makes the next sort place the
particles exactly where an uninterrupted run would).  L, Sp, the cell table and
the permutation are rebuilt by every substep and are not stored.  A checkpoint
loads into a simulation created from the same scene (same particle count, grid
and materials); the substep count is kept so that a driver can place time-
dependent inputs such as a moving collider.  Reference: main.js:289-313 reads
the same buffers back for the render and the HUD.
"""
import numpy as np

FIELDS = ("POS", "VEL", "F", "AUX", "ID")


def save(sim, path, substeps=0):
    """Write sim's state after `substeps` substeps to path (.npz, uncompressed)."""
    arrays = {f: sim.read(f) for f in FIELDS}
    np.savez(path, substeps=np.int64(substeps), N=np.int64(sim.N), **arrays)


def load(sim, path):
    """Restore a checkpoint into sim; returns its substep count.  Raises ValueError
    when the checkpoint was written by a simulation of another size."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["N"]) != sim.N:
            raise ValueError(f"checkpoint holds {int(z['N'])} particles, the simulation {sim.N}")
        for f in FIELDS:
            sim.write(f, z[f])
        return int(z["substeps"])
