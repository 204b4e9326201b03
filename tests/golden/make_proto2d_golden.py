"""Golden vectors of the reference's 2D prototype, prototype/xpbi2d.py.

Run in the build container only (the reference is not on the GPU box, and this
script never travels with a test run): python tests/golden/make_proto2d_golden.py

The prototype is a script with no API: it builds its 666-particle sand column
(xpbi2d.py:29-58) and runs its main loop (xpbi2d.py:123-241) for 15,001 steps,
writing PNG frames into ./frames.  This runs it unmodified, from a temporary
working directory, with two pass-through wrappers that only observe:

* scipy.spatial.cKDTree, called once at the start of every step with the
  positions x^n (xpbi2d.py:125): its first call records the initial positions;
* numpy.clip, called by the script once at the end of every step on the
  updated positions (xpbi2d.py:234; calls from inside libraries pass through
  unrecorded): it records x, v and F after the step (v and F are the
  script's globals at that point) and the wall time of the step, and after
  STEPS steps stops the script.

It writes proto2d_golden.npz beside itself: the initial positions and the state
after steps 1, 2, 5, 10 and 20, and the script's mean wall time per step over
steps 3-20 on this container's CPU (numpy + scipy, one process).  Nothing of
the reference's source is stored: only inputs and outputs of its run.
"""
import os
import runpy
import sys
import tempfile
import time

import numpy as np
import scipy.spatial

REF = "/root/reference/prototype/xpbi2d.py"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "proto2d_golden.npz")
STEPS = 20
KEEP = (1, 2, 5, 10, 20)


class _Stop(Exception):
    pass


def main():
    rec = {"x0": None, "t": []}
    real_tree, real_clip = scipy.spatial.cKDTree, np.clip

    def tree(data, *a, **kw):
        if rec["x0"] is None:
            rec["x0"] = np.array(data, np.float64)
        return real_tree(data, *a, **kw)

    def clip(a, lo, hi, *args, **kw):
        out = real_clip(a, lo, hi, *args, **kw)
        g = sys._getframe(1).f_globals
        if g.get("__name__") != "xpbi2d_prototype":
            return out  # a call from inside a library, not the script's position update
        rec["t"].append(time.perf_counter())
        k = len(rec["t"])
        if k in KEEP:
            rec[f"x{k}"] = np.array(out, np.float64)
            rec[f"v{k}"] = np.array(g["v"], np.float64)
            rec[f"F{k}"] = np.array(g["F"], np.float64)
        if k >= STEPS:
            raise _Stop()
        return out

    cwd = os.getcwd()
    scipy.spatial.cKDTree, np.clip = tree, clip
    try:
        with tempfile.TemporaryDirectory() as tmp:
            os.chdir(tmp)
            try:
                runpy.run_path(REF, run_name="xpbi2d_prototype")
            except _Stop:
                pass
    finally:
        os.chdir(cwd)
        scipy.spatial.cKDTree, np.clip = real_tree, real_clip
    t = np.array(rec["t"])
    ms = float(np.mean(np.diff(t[2:]))) * 1e3
    arrays = {k: v for k, v in rec.items() if k != "t"}
    np.savez_compressed(OUT, steps=np.array(KEEP), ms_per_step=np.array(ms), **arrays)
    print(f"{OUT}: N = {arrays['x0'].shape[0]}, {ms:.2f} ms per step (steps 3-{STEPS})")


if __name__ == "__main__":
    main()
