"""How far any warm start of the Jacobi multipliers can cut the iterations to
residual 1e-4: the ceiling experiment, run on the CPU oracle (test
infrastructure; nothing in the product imports it).

    python oracle/warmstart_ceiling.py [--nx 24] [--ny 48] [--r 0.005] [--settle 300]
                                       [--substeps 4] [--cap 200] [--tol 1e-4] [--threads 8]

A Drucker-Prager column (phi 30 deg, E 3.5e5, nu 0.3, rho 2000, dt 1/6000, the
benchmark column's material and step at a smaller size) is settled with 20
fixed iterations per substep.  Then, for each of `substeps` consecutive
substeps, from the same state:

* plain: lambda = 0, Jacobi until ||num|| <= tol ||C at lambda = 0|| (cap);
* exact: the first iteration applies lambda0 = that substep's converged lambda
  (lambda* of the plain run, matched by particle id): the best any predictor of
  lambda0 alone can do;
* prev: lambda0 = the previous substep's converged lambda (free at run time);
* half: lambda0 = lambda* / 2 (a predictor with 50 % error everywhere).

Warm starts count their first iteration.  Each line gives the iterations and
the residual after the first iteration; the last line the iteration ratios
plain / variant over all substeps.  The state advances with the plain run.

The residual after the exact warm step is not zero: lambda* belongs to the
velocities reached by the plain run's iterates, and one Jacobi step from the
predicted velocities with lambda* applied lands near but not on them; Jacobi
then needs most of its iterations for the last decades of the residual.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(REPO, "xpbi-webgpu_amd"), os.path.join(REPO, "tests")]

import oracle_py as O  # noqa: E402
from xpbi import scenes  # noqa: E402
from helpers import mats_record, params_record  # noqa: E402


def by_id(state, values):
    out = np.empty(state.N, values.dtype)
    out[state.id] = values
    return out


def one(state, P, M, cap, tol, threads, lam0=None):
    """One substep from a copy of state; returns (new state, iterations, residual
    after the first iteration, lambda by id at the end of the solve)."""
    s = state.copy()
    resid = np.zeros(2 * cap)
    if lam0 is None:
        n = O.substep(s, P, M, cap, False, nthreads=threads, resid=resid, tol=tol, collect=True)
    else:
        n = O.substep(s, P, M, cap, False, nthreads=threads, resid=resid, tol=tol, lam0_by_id=lam0)
    c0 = resid[1] if lam0 is None else None
    return s, n, resid, by_id(s, s.aux[:, 1].copy()), c0


def experiment(nx=24, ny=48, r=0.005, settle=300, substeps=4, cap=200, tol=1e-4, threads=8, log=print):
    sc = scenes.sand_column(nx, ny, nx, r, [max(0.8, 3 * nx * r)] * 3, seed=11, iterations=20)
    P, Cn = params_record(sc)
    M = mats_record(sc)
    st = O.initial_state(sc, Cn)
    t0 = time.time()
    for _ in range(settle):
        O.substep(st, P, M, 20, False, nthreads=threads)
    log(f"column {nx}x{ny}x{nx} at r = {r} ({st.N} particles), settled {settle} substeps in {time.time() - t0:.0f} s")
    rows = []
    for k in range(substeps):
        lam_prev = by_id(st, st.aux[:, 1].copy())
        nxt, n0, res0, lam_star, c0 = one(st, P, M, cap, tol, threads)
        r1 = lambda res, it=0: float(np.sqrt(res[2 * it] / c0)) if c0 > 0 else float("nan")
        row = dict(plain=n0, plain_r1=r1(res0, 1))
        for name, l0 in (("exact", lam_star), ("prev", lam_prev), ("half", 0.5 * lam_star)):
            _, n, res, _, _ = one(st, P, M, cap, tol, threads, l0)
            # res[0]: the residual the first plain iteration after the warm step
            # measures, as res0[2] does after plain's first iteration
            row[name], row[name + "_r1"] = n, r1(res)
        rows.append(row)
        log(f"substep {k}: " + " | ".join(f"{v} {row[v]} (r {row[v + '_r1']:.3g})"
                                          for v in ("plain", "exact", "prev", "half")))
        st = nxt
    tot = {v: sum(rw[v] for rw in rows) for v in ("plain", "exact", "prev", "half")}
    ratios = {v: tot["plain"] / tot[v] for v in ("exact", "prev", "half")}
    log("iterations over all substeps: " + ", ".join(f"{v} {tot[v]}" for v in tot) + "; plain / variant: " +
        ", ".join(f"{v} {ratios[v]:.3f}" for v in ratios))
    return rows, ratios


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=24)
    ap.add_argument("--ny", type=int, default=48)
    ap.add_argument("--r", type=float, default=0.005)
    ap.add_argument("--settle", type=int, default=300)
    ap.add_argument("--substeps", type=int, default=4)
    ap.add_argument("--cap", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    O.build()
    experiment(a.nx, a.ny, a.r, a.settle, a.substeps, a.cap, a.tol, threads=a.threads,
               log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
