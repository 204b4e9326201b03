"""The 2D fp64 engine (xpbi2d_*) against the oracle's restatement of
prototype/xpbi2d.py on the same inputs.

Bar: float64, every sum in the same order; the two may differ only where the
device's and numpy's log/exp round apart (a few ulp), so positions, velocities, F
and lambda agree to 1e-11 relative after 20 substeps of the 3,200-particle column
and after 150 of the prototype's own 666-particle column (chaotic growth of the
ulp differences stays below that over these spans)."""
import numpy as np
import pytest

import proto2d as o
from xpbi import XpbiError
from xpbi import proto2d as host

pytestmark = pytest.mark.gpu
TOL = 1e-11


def _oracle_prm(p):
    return o.make_params(p["h"], p["E"], p["nu"], p["rho"], p["phi_deg"], p["dt"], p["iterations"], p["gravity"],
                         p["xsph_c"], p["domain"], p["omega"])


def _compare(sim, st, what):
    errs = {}
    for k, ref in (("X", st["x"]), ("V", st["v"]), ("F", st["F"]), ("LAMBDA", st["lam"])):
        got = sim.read(k)
        scale = max(np.abs(ref).max(), 1e-30)
        errs[k] = float(np.abs(got - ref).max() / scale)
    print(what, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL, (what, k, v)


@pytest.mark.parametrize("case,chunks", [("column40x80", [1, 4, 15]), ("prototype", [10, 40, 100])])
def test_engine_matches_oracle(case, chunks):
    if case == "column40x80":
        p, x = host.column40x80_params(), host.column40x80_positions()
    else:
        p, x = host.prototype_params(), host.prototype_positions()
    prm = _oracle_prm(p)
    st = o.initial_state(x)
    sim = host.Prototype2D(x, p)
    done = 0
    for k in chunks:
        print(f"{case}: substeps {done}..{done + k}", flush=True)
        sim.step(k)
        o.run(prm, st, k)
        done += k
        _compare(sim, st, f"{case} after {done}")
    assert np.abs(st["v"]).max() > 1e-3  # the column moved
    sim.close()


def test_rest_lattice_and_momentum_on_device():
    p = dict(host.column40x80_params(), domain=[4.0, 4.0], gravity=[0.0, 0.0])
    ax = np.arange(12) * 0.01 + 1.5
    X, Y = np.meshgrid(ax, ax)
    lat = np.stack([X.ravel(), Y.ravel()], 1)
    sim = host.Prototype2D(lat, p)
    sim.step(3)
    assert np.array_equal(sim.x, lat) and not sim.v.any()
    sim.close()
    # a disturbed blob away from the walls: internal impulses move no momentum
    g = np.random.default_rng(8)
    x = host.column(16, 16, 0.01, 1.5, 1.5, seed=7) + g.normal(0, 0.002, (256, 2))
    v = g.normal(0, 0.05, (256, 2))
    p = dict(host.column40x80_params(), domain=[4.0, 4.0])
    sim = host.Prototype2D(x, p)
    sim.write("V", v)
    sim.step(1)
    want = v.sum(0) + 256 * p["dt"] * np.asarray(p["gravity"])
    assert np.max(np.abs(sim.v.sum(0) - want)) < 1e-12 * 256
    assert np.abs(sim.lam).max() > 0
    sim.close()


def test_neighbour_overflow_is_an_error():
    x = 0.5 + np.random.default_rng(1).uniform(-0.002, 0.002, (60, 2))  # 60 particles within 2h
    sim = host.Prototype2D(x, host.column40x80_params())
    with pytest.raises(XpbiError, match="neighbours"):
        sim.step(1)
    sim.close()


def test_cooperative_grid_equals_single_workgroup():
    """The cooperative grid (one thread per particle, grid barriers) and one workgroup
    (workgroup barriers) run the same arithmetic: identical state, bit for bit."""
    p, x = host.column40x80_params(), host.column40x80_positions()
    out = {}
    for single in (False, True):
        sim = host.Prototype2D(x, p)
        sim.set_workgroups(single)
        ms = sim.time(20)
        print(f"{'single workgroup' if single else 'cooperative grid'}: {ms / 20:.3f} ms per substep", flush=True)
        out[single] = {k: sim.read(k) for k in ("X", "V", "F", "LAMBDA")}
        sim.close()
    for k in out[False]:
        assert np.array_equal(out[False][k], out[True][k]), k


def test_engine_matches_the_reference_prototype_run():
    """The 2D engine against the reference prototype script's own run: its initial
    column and its states after steps 1, 2, 5, 10 and 20, recorded in the build
    container (tests/golden/make_proto2d_golden.py), to 1e-11 relative."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "proto2d_golden.npz"))
    x = host.prototype_positions()
    assert np.array_equal(x, g["x0"]), "initial column differs from the script's"
    sim = host.Prototype2D(x, host.prototype_params())
    done = 0
    for k in g["steps"]:
        sim.step(int(k) - done)
        done = int(k)
        for f, key in (("X", "x"), ("V", "v"), ("F", "F")):
            ref = g[f"{key}{k}"]
            got = sim.read(f).reshape(ref.shape)
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            print(f"step {k} {f}: rel err {err:.2e}", flush=True)
            assert err <= TOL, (k, f, err)
    sim.close()
