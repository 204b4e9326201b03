"""Pins of the 2D fp64 oracle (oracle/proto2d.py, the restatement of
prototype/xpbi2d.py) by what the mathematics fixes: LAPACK's singular values,
the kernel's normalisation and gradient, brute-force and k-d-tree pair sets,
the return map's cone, the energy's derivative, and whole-substep invariants
(a rest lattice stays at rest, free fall, momentum of the internal impulses);
and by the reference prototype's own run (golden vectors recorded from the
unmodified script)."""
import math
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import proto2d as o
from xpbi import proto2d as host

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mats(seed):
    g = np.random.default_rng(seed)
    A = g.standard_normal((400, 2, 2))
    A[0] = 0.0
    A[1] = [[2.0, 0.0], [0.0, 2.0]]
    A[2] = [[1.0, 2.0], [2.0, 4.0]]  # rank 1
    A[3] = [[0.0, 1.0], [1.0, 0.0]]  # reflection
    A[4] = [[math.cos(0.3), -math.sin(0.3)], [math.sin(0.3), math.cos(0.3)]]
    A[5] = [[1.0, 1e-9], [0.0, 1e-7]]
    A[6:50] = np.eye(2) + 1e-3 * g.standard_normal((44, 2, 2))  # near identity, like F
    return A


def test_svd2_against_lapack():
    A = _mats(1)
    U, S, V = o.svd2(A)
    ref = np.linalg.svd(A, compute_uv=False)
    scale = np.maximum(ref[:, :1], 1e-300)
    assert np.all(np.abs(S - ref) <= 1e-15 * scale + 1e-300)  # a few ulp of sigma_1
    assert np.all(S[:, 0] >= S[:, 1]) and np.all(S >= 0)
    assert np.max(np.abs(o.usv(U, S, V) - A)) <= 4e-15 * max(1.0, np.abs(A).max())
    I = np.eye(2)
    assert np.max(np.abs(np.einsum("nji,njk->nik", U, U) - I)) < 1e-14
    assert np.max(np.abs(np.einsum("nji,njk->nik", V, V) - I)) < 1e-14
    assert np.allclose(np.linalg.det(V), 1.0)
    dA = np.linalg.det(A)
    dU = np.linalg.det(U)
    assert np.all(dU[dA < 0] < 0) and np.all(dU[dA >= 0] > 0)


def test_wendland_2d_normalisation_and_gradient():
    prm = o.column40x80_params()
    h = prm["h"]
    n = 801
    ax = np.linspace(-2 * h, 2 * h, n)
    X, Y = np.meshgrid(ax, ax)
    r = np.hypot(X, Y).ravel()
    dA = (ax[1] - ax[0]) ** 2
    assert abs(o.kernel_W(prm, r).sum() * dA - 1.0) < 2e-4
    g = np.random.default_rng(2)
    d = g.uniform(-1.9 * h, 1.9 * h, (200, 2))
    d = d[np.hypot(d[:, 0], d[:, 1]) < 1.95 * h]
    rr = np.hypot(d[:, 0], d[:, 1])
    gw = o.kernel_gradW(prm, d, rr)
    e = 1e-7 * h
    for k in range(2):
        dp, dm = d.copy(), d.copy()
        dp[:, k] += e
        dm[:, k] -= e
        fd = (o.kernel_W(prm, np.hypot(dp[:, 0], dp[:, 1])) - o.kernel_W(prm, np.hypot(dm[:, 0], dm[:, 1]))) / (2 * e)
        assert np.max(np.abs(fd - gw[:, k])) < 1e-6 * np.abs(gw).max()
    # antisymmetry and the support edge
    assert np.array_equal(o.kernel_gradW(prm, -d, rr), -gw)
    assert o.kernel_W(prm, np.array([2 * h, 3 * h])).tolist() == [0.0, 0.0]


def test_pairs_against_brute_force_and_kdtree():
    x = host.column40x80_positions(seed=3)[:900]
    s = 0.02
    P, B = o.pairs(x, s)
    d = x[:, None, :] - x[None, :, :]
    r2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    bp, bb = np.nonzero((r2 <= s * s) & ~np.eye(len(x), dtype=bool))
    assert np.array_equal(P, bp) and np.array_equal(B, bb)  # np.nonzero is (p, b)-sorted
    kd = cKDTree(x).query_pairs(s * (1 - 1e-12), output_type="ndarray")
    inner = set(map(tuple, kd))
    got = set(zip(P.tolist(), B.tolist()))
    assert all((a, b) in got and (b, a) in got for a, b in inner)


def test_drucker_prager_cone_and_tip():
    prm = o.column40x80_params()
    g = np.random.default_rng(4)
    S = np.exp(g.normal(0, 0.02, (500, 2)))
    out = o.drucker_prager(prm, S)
    eps = np.log(S)
    tr = eps.sum(1)
    lo = np.log(out)
    dev = lo - lo.mean(1, keepdims=True)
    dn = np.hypot(dev[:, 0], dev[:, 1])
    k = prm["alpha_dp"] * (2 * prm["lam"] + 2 * prm["mu"]) / (2 * prm["mu"])
    tension = tr > 0
    assert np.allclose(out[tension], 1.0)
    dev0 = eps - eps.mean(1, keepdims=True)
    yielded = (~tension) & (np.hypot(dev0[:, 0], dev0[:, 1]) + k * tr > 0)
    elastic = (~tension) & ~yielded
    assert yielded.sum() > 20 and elastic.sum() > 20
    # on the cone, with the trace kept
    assert np.max(np.abs(dn[yielded] + k * tr[yielded])) < 1e-12
    assert np.allclose(lo[yielded].sum(1), tr[yielded], atol=1e-14)
    assert np.allclose(out[elastic], S[elastic], rtol=1e-14)


def test_stvk_energy_and_derivative():
    prm = o.column40x80_params()
    psi, _ = o.stvk(prm, np.ones((1, 2)))
    assert psi[0] == 0.0
    g = np.random.default_rng(5)
    S = np.exp(g.normal(0, 0.1, (50, 2)))
    psi, dpsi = o.stvk(prm, S)
    e = 1e-6
    for k in range(2):
        Sp, Sm = S.copy(), S.copy()
        Sp[:, k] *= 1 + e
        Sm[:, k] *= 1 - e
        fd = (o.stvk(prm, Sp)[0] - o.stvk(prm, Sm)[0]) / (2 * e * S[:, k])
        assert np.max(np.abs(fd - dpsi[:, k]) / np.abs(dpsi).max()) < 1e-7


def _free(prm, x):
    prm = dict(prm)
    prm["domain"] = np.array([50.0, 50.0])
    return prm, o.initial_state(x + 20.0)


def _lattice(n, h):
    ax = np.arange(n) * h
    X, Y = np.meshgrid(ax, ax)
    return np.stack([X.ravel(), Y.ravel()], 1)


def test_rest_lattice_stays_at_rest():
    prm, st = _free(o.column40x80_params(), _lattice(12, 0.01))
    prm["gravity"] = np.zeros(2)
    x0 = st["x"].copy()
    o.run(prm, st, 3)
    assert np.array_equal(st["x"], x0) and not st["v"].any() and np.array_equal(st["F"], np.tile(np.eye(2), (144, 1, 1)))


def test_free_fall_without_neighbours():
    prm, st = _free(o.column40x80_params(), np.stack([np.arange(50) * 0.1, np.zeros(50)], 1))
    o.run(prm, st, 10)
    assert np.allclose(st["v"][:, 1], 10 * prm["dt"] * prm["gravity"][1], rtol=1e-13)
    assert np.array_equal(st["F"], np.tile(np.eye(2), (50, 1, 1)))


@pytest.mark.parametrize("params", ["prototype", "column40x80"])
def test_internal_impulses_conserve_momentum(params):
    """Away from the walls, with V^n uniform (F = I), the constraint gather, the
    position correction and XSPH move no net momentum: sum v changes by N g dt."""
    prm = o.prototype_params() if params == "prototype" else o.column40x80_params()
    h = prm["h"]
    x = host.column(16, 16, h, 0.0, 0.0, seed=7)
    g = np.random.default_rng(8)
    x = x + g.normal(0, 0.2 * h, x.shape)  # compressed and stretched spots: active constraints, contacts
    prm, st = _free(prm, x)
    st["v"] = g.normal(0, 0.05, st["v"].shape)
    v0 = st["v"].sum(0)
    o.substep(prm, st)
    want = v0 + len(x) * prm["dt"] * prm["gravity"]
    assert np.max(np.abs(st["v"].sum(0) - want)) < 1e-12 * len(x)
    assert np.abs(st["lam"]).max() > 0  # constraints were active


def test_host_and_oracle_columns_agree():
    assert np.array_equal(host.column40x80_positions(), o.column40x80_state()["x"])
    assert np.array_equal(host.prototype_positions(), o.prototype_state()["x"])
    assert len(host.prototype_positions()) == 666 and len(host.column40x80_positions()) == 3200


def test_oracle_matches_the_reference_prototype_run():
    """The reference's own prototype script, run unmodified in the build container
    (tests/golden/make_proto2d_golden.py), recorded its initial column and its state
    after steps 1, 2, 5, 10 and 20 (tests/golden/proto2d_golden.npz).  The oracle,
    from the same parameters and column, lands on the same states: positions,
    velocities and F to 1e-12 relative (numpy's LAPACK SVD and kd-tree pair order
    against the oracle's closed-form 2x2 SVD and sorted pairs)."""
    g = np.load(os.path.join(GOLDEN, "proto2d_golden.npz"))
    st = o.prototype_state()
    assert np.array_equal(st["x"], g["x0"]), "initial column differs from the script's"
    prm, done = o.prototype_params(), 0
    for k in g["steps"]:
        o.run(prm, st, int(k) - done)
        done = int(k)
        for f in ("x", "v", "F"):
            ref = g[f"{f}{k}"]
            err = np.abs(st[f] - ref).max() / np.abs(ref).max()
            assert err <= 1e-12, (k, f, err)
