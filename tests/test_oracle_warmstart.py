"""The oracle's learned warm-start machinery (CPU): the C++ MLP forward against the
numpy oracle, the HB bracket hint against scipy's root finder, the residual loop
against fixed iteration counts, and the zero network against plain Jacobi."""
import math

import numpy as np
import pytest
from scipy import optimize

import mlp_oracle as mo
import xpbi
from xpbi import scenes
from xpbi.mlp import init_params
from helpers import mats_record, params_record
from test_oracle_math import _rand_sig, mat


def test_cpp_mlp_forward_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    p = init_params(3)
    d = mo.split(p)
    d["b1"], d["b2"], d["b3"] = rng.normal(0, 0.3, 64), rng.normal(0, 0.3, 64), rng.normal(0, 0.3, 2)
    p = mo.join(d).astype(np.float32)
    X = rng.normal(0, 2.0, size=(3000, 12))
    y = oracle.mlp_forward(p, X)
    yo, _ = mo.forward(p, X)
    scale = np.abs(yo).max()
    bad = (np.abs(y - yo) > 1e-6 * scale).any(axis=1).mean()
    assert bad <= 1e-3 and np.abs(y - yo).max() <= 2e-2 * scale


@pytest.mark.parametrize("hexp", [0.5, 2.0])
def test_hb_hint_keeps_the_root(oracle, hexp):
    dt = 1e-3
    m, mu, lam = mat(4, E=2250.0, nu=0.125, p0=20.0, p1=hexp, p2=25.0)
    sigY, eta = float(m.p0), float(m.p2)
    sig = _rand_sig(3000, 0.03, 7)
    plain, t = oracle.hb_return_map(m, sig, -1.0, dt)
    ok = ~np.isnan(t)
    assert ok.sum() > 100 and (t[ok] >= 0).all() and (t[ok] <= 1).all()
    rng = np.random.default_rng(1)
    # a hint within the +-1/64 window: 13 halvings from the narrow bracket
    near = np.where(ok, np.clip(t + rng.uniform(-0.012, 0.012, t.size), 0, 1), -1.0)
    out, t2 = oracle.hb_return_map(m, sig, near, dt)
    # a wrong hint: the full bracket, exactly the plain result
    far = np.where(ok, np.where(t > 0.5, t - 0.3, t + 0.3), -1.0)
    out3, t3 = oracle.hb_return_map(m, sig, far, dt)
    np.testing.assert_array_equal(out3, plain)
    np.testing.assert_array_equal(out[~ok], plain[~ok])
    J = sig.prod(axis=1)
    bbar = sig ** 2 * J[:, None] ** (-2 / 3)
    trb = bbar.sum(axis=1)
    sPre = mu * np.linalg.norm(bbar - trb[:, None] / 3, axis=1)
    sY = math.sqrt(2 / 3) * sigY
    muT = trb / 3 * mu
    for i in np.nonzero(ok)[0][:300]:
        g = lambda s: eta ** (1 / hexp) * (s - sPre[i]) + 2 * muT[i] * dt * max(s - sY, 0.0) ** (1 / hexp)
        root = optimize.brentq(g, sY, sPre[i], xtol=1e-14)
        s_hint = sY + t2[i] * (sPre[i] - sY)
        assert abs(s_hint - root) <= (sPre[i] - sY) * 2.0 ** -17 + 1e-9 * sPre[i]
    np.testing.assert_allclose(out[ok], plain[ok], rtol=1e-5)


_LANDED = {}


def _landed(oracle, dtype=np.float32):
    """A 2000-particle DP column a few substeps after it lands (constraints active)."""
    if not _LANDED:
        scene = scenes.sand_column(10, 20, 10, 0.01, [0.4, 0.35, 0.4], seed=4, iterations=20)
        P, C = params_record(scene)
        st = oracle.initial_state(scene, C)
        mats = mats_record(scene)
        for _ in range(240):  # free fall ends near substep 180
            oracle.substep(st, P, mats, 10, False)
        _LANDED.update(scene=scene, P=P, mats=mats, st=st)
    L = _LANDED
    return L["scene"], L["P"], L["mats"], L["st"].copy(dtype)


def test_residual_loop_equals_fixed_count(oracle):
    scene, P, mats, st = _landed(oracle)
    a = st.copy()
    r = np.zeros(2 * 200)
    n = oracle.substep(a, P, mats, 200, False, tol=1e-3, resid=r)
    rel = np.sqrt(r[0:2 * n:2] / r[1])
    assert 1 < n < 200 and rel[n - 1] <= 1e-3 and (rel[: n - 1] > 1e-3).all()
    b = st.copy()
    assert oracle.substep(b, P, mats, n, False) == n
    for f in ("pos", "velA", "F", "aux"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def omega_network(omega=0.5):
    """y = (omega, -1): lambda0 = omega d0 (plain Jacobi's first step), no HB hint."""
    p = np.zeros(5122, np.float32)
    p[5120], p[5121] = omega, -1.0
    return p


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_omega_network_is_plain_jacobi(oracle, dtype):
    """A network predicting rho = omega makes the learned first iteration plain Jacobi's: the
    substeps agree bit for bit and count the same iterations."""
    scene, P, mats, st = _landed(oracle, dtype)
    for tol in (0.0, 1e-3):
        a, b = st.copy(), st.copy()
        for _ in range(2):
            na = oracle.substep(a, P, mats, 40, False, tol=tol, warm=omega_network())
            nb = oracle.substep(b, P, mats, 40, False, tol=tol)
            assert na == nb
            for f in ("pos", "velA", "F", "aux"):
                assert np.array_equal(getattr(a, f), getattr(b, f)), (tol, f)


def test_collect_pairs(oracle):
    scene, P, mats, st = _landed(oracle)
    feat = np.zeros((st.N, 16), np.float32)
    tgt = np.zeros((st.N, 2), np.float32)
    n = oracle.substep(st, P, mats, 200, False, tol=1e-4, collect=True, feat=feat, target=tgt)
    active = feat[:, 12] != 0
    assert active.mean() > 0.5
    assert np.isfinite(tgt[active, 0]).all() and np.isnan(tgt[~active, 0]).all()
    # the converged multiplier of a landed column is a few plain first steps (omega = 0.5)
    med = np.median(tgt[active, 0])
    assert 0.2 < med < 20, med
    # features are finite; the constraint column holds C, the neighbour column n/40
    assert np.isfinite(feat[active, :12]).all()
    assert (feat[active, 14] >= 1e-7).all()
    # (about 33 support neighbours inside the column, fewer on its surface)
    assert (feat[active, 9] > 0).all() and 0.6 < np.median(feat[active, 9]) < 1.2


def test_given_lambda0_equals_learned_start(oracle):
    """The given-lambda0 start (lambda0 per particle id, the ceiling experiments of
    oracle/warmstart_ceiling.py) fed the learned start's own lambda0 = clamp(y0, 0, 2) d0
    of each particle reproduces the learned substep bit for bit; fed zero it applies no
    multiplier in its first iteration."""
    scene, P, mats, st = _landed(oracle)
    net = init_params(5)
    net[5120] = 0.7  # output bias: rho_hat near the plain first step, a non-trivial lambda0
    a = st.copy()
    feat = np.zeros((st.N, 16), np.float32)
    na = oracle.substep(a, P, mats, 60, False, tol=1e-3, warm=net, feat=feat)
    y = oracle.mlp_forward(net, feat[:, :12].astype(np.float64))
    y0 = y[:, 0].astype(np.float32)
    lam0 = np.where((feat[:, 12] == 0) | np.isnan(y0), np.float32(0),
                    np.clip(y0, np.float32(0), np.float32(2)) * feat[:, 12]).astype(np.float32)
    assert (lam0 != 0).mean() > 0.5
    by_id = np.zeros(st.N, np.float32)
    by_id[a.id] = lam0  # the substep's sorted order is the order a ends in
    b = st.copy()
    nb = oracle.substep(b, P, mats, 60, False, tol=1e-3, lam0_by_id=by_id)
    assert na == nb
    for f in ("pos", "velA", "F", "aux"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    # lambda0 = 0 with one iteration: the given start is the only iteration and
    # leaves every multiplier at zero; a given lambda0 = 1 leaves it at 1 wherever
    # the constraint is active
    c = st.copy()
    assert oracle.substep(c, P, mats, 1, False, lam0_by_id=np.zeros(st.N, np.float32)) == 1
    assert not c.aux[:, 1].any()
    c = st.copy()
    f1 = np.zeros((st.N, 16), np.float32)
    oracle.substep(c, P, mats, 1, False, lam0_by_id=np.ones(st.N, np.float32), feat=f1)
    assert np.array_equal(c.aux[:, 1], (f1[:, 12] != 0).astype(np.float32))
