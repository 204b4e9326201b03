"""Pins of the CPU oracle's math library against what the mathematics fixes.

Each check is independent of the oracle's own formulas: integrals of the SPH
kernels, finite differences, LAPACK singular values, geometric properties of the
return mappings (the projected state lies on the yield surface), brute-force
grid hashing.  These are what make the oracle trustworthy as the parity judge.
"""
import math

import numpy as np
import pytest
from scipy import integrate, optimize

import xpbi
from xpbi import scenes
from helpers import params_record


def mat(kind, E=1e4, nu=0.3, density=1000.0, **p):
    """One 64-B record; mu and lambda are returned as the f32 values the record holds."""
    m = dict(kind=kind, E=E, nu=nu, density=density, **p)
    r = xpbi.pack_materials([m], 1e-6)[0]
    return r, float(r.mu), float(r.lam)


# ---------------------------------------------------------------- SPH kernels
@pytest.mark.parametrize("h", [0.005, 0.01, 0.016])
def test_wendland_normalised(oracle, h):
    f = lambda r: 4 * math.pi * r * r * oracle.kernels(h, np.array([r]), np.zeros((1, 3)))[0][0]
    val, _ = integrate.quad(f, 0, 2 * h, limit=200)
    assert abs(val - 1.0) < 1e-7


@pytest.mark.parametrize("h", [0.005, 0.01])
def test_poly6_and_spiky_normalised(oracle, h):
    H = 2 * h
    f = lambda r: 4 * math.pi * r * r * oracle.kernels(h, np.array([r]), np.zeros((1, 3)))[2][0]
    assert abs(integrate.quad(f, 0, H, limit=200)[0] - 1.0) < 1e-7
    # the spiky gradient is the derivative of 15/(pi H^6) (H - r)^3, a unit-mass kernel
    spiky = lambda r: 15 / (math.pi * H ** 6) * (H - r) ** 3
    assert abs(integrate.quad(lambda r: 4 * math.pi * r * r * spiky(r), 0, H)[0] - 1.0) < 1e-12
    rng = np.random.default_rng(0)
    d = rng.normal(size=(50, 3))
    d = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.05, 0.95, (50, 1)) * H
    r = np.linalg.norm(d, axis=1)
    g = oracle.kernels(h, r, d)[3]
    eps = 1e-7 * H
    dW = (spiky(r + eps) - spiky(r - eps)) / (2 * eps)
    np.testing.assert_allclose(g, d / r[:, None] * dW[:, None], rtol=1e-6)


def test_wendland_gradient_is_derivative(oracle):
    h = 0.01
    rng = np.random.default_rng(1)
    d = rng.normal(size=(200, 3))
    d = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.02, 1.98, (200, 1)) * h
    r = np.linalg.norm(d, axis=1)
    _, g, _, _ = oracle.kernels(h, r, d)
    eps = 1e-7 * h
    Wp = oracle.kernels(h, r + eps, d)[0]
    Wm = oracle.kernels(h, r - eps, d)[0]
    np.testing.assert_allclose(g, d / r[:, None] * ((Wp - Wm) / (2 * eps))[:, None], rtol=1e-6, atol=1e-3)
    # antisymmetric, and zero at r >= 2h and r < 1e-9
    np.testing.assert_array_equal(oracle.kernels(h, r, -d)[1], -g)
    far = np.array([[2 * h, 0, 0], [0, 3 * h, 0], [1e-10, 0, 0]])
    assert not oracle.kernels(h, np.linalg.norm(far, axis=1), far)[1].any()
    # f32 restatement agrees with f64 to f32 precision
    g32 = oracle.kernels(h, r, d, dtype=np.float32)[1]
    np.testing.assert_allclose(g32, g, rtol=2e-5, atol=1e-2)


# ---------------------------------------------------------------- grid hash
def test_cell_index_brute_force(oracle):
    scene = scenes.sand_column(4, 4, 4, 0.01, [0.3, 0.2, 0.4], seed=1)
    P, C = params_record(scene)
    o = np.array(P.gridOrigin[:], np.float32)
    cs = np.float32(P.cellSize)
    dims = np.array(P.gridDims[:])
    rng = np.random.default_rng(2)
    p = rng.uniform(-0.2, 0.6, (5000, 3)).astype(np.float32)
    p[:10] = [[np.nan, 0, 0], [0, np.inf, 0], [-np.inf, 0, 0], [1e30, 1e30, 1e30], [-1e30, 0, 0],
              [0, 0, 0], o, o + cs, [0.3, 0.2, 0.4], [1e-30, -1e-30, 0]]
    with np.errstate(invalid="ignore", over="ignore"):
        q = np.floor((p - o) / cs)
    q = np.where(np.isnan(q), 0, q)
    q = np.clip(q, -2.0 ** 31, 2.0 ** 31 - 1)
    c = np.clip(q.astype(np.int64), 0, dims - 1)
    ref = c[:, 0] + c[:, 1] * dims[0] + c[:, 2] * dims[0] * dims[1]
    np.testing.assert_array_equal(oracle.cell_index(P, p), ref.astype(np.uint32))
    assert ref.max() < C


# ---------------------------------------------------------------- svd3
def _svd_cases(n=400, seed=3):
    rng = np.random.default_rng(seed)
    cases = [np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3)),          # near identity (the common case)
             rng.normal(size=(n, 3, 3)),                              # general
             np.eye(3) * rng.uniform(0.5, 2, (n, 1, 1)) + 0.3 * rng.normal(size=(n, 3, 3))]
    inv = np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3))
    inv[:, :, 2] *= -1                                                # det < 0 (inverted)
    cases.append(inv)
    diag = np.zeros((n, 3, 3))
    diag[:, [0, 1, 2], [0, 1, 2]] = rng.uniform(0.5, 1.5, (n, 3))
    cases.append(diag)
    return np.concatenate(cases)


def test_svd3_reconstructs_and_matches_lapack(oracle):
    F = _svd_cases()
    U, sig, V = oracle.svd3(F)
    rec = np.einsum("nij,nj,nkj->nik", U, sig, V)
    np.testing.assert_allclose(rec, F, atol=1e-9)
    I = np.eye(3)
    np.testing.assert_allclose(np.einsum("nji,njk->nik", U, U), np.broadcast_to(I, U.shape), atol=1e-9)
    np.testing.assert_allclose(np.einsum("nji,njk->nik", V, V), np.broadcast_to(I, V.shape), atol=1e-9)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-9)
    np.testing.assert_allclose(np.linalg.det(V), 1.0, atol=1e-9)
    detF = np.linalg.det(F)
    assert np.array_equal(sig[:, 2] < 0, detF < 0)
    assert (sig[:, 0] >= sig[:, 1]).all() and (sig[:, 1] >= np.abs(sig[:, 2])).all()
    lap = np.linalg.svd(F, compute_uv=False)
    np.testing.assert_allclose(np.abs(sig), lap, rtol=1e-9, atol=1e-12)


def test_svd3_f32_accuracy(oracle):
    F = _svd_cases(seed=4)
    F = F[np.linalg.cond(F) < 20]
    U, sig, V = oracle.svd3(F.astype(np.float32), dtype=np.float32)
    rec = np.einsum("nij,nj,nkj->nik", U.astype(np.float64), sig, V)
    np.testing.assert_allclose(rec, F.astype(np.float32), atol=5e-5)
    lap = np.linalg.svd(F.astype(np.float32).astype(np.float64), compute_uv=False)
    np.testing.assert_allclose(np.abs(sig), lap, rtol=1e-4, atol=1e-5)


# ---------------------------------------------------------------- StVK-Hencky
def test_stvk_energy(oracle):
    m, mu, lam = mat(0)
    sig = np.array([[1.0, 1.0, 1.0], [1.1, 0.9, 1.05], [0.8, 1.2, 1.3]])
    _, _, psi, dpsi = oracle.return_map(m, sig, 0.0, 1e-3)
    e = np.log(sig)
    tr = e.sum(axis=1)
    np.testing.assert_allclose(psi, mu * (e * e).sum(axis=1) + 0.5 * lam * tr * tr, rtol=1e-6)
    assert psi[0] == 0 and not dpsi[0].any()
    # finite difference of psi in each singular value
    for k in range(3):
        d = np.zeros(3)
        d[k] = 1e-6
        pp = oracle.return_map(m, sig + d, 0.0, 1e-3)[2]
        pm = oracle.return_map(m, sig - d, 0.0, 1e-3)[2]
        np.testing.assert_allclose(dpsi[1:, k], ((pp - pm) / 2e-6)[1:], rtol=1e-5)


# ---------------------------------------------------------------- return maps
def _rand_sig(n, spread, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(scale=spread, size=(n, 3)))


def test_drucker_prager(oracle):
    m, mu, lam = mat(1, p0=scenes.dp_alpha(30), p1=0.0)
    alpha = float(m.p0)
    sig = _rand_sig(3000, 0.05, 5)
    out = oracle.return_map(m, sig, 0.0, 1e-3)[0]
    e, eo = np.log(sig), np.log(out)
    tr = e.sum(axis=1)
    dev = e - tr[:, None] / 3
    dg = np.linalg.norm(dev, axis=1) + alpha * (3 * lam + 2 * mu) / (2 * mu) * tr
    tension = tr > 0
    elastic = (~tension) & (dg <= 0)
    cone = (~tension) & (dg > 0)
    assert tension.any() and elastic.any() and cone.any()
    np.testing.assert_array_equal(out[tension], 1.0)
    np.testing.assert_array_equal(out[elastic], sig[elastic])
    tro = eo[cone].sum(axis=1)
    np.testing.assert_allclose(tro, tr[cone], atol=1e-12)                      # volume kept
    devo = eo[cone] - tro[:, None] / 3
    on_cone = np.linalg.norm(devo, axis=1) + alpha * (3 * lam + 2 * mu) / (2 * mu) * tro
    np.testing.assert_allclose(on_cone, 0.0, atol=1e-12)                       # on the yield cone


def test_von_mises(oracle):
    m, mu, lam = mat(3, p0=50.0)
    sigY = float(m.p0)
    sig = _rand_sig(3000, 0.01, 6)
    out = oracle.return_map(m, sig, 0.0, 1e-3)[0]
    e, eo = np.log(sig), np.log(out)
    dn = np.linalg.norm(e - e.mean(axis=1, keepdims=True), axis=1)
    yielded = dn > sigY / (2 * mu)
    assert yielded.any() and (~yielded).any()
    np.testing.assert_array_equal(out[~yielded], sig[~yielded])
    dno = np.linalg.norm(eo - eo.mean(axis=1, keepdims=True), axis=1)
    np.testing.assert_allclose(dno[yielded], sigY / (2 * mu), rtol=1e-9)
    np.testing.assert_allclose(eo.sum(axis=1), e.sum(axis=1), atol=1e-12)


@pytest.mark.parametrize("hexp", [1.0, 0.5, 2.0])
def test_herschel_bulkley(oracle, hexp):
    dt = 1e-3
    m, mu, lam = mat(4, E=2250.0, nu=0.125, p0=20.0, p1=hexp, p2=25.0)
    sigY, hexp, eta = float(m.p0), float(m.p1), float(m.p2)
    sig = _rand_sig(2000, 0.03, 7)
    out = oracle.return_map(m, sig, 0.0, dt)[0]
    J = sig.prod(axis=1)
    np.testing.assert_allclose(out.prod(axis=1), J, rtol=1e-9)                # volume kept
    bbar = sig ** 2 * J[:, None] ** (-2 / 3)
    trb = bbar.sum(axis=1)
    devb = bbar - trb[:, None] / 3
    sPre = mu * np.linalg.norm(devb, axis=1)
    sY = math.sqrt(2 / 3) * sigY
    yielded = sPre > sY
    assert yielded.any() and (~yielded).any()
    np.testing.assert_array_equal(out[~yielded], sig[~yielded])
    # recover s_new from the output: the deviator/trace ratio of b_bar survives the det normalisation
    bo = out ** 2 * J[:, None] ** (-2 / 3)
    ratio = np.linalg.norm(bo - bo.sum(axis=1, keepdims=True) / 3, axis=1) / bo.sum(axis=1)
    s_out = sPre * ratio * trb / np.linalg.norm(devb, axis=1)
    muT = trb / 3 * mu
    for i in np.nonzero(yielded)[0][:200]:
        g = lambda s: eta ** (1 / hexp) * (s - sPre[i]) + 2 * muT[i] * dt * max(s - sY, 0.0) ** (1 / hexp)
        root = optimize.brentq(g, sY, sPre[i], xtol=1e-14)
        tol = (sPre[i] - sY) * 2.0 ** -17 if abs(hexp - 1) > 1e-3 else 1e-9 * sPre[i]
        assert abs(s_out[i] - root) <= tol + 1e-9 * sPre[i]


def test_snow_clamp_and_hardening(oracle):
    m, mu, lam = mat(5, p0=3.5e-2, p1=8e-3, p2=10.0)
    thC, thS, xi = float(m.p0), float(m.p1), float(m.p2)
    sig = _rand_sig(2000, 0.03, 8)
    state = np.random.default_rng(9).uniform(0.8, 1.2, 2000)
    out, st, _, _ = oracle.return_map(m, sig, state, 1e-3)
    assert (out >= 1 - thC - 1e-15).all() and (out <= 1 + thS + 1e-15).all()
    np.testing.assert_array_equal(out, np.clip(sig, 1 - thC, 1 + thS))
    Jp = state * sig.prod(axis=1) / out.prod(axis=1)
    np.testing.assert_allclose(st, np.clip(Jp, 0.2, 2.0), rtol=1e-12)
    # the harden flag scales mu, lam by exp(min(xi (1 - Jp), 5)): Psi scales by the same factor
    psi1 = oracle.return_map(m, sig, state, 1e-3, harden=False)[2]
    psi2 = oracle.return_map(m, sig, state, 1e-3, harden=True)[2]
    np.testing.assert_allclose(psi2, psi1 * np.exp(np.minimum(xi * (1 - state), 5)), rtol=1e-9)


def test_nacc(oracle):
    m, mu, lam = mat(2, E=2000.0, p0=-0.008, p1=0.3, p2=5.0, p3=1.85)
    beta, xi, M = float(m.p1), float(m.p2), float(m.p3)
    kappa = 2 / 3 * mu + lam
    sig = _rand_sig(4000, 0.04, 10)
    alpha = np.random.default_rng(11).uniform(-0.2, 0.1, 4000)
    out, aout, _, _ = oracle.return_map(m, sig, alpha, 1e-3)
    a = np.clip(alpha, -0.25, 0.4)
    p0 = kappa * (1e-5 + np.sinh(np.minimum(xi * np.maximum(-a, 0), 4)))
    J = sig.prod(axis=1)
    p = -0.5 * kappa * (J * J - 1)
    c1, c2 = p > p0, p < -beta * p0
    assert c1.any() and c2.any()
    # tips: the output pressure is exactly the tip pressure
    Jo = out.prod(axis=1)
    po = -0.5 * kappa * (Jo * Jo - 1)
    np.testing.assert_allclose(po[c1], p0[c1], rtol=1e-8, atol=1e-8 * kappa)
    np.testing.assert_allclose(po[c2], -beta * p0[c2], rtol=1e-8, atol=1e-8 * kappa)
    np.testing.assert_allclose(aout[c1 | c2], (a + np.log(J / Jo))[c1 | c2], rtol=1e-9, atol=1e-12)
    # case 3: q back on the ellipse at the trial pressure (q from the output deviator with trial J)
    b2 = sig ** 2
    sHat = mu * J[:, None] ** (-2 / 3) * (b2 - b2.sum(axis=1, keepdims=True) / 3)
    q = math.sqrt(1.5) * np.linalg.norm(sHat, axis=1)
    bq = 1 + 2 * beta
    y = bq * q * q + M * M * (p + beta * p0) * (p - p0)
    c3 = (~c1) & (~c2) & (y > 1e-4)
    assert c3.any()
    bo = out ** 2
    qo = math.sqrt(1.5) * mu * J ** (-2 / 3) * np.linalg.norm(bo - bo.sum(axis=1, keepdims=True) / 3, axis=1)
    yo = bq * qo * qo + M * M * (p + beta * p0) * (p - p0)
    scale = M * M * np.abs((p + beta * p0) * (p - p0)) + 1.0
    np.testing.assert_allclose(yo[c3] / scale[c3], 0.0, atol=1e-8)
    np.testing.assert_allclose(bo.sum(axis=1)[c3], b2.sum(axis=1)[c3], rtol=1e-12)  # tr(b) kept
    inside = (~c1) & (~c2) & (y <= 1e-4)
    np.testing.assert_array_equal(out[inside], sig[inside])
