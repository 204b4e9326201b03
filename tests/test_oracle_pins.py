"""Closed-form and first-principles pins of the oracle's solver functions (no device).

The reference has no tests or vectors for its solver kernels, so each function of
the CPU oracle that the whole-substep invariants leave open is pinned here by what
mathematics or physics fixes, on configurations built so that the quantity is
isolated:

* Bonet-Lok correction (computeL, simkernels.js:189-221) and the velocity gradient
  (simkernels.js:224-245): for a linear velocity field v = G x the corrected
  gradient is G P, P the projection onto the right singular vectors of
  A = sum_b V_b gradW(x_i - x_b) (x) (x_b - x_i) whose singular values exceed 0.2
  (A and its SVD from numpy/LAPACK, the kernel gradient pinned in
  test_oracle_math); a particle 1.1 spacings above a block keeps one of its three
  directions;
* the damped XPBD multiplier (simkernels.js:255-308) and the impulse it applies
  (solveB, simkernels.js:315-366): for one active constraint among neighbours of
  zero stiffness, the first Jacobi step from lambda = 0 is
  d-lambda = -(C + gamma dC/dv . v) / ((1 + gamma) sum_k |dC/dv_k|^2 / (m_k dt^2) + alpha~)
  with alpha~ = 1 / (V0 dt^2) (the compliance of the elastic energy V0 Psi =
  V0 C^2 / 2; Macklin, Mueller and Chentanez 2016, damped XPBD), and every velocity
  changes by d-lambda dC/dv_k / (m_k dt^2); the derivatives dC/dv_k are central
  finite differences of the oracle's own C (features row 14), not its formula;
* walls (simkernels.js:369-397): a particle crossing the floor leaves with
  v_n = (bound - x_n) / dt and its tangential speed cut by
  min(friction |delta v_n|, |v_t|);
* the kinematic sphere (simkernels.js:398-404): inside it the relative outward
  speed becomes min((R - d) / dt, 0.6), tangential motion untouched;
* contacts (simkernels.js:341-355): an isolated pair closer than 0.75 h is pushed
  apart by (0.75 h - d) shared as m_b / (m_a + m_b), so that momentum is kept;
* PBF (simkernels.js:407-477): with J the Jacobian of the density constraints
  C_j = rho_j / rho0 - 1 (spiky gradients, the PBF convention), every compressed
  particle has lambda_j (|J_j|^2 + eps) = -C_j and the position correction is
  J^T lambda; s_corr of a pair at distance q H is -k ((1 - q^2) / 0.96)^12;
* XSPH (simkernels.js:480-507): an isolated pair's velocities move towards each
  other by c V_b (v_b - v_i) W(r), c ten times larger for water, and a solid and a
  water particle do not interact.

test_oracle_mutations.py plants plausible mistakes in a copy of the oracle and
checks that each fails one of these pins.
"""
import numpy as np
import pytest

from xpbi import scenes
from helpers import F3, mats_record, params_record

H = 0.01  # particle spacing h of every configuration here; support 2h
HK = float(__import__("numpy").float32(H))  # h as the substep sees it (an f32 parameter)
E0 = dict(kind=0, E=0.0, nu=0.3, density=1000.0, boundFriction=0.4)  # elastic, zero stiffness


def _scene(pos, mat, materials, vel=None, dom=(0.6, 0.6, 0.6), **solver):
    s = scenes.base(H)
    s.update(domainMin=[0.0, 0.0, 0.0], domainMax=list(dom), gravity=[0.0, 0.0, 0.0], dt=1e-3, xsphC=0.0,
             iterations=1)
    s.update(solver)
    mat = np.asarray(mat, np.uint32)
    s.update(positions=np.asarray(pos, np.float32), matIds=mat, tags=mat.astype(np.float32),
             states=np.zeros(mat.size, np.float32), materials=materials, hasWater=False)
    if vel is not None:
        s["velocities"] = np.asarray(vel, np.float32)
    return s


def _run(oracle, s, iters, stop, water=False, F_of=None, vel=None, collect=False):
    """f64 oracle substep from the scene's initial state; F_of: {particle: 3x3 F}."""
    P, C = params_record(s)
    st = oracle.initial_state(s, C).copy(np.float64)
    if vel is not None:
        st.velA[:, :3] = vel
    for i, F in (F_of or {}).items():
        st.F[i] = np.concatenate([np.append(F[:, j], 0.0) for j in range(3)])
    feat = np.zeros((st.N, 16)) if collect else None
    oracle.substep(st, P, mats_record(s), iters, water, stop=stop, nthreads=1, collect=collect, feat=feat)
    return st, P, feat


def _row(st, pid):
    return int(np.where(st.id == pid)[0][0])


def _lattice(n, rng, origin=0.2):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * H
    return g + origin + rng.uniform(-0.1 * H, 0.1 * H, g.shape)


def _cap_matrix(oracle, xs, Vb, i):
    """A of particle i (numpy, kernel gradient from the pinned oracle kernel)."""
    d = xs[i] - xs
    r = np.linalg.norm(d, axis=1)
    m = (r < 2 * HK) & (np.arange(len(xs)) != i)
    _, gW, _, _ = oracle.kernels(HK, r[m], d[m])
    return np.einsum("k,ka,kb->ab", Vb[m], gW, -d[m])


# ---------------------------------------------------------------- Bonet-Lok
def pin_bonet_lok_linear_field(oracle):
    rng = np.random.default_rng(0)
    x = _lattice(8, rng)
    x = np.vstack([x, [[0.2 + 3.5 * H, 0.2 + 7 * H + 1.1 * H, 0.2 + 3.5 * H]]])  # 1.1 h above the top face
    G = np.array([[0.3, -0.5, 0.2], [0.1, 0.4, -0.6], [0.7, 0.2, -0.1]])
    s = _scene(x, np.zeros(len(x)), [E0])
    x32 = x.astype(np.float32).astype(np.float64)
    st, P, _ = _run(oracle, s, 0, "UPDATE_F", vel=x32 @ G.T)  # exactly linear in the f64 state
    xs, Vb = st.pos[:, :3], st.velA[:, 3]
    dt = float(np.float32(P.dt))
    Fm = F3(st.F).reshape(-1, 3, 3).transpose(0, 2, 1)  # row-major F per particle
    kept = []
    for i in range(st.N):
        A = _cap_matrix(oracle, xs, Vb, i)
        _, sig, Vh = np.linalg.svd(A)
        Pk = sum(np.outer(Vh[k], Vh[k]) for k in range(3) if sig[k] > 0.2) if (sig > 0.2).any() else np.zeros((3, 3))
        kept.append(int((sig > 0.2).sum()))
        grad = (Fm[i] - np.eye(3)) / dt
        err = np.abs(grad - G @ Pk).max()
        assert err <= 1e-7 * np.abs(G).max(), f"particle {i}: grad v {grad} vs G P {G @ Pk} (sigma {sig})"
    top = _row(st, len(x) - 1)
    assert kept[top] == 1, f"the particle above the block keeps {kept[top]} directions"
    assert kept.count(3) > 0.9 * st.N


# ---------------------------------------------------------------- damped XPBD
def pin_xpbd_first_step(oracle, kdamp):
    rng = np.random.default_rng(1)
    x = _lattice(6, rng)
    c = int(np.argmin(np.linalg.norm(x - x.mean(0), axis=1)))
    mat = np.zeros(len(x))
    mat[c] = 1
    E1 = dict(kind=0, E=2e5, nu=0.0, density=1000.0)
    eps = rng.normal(0, 1e-3, (3, 3))
    Fc = {c: np.eye(3) + 0.5 * (eps + eps.T)}
    v = rng.uniform(-0.005, 0.005, x.shape)
    s = _scene(x, mat, [E0, E1], vel=v, omega=1.0, kDamp=kdamp)
    st0, P, feat = _run(oracle, s, 0, "SOLVE", F_of=Fc, collect=True)
    i = _row(st0, c)
    C0 = feat[i, 14]
    assert C0 > 0.1
    dt, V0 = float(np.float32(P.dt)), float(np.float32(P.V0))
    mr = mats_record(s)
    mass = np.array([mr[0].mass, mr[1].mass], np.float64)[mat.astype(int)]
    # dC/dv_k by central differences of C (features row 14) in the initial order
    d = np.linalg.norm(x - x[c], axis=1)
    near = np.where(d < 2.2 * H)[0]
    grad = np.zeros_like(x)
    dv = 1e-6
    for k in near:
        for a in range(3):
            Cs = []
            for sgn in (1, -1):
                vp = v.copy()
                vp[k, a] += sgn * dv
                st, _, f = _run(oracle, dict(s, velocities=vp.astype(np.float32)), 0, "SOLVE", F_of=Fc,
                                collect=True)
                Cs.append(f[_row(st, c), 14])
            grad[k, a] = (Cs[0] - Cs[1]) / (2 * dv)
    v32 = v.astype(np.float32).astype(np.float64)
    alpha = 1.0 / (V0 * dt * dt)
    denom = np.sum(np.sum(grad * grad, axis=1) / (mass * dt * dt))
    dlam = -(C0 + kdamp * np.sum(grad * v32)) / ((1.0 + kdamp) * denom + alpha)
    st1, _, _ = _run(oracle, s, 1, "SOLVE", F_of=Fc)
    i1 = _row(st1, c)
    lam = st1.aux[i1, 1]
    assert abs(lam - dlam) <= 1e-5 * abs(dlam), f"lambda {lam:.6e}, damped XPBD step {dlam:.6e}"
    dv_expect = dlam * grad / (mass[:, None] * dt * dt)
    v_after = np.empty_like(v32)
    v_after[st1.id] = st1.velA[:, :3]
    err = np.abs((v_after - v32) - dv_expect).max()
    assert err <= 1e-5 * np.abs(dv_expect).max(), f"velocity change off by {err:.3e}"
    assert np.abs(dv_expect).max() > 1e-4


# ---------------------------------------------------------------- walls, sphere, contacts
def pin_floor_friction(oracle):
    for vx, fric in ((2.0, 0.4), (0.3, 0.4), (2.0, 0.0)):
        m = dict(E0, boundFriction=fric)
        x = np.array([[0.3, 0.0004, 0.3]])
        v = np.array([[vx, -1.0, 0.0]])
        s = _scene(x, [0], [m], vel=v)
        st, P, _ = _run(oracle, s, 1, "SOLVE")
        dt = float(np.float32(P.dt))
        xy = float(np.float32(0.0004))
        vn = (0.0 - xy) / dt
        dvn = abs(vn - (-1.0))
        vt = max(vx - fric * dvn, 0.0)
        got = st.velA[0, :3]
        assert abs(got[1] - vn) <= 1e-9 * abs(vn), f"normal speed {got[1]} vs {vn}"
        assert abs(got[0] - vt) <= 1e-6 * max(vx, 1.0), f"friction {fric}: tangential {got[0]} vs {vt}"
        assert got[2] == 0.0


def pin_collider(oracle):
    c, R = np.array([0.3, 0.3, 0.3]), 0.05
    for xoff, v, cv in (([0.0497, 0.0, 0.0], [-0.2, 0.1, 0.0], [0.0, 0.0, 0.0]),  # near the surface: (R - d)/dt
                        ([0.02, 0.01, 0.0], [0.0, 0.0, 0.3], [0.1, 0.0, 0.0]),    # deep: capped at 0.6
                        ([0.04, 0.0, 0.0], [1.5, 0.0, 0.0], [0.0, 0.0, 0.0])):    # inside, already leaving fast
        x = (c + np.array(xoff))[None, :]
        s = _scene(x, [0], [E0], vel=np.array([v]))
        s["collider"] = dict(pos=list(c), R=R, vel=cv)
        st, P, _ = _run(oracle, s, 1, "SOLVE")
        dt = float(np.float32(P.dt))
        x32 = x[0].astype(np.float32).astype(np.float64)
        v32 = np.array(v, np.float32).astype(np.float64)
        xq = x32 + dt * v32
        dvec = xq - c.astype(np.float32)
        dist = np.linalg.norm(dvec)
        n = dvec / dist
        goal = min((R - dist) / dt, 0.6)
        got = st.velA[0, :3]
        out_before = float(np.dot(v32 - np.array(cv, np.float32), n))
        out_after = float(np.dot(got - np.array(cv, np.float32), n))
        assert abs(out_after - max(out_before, goal)) <= 1e-6 * max(1.0, abs(goal)), (xoff, out_after, goal)
        tang = (got - v32) - np.dot(got - v32, n) * n
        assert np.abs(tang).max() <= 1e-9, "the sphere changed tangential motion"


def pin_contact_weights(oracle):
    ma = dict(E0, density=1000.0)
    mb = dict(E0, density=3000.0)
    x = np.array([[0.3, 0.3, 0.3], [0.3 + 0.5 * H, 0.3 + 0.1 * H, 0.3]])
    s = _scene(x, [0, 1], [ma, mb])
    st, P, _ = _run(oracle, s, 1, "SOLVE")
    dt, co = float(np.float32(P.dt)), float(np.float32(P.contactOmega))
    v = np.empty((2, 3))
    v[st.id] = st.velA[:, :3]
    mr = mats_record(s)
    m = np.array([mr[0].mass, mr[1].mass], np.float64)
    xs = x.astype(np.float32).astype(np.float64)
    dvec = xs[0] - xs[1]
    d = np.linalg.norm(dvec)
    push = (0.75 * float(np.float32(H)) - d) * co / dt
    exp0 = dvec / d * push * m[1] / (m[0] + m[1])
    assert np.abs(v[0] - exp0).max() <= 1e-9 * np.abs(exp0).max(), (v[0], exp0)
    assert np.abs(m[0] * v[0] + m[1] * v[1]).max() <= 1e-12 * m[1] * np.abs(v[1]).max(), "momentum"


# ---------------------------------------------------------------- PBF
W_ = dict(kind=6, E=100.0, nu=0.3, density=1000.0, boundFriction=0.0)


def _pbf(oracle, x, eps, k):
    s = _scene(x, np.zeros(len(x)), [W_], pbfEps=eps, sCorrK=k)
    s["hasWater"] = True
    st, P, _ = _run(oracle, s, 1, "SOLVE", water=True)
    mr = mats_record(s)[0]
    return st, P, float(mr.mass) / float(mr.density)  # m / rho0 as the oracle forms it


def pin_pbf_jacobian(oracle):
    rng = np.random.default_rng(4)
    x = 0.3 + rng.normal(0, 0.9 * H, (40, 3))
    for eps in (100.0, 0.0):
        st, P, V0 = _pbf(oracle, x, eps, 0.0)
        Hs = 2 * float(np.float32(H))
        dt = float(np.float32(P.dt))
        xs = st.pos[:, :3]
        N = st.N
        rho = np.zeros(N)
        J = np.zeros((N, N, 3))
        W0 = oracle.kernels(HK, np.zeros(1), np.zeros((1, 3)))[2][0]
        for j in range(N):
            d = xs[j] - xs
            r = np.linalg.norm(d, axis=1)
            m = (r < Hs) & (np.arange(N) != j)
            _, _, p6, sp = oracle.kernels(HK, r[m], d[m])
            rho[j] = V0 * (W0 + p6.sum())
            J[j, m] = -V0 * sp
            J[j, j] = V0 * sp.sum(axis=0)
        C = np.maximum(rho - 1.0, 0.0)
        lam = st.aux[:, 3]
        assert (C > 0).sum() >= 5 and (C == 0).sum() >= 5, "some particles compressed, some not"
        den = np.sum(J * J, axis=(1, 2)) + eps
        assert np.abs(lam * den + C).max() <= 1e-9 * C.max(), f"eps {eps}: lambda (|J_j|^2 + eps) != -C_j"
        dx = st.velA[:, :3] * dt
        dx_exp = np.einsum("j,jka->ka", lam, J)
        assert np.abs(dx - dx_exp).max() <= 1e-9 * np.abs(dx_exp).max(), f"eps {eps}: correction != J^T lambda"


def pin_s_corr(oracle):
    k = 0.1
    Hs = 2 * float(np.float32(H))
    for q in (0.05, 0.2, 0.5, 0.8):
        x = np.array([[0.3, 0.3, 0.3], [0.3 + q * Hs, 0.3, 0.3]])
        st, P, V0 = _pbf(oracle, x, 100.0, k)
        dt = float(np.float32(P.dt))
        assert not st.aux[:, 3].any(), "an isolated pair is below rest density"
        xs = st.pos[:, :3]
        i = _row(st, 0)
        d = xs[i] - xs[1 - i]
        r = np.linalg.norm(d)
        _, _, _, sp = oracle.kernels(HK, np.array([r]), d[None, :])
        s_exp = -float(np.float32(k)) * ((1.0 - (r / Hs) ** 2) / 0.96) ** 12
        dx = st.velA[i, :3] * dt
        coef = dx[0] / (V0 * sp[0, 0])  # the pair lies along x
        assert abs(coef - s_exp) <= 1e-9 * abs(s_exp) + 1e-15, f"q {q}: s_corr {coef} vs {s_exp}"
        assert np.abs(dx[1:]).max() <= 1e-12 * abs(dx[0])


def pin_xsph(oracle):
    """XSPH (simkernels.js:480-507) on isolated pairs 1.2 h apart: v_i gains
    c V_b (v_b - v_i) W(r), c = xsphC for solids and 10 xsphC for water, and a solid
    and a water particle leave each other alone."""
    W_ = dict(kind=6, E=100.0, nu=0.3, density=1000.0, boundFriction=0.0)
    va, vb = np.array([1.0, 0.0, 0.2]), np.array([0.0, 0.5, 0.0])
    for mats, cmul in (((E0, E0), 1.0), ((W_, W_), 10.0), ((E0, W_), 0.0)):
        x = np.array([[0.3, 0.3, 0.3], [0.3 + 1.2 * H, 0.3, 0.3]])
        s = _scene(x, [0, 1], list(mats), vel=np.array([va, vb]), xsphC=0.01)
        s["hasWater"] = mats[1]["kind"] == 6
        st, P, _ = _run(oracle, s, 0, "XSPH", water=s["hasWater"])
        V0, c = float(np.float32(P.V0)), float(np.float32(P.xsphC)) * cmul
        v = np.empty((2, 3))
        v[st.id] = st.velA[:, :3]
        xs = x.astype(np.float32).astype(np.float64)
        r = np.linalg.norm(xs[1] - xs[0])
        W = oracle.kernels(HK, np.array([r]), np.zeros((1, 3)))[0][0]
        v32 = np.array([va, vb], np.float32).astype(np.float64)
        exp = v32 + c * V0 * W * (v32[::-1] - v32)
        assert np.abs(v - exp).max() <= 1e-12, (mats[0]["kind"], mats[1]["kind"], v, exp)
        if cmul:
            assert np.abs(v - v32).max() > 1e-4, "the pair is smoothed"


PINS = {
    "xsph": pin_xsph,
    "bonet_lok": pin_bonet_lok_linear_field,
    "xpbd_undamped": lambda o: pin_xpbd_first_step(o, 0.0),
    "xpbd_damped": lambda o: pin_xpbd_first_step(o, 0.4),
    "floor_friction": pin_floor_friction,
    "collider": pin_collider,
    "contact_weights": pin_contact_weights,
    "pbf_jacobian": pin_pbf_jacobian,
    "s_corr": pin_s_corr,
}


@pytest.mark.parametrize("name", list(PINS))
def test_pin(oracle, name):
    PINS[name](oracle)
