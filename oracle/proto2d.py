"""CPU oracle of the 2D fp64 prototype -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker of the library's 2D engine; the product never imports it.

A numpy float64 restatement of prototype/xpbi2d.py:29-241 (velocity-based XPBD
with Jacobi iterations, Wendland C2 in 2D, Bonet-Lok corrected gradients, StVK
Hencky with C = sqrt(2 Psi), Drucker-Prager return map in the loop, the 0.75h
position correction, box walls with friction, XSPH, the final F update and the
clipped position update).  The prototype is a script with hard-coded constants;
here they are parameters (``prototype_params`` are its own values,
``column40x80_params`` the 40x80 column at r = 0.01 m) and the state is explicit.

Where the restatement fixes what the prototype leaves to a library, it says so:

* neighbour pairs: every ordered (p, b), b != p, with |x_p - x_b|^2 <= (2h)^2
  on x^n (the prototype's cKDTree ball query, xpbi2d.py:125-129), in the order
  (p ascending, b ascending), which fixes the order of every sum;
* the 2x2 SVD: ``svd2`` below, a symmetric Jacobi rotation of A^T A (Golub &
  Van Loan, Matrix Computations, section 8.5.2) with sigma_2 = |det A| / sigma_1,
  U a rotation times sign(det A) and V a rotation (the prototype calls LAPACK
  through numpy.linalg.svd, whose sign conventions differ but whose products
  U diag(f(S)) V^T are the same);
* det F is F00 F11 - F01 F10 (numpy.linalg.det goes through LU);
* t**4 and t**3 of the kernel are products (t*t)*(t*t) and (t*t)*t.

Every product and sum is written in the order the library's kernel evaluates
it, so the two differ only where numpy's and the device's log/exp round apart.
"""
import math

import numpy as np

DIM = 2


# ---------------------------------------------------------------- parameters

def make_params(h, E, nu, rho, phi_deg, dt, iterations, gravity, xsph_c, domain, omega=1.0):
    """Derived constants of xpbi2d.py:29-49."""
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    phi = math.radians(phi_deg)
    alpha_dp = math.sqrt(2.0 / 3.0) * 2.0 * math.sin(phi) / (3.0 - math.sin(phi))
    V0 = h * h
    return dict(h=float(h), support=2.0 * h, E=float(E), nu=float(nu), rho=float(rho), phi_deg=float(phi_deg),
                dt=float(dt), iterations=int(iterations), gravity=np.asarray(gravity, np.float64),
                xsph_c=float(xsph_c), domain=np.asarray(domain, np.float64), omega=float(omega),
                mu=mu, lam=lam, alpha_dp=alpha_dp, V0=V0, mass=rho * V0,
                W_SIG=7.0 / (4.0 * np.pi * h * h))


def prototype_params():
    """The prototype's own constants (xpbi2d.py:29-49): dx = 0.008, E = 1e4, rho = 1000,
    phi = 35 deg, dt = 2e-4, 8 iterations, no relaxation of Delta-lambda, box 1.0 x 0.6."""
    return make_params(0.008, 1e4, 0.3, 1000.0, 35.0, 2e-4, 8, [0.0, -9.8], 0.01, [1.0, 0.6])


def column40x80_params():
    """The 40 x 80 column at r = 0.01 m: E = 3.5e5 Pa, nu = 0.3, rho = 2000, phi = 30 deg,
    g = -9.81, dt = 1/3000 s, 20 Jacobi iterations with Delta-lambda under-relaxed by 0.5,
    in a 2.0 x 1.0 m box."""
    return make_params(0.01, 3.5e5, 0.3, 2000.0, 30.0, 1.0 / 3000.0, 20, [0.0, -9.81], 0.01, [2.0, 1.0],
                       omega=0.5)


def column(nx, ny, dx, x0, y0, seed):
    """Lattice column with U(-0.05 dx, 0.05 dx) jitter from default_rng(seed)
    (xpbi2d.py:52-56: rows of x first, then y; jitter drawn for all coordinates at once)."""
    xs, ys = np.meshgrid(np.arange(nx) * dx + x0, np.arange(ny) * dx + y0)
    x = np.stack([xs.ravel(), ys.ravel()], axis=1)
    rng = np.random.default_rng(seed)
    return x + rng.uniform(-0.05 * dx, 0.05 * dx, x.shape)


def prototype_state():
    """The prototype's initial column (xpbi2d.py:52-58): 18 x 37 particles."""
    dx = 0.008
    return initial_state(column(int(0.15 / dx), int(0.30 / dx), dx, 0.425, dx, 0))


def column40x80_state(seed=0):
    """40 x 80 column at r = 0.01 m, centred in x on the 2.0 m floor, first row at y = r."""
    r = 0.01
    return initial_state(column(40, 80, r, 1.0 - 20 * r + 0.5 * r, r, seed))


def initial_state(x):
    N = len(x)
    return dict(x=np.array(x, np.float64), v=np.zeros((N, 2)), F=np.tile(np.eye(2), (N, 1, 1)), lam=np.zeros(N))


# ---------------------------------------------------------------- pieces

def pairs(x, support):
    """Ordered pairs (P, B), b != p, |x_p - x_b|^2 <= support^2, sorted by (p, b).
    Candidates come from the 3x3 cells of side `support` around each particle."""
    N = len(x)
    cs = support
    c = np.floor(x / cs).astype(np.int64)
    c -= c.min(axis=0)
    ncx = int(c[:, 0].max()) + 1
    key = c[:, 1] * ncx + c[:, 0]
    order = np.argsort(key, kind="stable")
    skey = key[order]
    Ps, Bs = [], []
    for dy in (-1, 0, 1):
        for dxo in (-1, 0, 1):
            cx, cy = c[:, 0] + dxo, c[:, 1] + dy
            ok = (cx >= 0) & (cx < ncx) & (cy >= 0)
            k2 = cy * ncx + cx
            lo = np.searchsorted(skey, k2, "left")
            hi = np.searchsorted(skey, k2, "right")
            cnt = np.where(ok, hi - lo, 0)
            p = np.repeat(np.arange(N), cnt)
            within = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            b = order[np.repeat(lo, cnt) + within]
            Ps.append(p)
            Bs.append(b)
    P, B = np.concatenate(Ps), np.concatenate(Bs)
    d0 = x[P, 0] - x[B, 0]
    d1 = x[P, 1] - x[B, 1]
    keep = (P != B) & (d0 * d0 + d1 * d1 <= support * support)
    P, B = P[keep], B[keep]
    s = np.lexsort((B, P))
    return P[s], B[s]


def kernel_W(prm, r):
    """Wendland C2 in 2D, support 2h (xpbi2d.py:64-72)."""
    h = prm["h"]
    q = r / h
    t = 1.0 - 0.5 * q
    t2 = t * t
    return np.where(q < 2.0, prm["W_SIG"] * (t2 * t2) * (2.0 * q + 1.0), 0.0)


def kernel_gradW(prm, dxv, r):
    """grad_{x_p} W(|x_p - x_b|) for dxv = x_p - x_b (xpbi2d.py:74-88)."""
    h = prm["h"]
    q = r / h
    t = 1.0 - 0.5 * q
    dwdq = prm["W_SIG"] * (-5.0 * q) * ((t * t) * t)
    m = (q < 2.0) & (r > 1e-12)
    coef = np.where(m, dwdq / (h * np.where(m, r, 1.0)), 0.0)
    return dxv * coef[:, None]


def svd2(A):
    """Batched 2x2 SVD A = U diag(S) V^T, S descending and >= 0, det V = +1,
    det U = sign(det A) (+1 for det A = 0).  Returns U, S, V (V, not V^T)."""
    a00, a01, a10, a11 = A[:, 0, 0], A[:, 0, 1], A[:, 1, 0], A[:, 1, 1]
    m00 = a00 * a00 + a10 * a10
    m01 = a00 * a01 + a10 * a11
    m11 = a01 * a01 + a11 * a11
    nz = m01 != 0.0
    m01s = np.where(nz, m01, 1.0)
    tau = (m11 - m00) / (2.0 * m01s)
    t = np.where(tau >= 0.0, 1.0, -1.0) / (np.abs(tau) + np.sqrt(1.0 + tau * tau))
    t = np.where(nz, t, 0.0)
    c = 1.0 / np.sqrt(1.0 + t * t)
    s = t * c
    e0 = m00 - t * m01
    e1 = m11 + t * m01
    sw = e1 > e0
    # columns v0 = (c, -s), v1 = (s, c); when e1 > e0: v0 <- v1, v1 <- -v0
    v00 = np.where(sw, s, c)
    v10 = np.where(sw, c, -s)
    v01 = np.where(sw, -c, s)
    v11 = np.where(sw, s, c)
    w0 = a00 * v00 + a01 * v10
    w1 = a10 * v00 + a11 * v10
    s0 = np.sqrt(w0 * w0 + w1 * w1)
    det = a00 * a11 - a01 * a10
    pos = s0 > 0.0
    s0s = np.where(pos, s0, 1.0)
    u00 = np.where(pos, w0 / s0s, 1.0)
    u10 = np.where(pos, w1 / s0s, 0.0)
    s1 = np.where(pos, np.abs(det) / s0s, 0.0)
    sg = np.where(det < 0.0, -1.0, 1.0)
    u01 = sg * -u10
    u11 = sg * u00
    U = np.stack([np.stack([u00, u01], -1), np.stack([u10, u11], -1)], -2)
    V = np.stack([np.stack([v00, v01], -1), np.stack([v10, v11], -1)], -2)
    return U, np.stack([s0, s1], -1), V


def usv(U, S, V):
    """U diag(S) V^T with each entry (U_i0 S_0) V_k0 + (U_i1 S_1) V_k1."""
    out = np.empty(U.shape)
    for i in range(2):
        for k in range(2):
            out[:, i, k] = (U[:, i, 0] * S[:, 0]) * V[:, k, 0] + (U[:, i, 1] * S[:, 1]) * V[:, k, 1]
    return out


def mat2(A, B):
    """A B, entry (i, k) = A_i0 B_0k + A_i1 B_1k."""
    out = np.empty(A.shape)
    for i in range(2):
        for k in range(2):
            out[:, i, k] = A[:, i, 0] * B[:, 0, k] + A[:, i, 1] * B[:, 1, k]
    return out


def matvec(A, g):
    return np.stack([A[:, 0, 0] * g[:, 0] + A[:, 0, 1] * g[:, 1], A[:, 1, 0] * g[:, 0] + A[:, 1, 1] * g[:, 1]], -1)


def drucker_prager(prm, S):
    """Return map on the singular values (xpbi2d.py:94-110)."""
    mu, lam, d = prm["mu"], prm["lam"], float(DIM)
    eps = np.log(np.maximum(S, 1e-10))
    tr = eps[:, 0] + eps[:, 1]
    dev = eps - (tr / d)[:, None]
    dev_norm = np.sqrt(dev[:, 0] * dev[:, 0] + dev[:, 1] * dev[:, 1])
    dg = dev_norm + ((prm["alpha_dp"] * (d * lam + 2.0 * mu)) / (2.0 * mu)) * tr
    case2 = tr > 0.0
    case3 = (~case2) & (dg > 0.0) & (dev_norm > 1e-12)
    scale = 1.0 - dg / np.where(case3, dev_norm, 1.0)
    out = np.where(case3[:, None], dev * scale[:, None] + (tr / d)[:, None], eps)
    out = np.where(case2[:, None], 0.0, out)
    return np.exp(out)


def stvk(prm, S):
    """Hencky StVK Psi and dPsi/dSigma (xpbi2d.py:112-118)."""
    mu, lam = prm["mu"], prm["lam"]
    Sc = np.maximum(S, 1e-10)
    lg = np.log(Sc)
    tr = lg[:, 0] + lg[:, 1]
    psi = mu * (lg[:, 0] * lg[:, 0] + lg[:, 1] * lg[:, 1]) + (0.5 * lam) * (tr * tr)
    dpsi = ((2.0 * mu) * lg + (lam * tr)[:, None]) / Sc
    return psi, dpsi


def _grad_v(prm, st, P, B, gt, Vn):
    """sum_b (V_b (v_b - v_p)_i) gt_j (xpbi2d.py:157-160)."""
    N = len(Vn)
    v = st["v"]
    dv = v[B] - v[P]
    gv = np.zeros((N, 2, 2))
    for i in range(2):
        for j in range(2):
            np.add.at(gv[:, i, j], P, (Vn[B] * dv[:, i]) * gt[:, j])
    return gv


def _f_trial(prm, gv, F):
    M = np.empty(gv.shape)
    dt = prm["dt"]
    for i in range(2):
        for j in range(2):
            M[:, i, j] = (1.0 if i == j else 0.0) + dt * gv[:, i, j]
    return mat2(M, F)


def bonet_lok(prm, x, P, B, Vn):
    """L_p = pinv(sum_b V_b gradW (x) (x_b - x_p)) (xpbi2d.py:142-148); also the pair gradients."""
    N = len(x)
    dxv = x[P] - x[B]
    r = np.sqrt(dxv[:, 0] * dxv[:, 0] + dxv[:, 1] * dxv[:, 1])
    gW = kernel_gradW(prm, dxv, r)
    A = np.zeros((N, 2, 2))
    for i in range(2):
        for j in range(2):
            np.add.at(A[:, i, j], P, (Vn[B] * gW[:, i]) * (-dxv[:, j]))
    U, S, V = svd2(A)
    thr = 1e-6 * np.maximum(S[:, 0], 1e-12)
    Sinv = np.where(S > thr[:, None], 1.0 / np.where(S > thr[:, None], S, 1.0), 0.0)
    L = usv(V, Sinv, U)  # V diag(Sinv) U^T
    return L, dxv, r, gW


def substep(prm, st, collect=None):
    """One substep of xpbi2d.py:123-241 on the state dict (x, v, F, lam), in place.
    With ``collect`` a dict, records the first iteration's Delta-lambda (d0) and
    the substep's final lambda per particle."""
    x, F = st["x"], st["F"]
    N = len(x)
    dt, h, mass, V0 = prm["dt"], prm["h"], prm["mass"], prm["V0"]
    P, B = pairs(x, prm["support"])
    Vn = V0 * (F[:, 0, 0] * F[:, 1, 1] - F[:, 0, 1] * F[:, 1, 0])
    L, dxv, r, gW = bonet_lok(prm, x, P, B, Vn)
    Wv = kernel_W(prm, r)
    gt = matvec(L[P], gW)
    st["v"] = st["v"] + dt * prm["gravity"][None, :]
    lam_p = np.zeros(N)
    alpha_t = (1.0 / V0) / (dt * dt)
    mdt = mass * dt
    dom = prm["domain"]
    for it in range(prm["iterations"]):
        gv = _grad_v(prm, st, P, B, gt, Vn)
        Ftr = _f_trial(prm, gv, F)
        U, S, V = svd2(Ftr)
        Sp = drucker_prager(prm, S)
        psi, dpsi = stvk(prm, Sp)
        C = np.sqrt(np.maximum(2.0 * psi, 0.0))
        Cc = np.maximum(C, 1e-8)
        Pk = usv(U, dpsi, V) / Cc[:, None, None]
        G = np.empty(Pk.shape)
        for i in range(2):
            for k in range(2):
                G[:, i, k] = Pk[:, i, 0] * F[:, k, 0] + Pk[:, i, 1] * F[:, k, 1]  # Pk F^T
        gC = Vn[B][:, None] * matvec(G[P], gt)
        gCp = np.zeros((N, 2))
        for i in range(2):
            np.add.at(gCp[:, i], P, -gC[:, i])
        denom = np.zeros(N)
        np.add.at(denom, P, (gC[:, 0] * gC[:, 0] + gC[:, 1] * gC[:, 1]) / mass)
        denom = denom + (gCp[:, 0] * gCp[:, 0] + gCp[:, 1] * gCp[:, 1]) / mass
        dlam = (-C - alpha_t * lam_p) / (denom + alpha_t)
        dlam = np.where(C > 1e-9, prm["omega"] * dlam, 0.0)
        if collect is not None and it == 0:
            collect["d0"] = dlam.copy()
        lam_p = lam_p + dlam
        dva = np.zeros((N, 2))
        for i in range(2):
            np.add.at(dva[:, i], B, (gC[:, i] * dlam[P]) / mdt)
        dva = dva + (gCp * dlam[:, None]) / mdt
        st["v"] = st["v"] + dva
        # position correction on candidate positions (xpbi2d.py:189-201)
        v = st["v"]
        xc = x + dt * v
        dxc = xc[P] - xc[B]
        dc = np.sqrt(dxc[:, 0] * dxc[:, 0] + dxc[:, 1] * dxc[:, 1])
        Cd = dc - (h - 0.25 * h)
        act = Cd < 0.0
        n_pb = np.where((dc > 1e-12)[:, None], dxc / np.maximum(dc, 1e-12)[:, None], 0.0)
        corr = np.where(act[:, None], (-0.5 * Cd)[:, None] * n_pb, 0.0)
        dvp = np.zeros((N, 2))
        for i in range(2):
            np.add.at(dvp[:, i], P, corr[:, i])
            np.add.at(dvp[:, i], B, -corr[:, i])
        v = v + (dvp / dt) * 0.5
        # walls (xpbi2d.py:204-220)
        xc = x + dt * v
        for d in range(2):
            lo = xc[:, d] < h
            hi = xc[:, d] > dom[d] - h
            td = 1 - d
            for m, target in ((lo, h), (hi, dom[d] - h)):
                vn = (target - x[:, d]) / dt
                dvn = vn - v[:, d]
                vt = v[:, td]
                fric = np.maximum(0.0, 1.0 - (0.4 * np.abs(dvn)) / (np.abs(vt) + 1e-9))
                v[:, td] = np.where(m, vt * fric, vt)
                v[:, d] = np.where(m, vn, v[:, d])
        st["v"] = v
    if collect is not None:
        collect["lam"] = lam_p.copy()
    st["lam"] = lam_p
    # XSPH (xpbi2d.py:222-225)
    v = st["v"]
    dv = v[B] - v[P]
    dvx = np.zeros((N, 2))
    for i in range(2):
        np.add.at(dvx[:, i], P, (Vn[B] * dv[:, i]) * Wv)
    v = v + prm["xsph_c"] * dvx
    st["v"] = v
    # F update with the return map (xpbi2d.py:227-233)
    gv = _grad_v(prm, st, P, B, gt, Vn)
    U, S, V = svd2(_f_trial(prm, gv, F))
    st["F"] = usv(U, drucker_prager(prm, S), V)
    # positions (xpbi2d.py:235-237)
    st["x"] = np.clip(x + dt * v, 0.5 * h, dom - 0.5 * h)
    return st


def run(prm, st, substeps):
    for _ in range(substeps):
        substep(prm, st)
    return st
