"""Scene generators with a seeded, counter-based PRNG.

Mirrors web/js/scenes.js:7-176 (lattice samplers, material presets, solver
defaults, the eight demo scenes) with one change: the reference jitters with
unseeded ``Math.random`` (scenes.js:13-15, :31-33; sim.js:63), which no test can
reproduce, so every random number here comes from ``rand_u32(seed, stream, k)``
-- a lowbias32 hash of a counter -- which the JavaScript host computes
bit-identically with ``Math.imul`` (xpbi-webgpu_amd/js/scenes.js).

Positions are formed in float64 in the reference's operation order
(``min + (i + 0.5) * h + (u - 0.5) * jitter * h``) and rounded to f32 once.
The benchmark configurations (sand columns, the mixed dam breach, the von
Mises block) are built from the same samplers.
"""
import math

import numpy as np

MAT = dict(ELASTIC=0, DP=1, NACC=2, VM=3, HB=4, SNOW=5, WATER=6)

_M32 = 0xFFFFFFFF


def _lowbias32(x):
    """lowbias32 integer hash (C. Wellons), on uint32 numpy arrays."""
    x = np.asarray(x, dtype=np.uint32)
    x = x ^ (x >> np.uint32(16))
    x = x * np.uint32(0x7FEB352D)
    x = x ^ (x >> np.uint32(15))
    x = x * np.uint32(0x846CA68B)
    x = x ^ (x >> np.uint32(16))
    return x


def stream_key(seed, stream):
    k = _lowbias32(np.uint32(seed & _M32))
    return _lowbias32((k + np.uint32((stream * 0x632BE5AB) & _M32)).astype(np.uint32))


def rand_u32(seed, stream, counter):
    """32-bit random words for integer counters (vectorised)."""
    with np.errstate(over="ignore"):
        c = np.asarray(counter, dtype=np.uint64) & np.uint64(_M32)
        c = c.astype(np.uint32)
        key = stream_key(seed, stream)
        return _lowbias32((c * np.uint32(0x9E3779B9) + key).astype(np.uint32))


def rand_unit(seed, stream, counter):
    """Uniform [0, 1) doubles with 24 random bits (exact in f32 and f64)."""
    return (rand_u32(seed, stream, counter) >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


class Rng:
    """Sequential view of one stream: draws consume consecutive counters."""

    def __init__(self, seed, stream):
        self.seed, self.stream, self.k = seed, stream, 0

    def take(self, n):
        u = rand_unit(self.seed, self.stream, np.arange(self.k, self.k + n, dtype=np.uint64))
        self.k += n
        return u


# ---------------------------------------------------------------- samplers
def sample_box(out, mn, mx, h, mat_idx, rng, jitter=0.2, state0=0.0):
    """scenes.js:7-20 -- lattice at spacing h, x fastest, seeded jitter."""
    n = [max(1, int(math.floor((mx[i] - mn[i]) / h + 0.5))) for i in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    i, j, k = i.ravel().astype(np.float64), j.ravel().astype(np.float64), k.ravel().astype(np.float64)
    cnt = i.size
    u = rng.take(3 * cnt).reshape(cnt, 3)
    x = mn[0] + (i + 0.5) * h + (u[:, 0] - 0.5) * jitter * h
    y = mn[1] + (j + 0.5) * h + (u[:, 1] - 0.5) * jitter * h
    z = mn[2] + (k + 0.5) * h + (u[:, 2] - 0.5) * jitter * h
    out["pos"].append(np.stack([x, y, z], axis=1))
    out["mat"].append(np.full(cnt, mat_idx, dtype=np.uint32))
    out["state"].append(np.full(cnt, state0, dtype=np.float64))
    return cnt


def sample_sphere(out, c, r, h, mat_idx, rng, jitter=0.2, state0=0.0):
    """scenes.js:22-38 -- lattice points inside the ball, jitter drawn per kept point."""
    n = int(math.ceil(2 * r / h))
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    px = c[0] - r + i.ravel() * h
    py = c[1] - r + j.ravel() * h
    pz = c[2] - r + k.ravel() * h
    d = np.sqrt((px - c[0]) ** 2 + (py - c[1]) ** 2 + (pz - c[2]) ** 2)
    keep = d <= r
    px, py, pz = px[keep], py[keep], pz[keep]
    cnt = px.size
    u = rng.take(3 * cnt).reshape(cnt, 3)
    out["pos"].append(np.stack([px + (u[:, 0] - 0.5) * jitter * h, py + (u[:, 1] - 0.5) * jitter * h,
                                pz + (u[:, 2] - 0.5) * jitter * h], axis=1))
    out["mat"].append(np.full(cnt, mat_idx, dtype=np.uint32))
    out["state"].append(np.full(cnt, state0, dtype=np.float64))
    return cnt


# ---------------------------------------------------------------- presets (scenes.js:40-80)
def dp_alpha(phi_deg):
    s = math.sin(phi_deg * math.pi / 180)
    return math.sqrt(2 / 3) * 2 * s / (3 - s)


def SAND(phi=35, cohesion=0.0, E=2000.0, nu=0.3, density=1.0):
    return dict(kind=MAT["DP"], E=E, nu=nu, density=density, p0=dp_alpha(phi), p1=cohesion,
                color=[0.85, 0.63, 0.28], boundFriction=0.5)


def SNOW_NACC():
    return dict(kind=MAT["NACC"], E=2000.0, nu=0.3, density=4.0, p0=-0.008, p1=0.05, p2=5.0, p3=1.85,
                color=[0.93, 0.95, 0.98], boundFriction=0.35)


def PLASTICINE(sigY=30.0, E=1200.0, nu=0.42, density=2.0):
    return dict(kind=MAT["VM"], E=E, nu=nu, density=density, p0=sigY, color=[0.36, 0.76, 0.72],
                boundFriction=0.4)


def GOO(hExp=1.0, sigY=45.0, eta=25.0, E=2250.0, nu=0.125, density=100.0):
    return dict(kind=MAT["HB"], E=E, nu=nu, density=density, p0=sigY, p1=hExp, p2=eta,
                color=[0.91, 0.45, 0.62], boundFriction=0.2)


def SNOW_CLASSIC(thetaC=3.5e-2, thetaS=8e-3, xi=10.0, E=9000.0, nu=0.2, density=4.0):
    return dict(kind=MAT["SNOW"], E=E, nu=nu, density=density, p0=thetaC, p1=thetaS, p2=xi,
                color=[0.96, 0.97, 1.0], boundFriction=0.35)


def WATER(density=1.0):
    return dict(kind=MAT["WATER"], E=100.0, nu=0.3, density=density, color=[0.25, 0.55, 0.95],
                boundFriction=0.0)


def ELASTIC():
    return dict(kind=MAT["ELASTIC"], E=900.0, nu=0.3, density=2.0, color=[0.62, 0.52, 0.86],
                boundFriction=0.3)


def base(h):
    """scenes.js:82-92 solver defaults (camera omitted: no renderer here)."""
    return dict(h=h, domainMin=[0.0, 0.0, 0.0], domainMax=[1.0, 1.0, 1.0], gravity=[0.0, -9.8, 0.0],
                substeps=8, iterations=3, timeScale=0.42, omega=0.5, contactOmega=0.4, xsphC=0.01,
                pbfEps=100.0, sCorrK=0.0001, kDamp=0.4)


def _finish(s, out, materials, seed, has_water, velocities=None, tag_seed_stream=99, info=""):
    pos = np.concatenate(out["pos"]).astype(np.float64)
    mat = np.concatenate(out["mat"])
    states = np.concatenate(out["state"])
    n = mat.size
    # pos.w = matIdx + colour fraction in [0, 0.49) (sim.js:63), from its own stream
    frac = rand_unit(seed, tag_seed_stream, np.arange(n, dtype=np.uint64)) * 0.49
    tags = (mat.astype(np.float64) + frac).astype(np.float32)
    s = dict(s)
    s.update(positions=pos.astype(np.float32), matIds=mat, tags=tags, states=states.astype(np.float32),
             materials=materials, hasWater=has_water, info=info, seed=seed)
    if velocities is not None:
        s["velocities"] = velocities.astype(np.float32)
    return s


def _new():
    return dict(pos=[], mat=[], state=[])


# ---------------------------------------------------------------- the eight reference scenes
def scene_sand_column(seed=0):
    s = base(0.016)
    out, rng = _new(), Rng(seed, 0)
    sample_box(out, [0.36, 0.015, 0.36], [0.64, 0.62, 0.64], s["h"], 0, rng)
    return _finish(s, out, [SAND(35)], seed, False, info="Sand: column collapse")


def scene_two_blocks(seed=0):
    s = base(0.017)
    out, rng = _new(), Rng(seed, 0)
    n1 = sample_box(out, [0.08, 0.015, 0.35], [0.38, 0.42, 0.65], s["h"], 0, rng)
    n2 = sample_box(out, [0.62, 0.015, 0.35], [0.92, 0.42, 0.65], s["h"], 0, rng)
    vel = np.zeros((n1 + n2, 3))
    vel[:n1, 0] = 1.6
    vel[n1:, 0] = -1.6
    return _finish(s, out, [SAND(35)], seed, False, vel, info="Sand: two blocks collide")


def scene_snowball(seed=0):
    s = base(0.018)
    s.update(substeps=9, iterations=3, timeScale=0.45)
    m = SNOW_NACC()
    out, rng = _new(), Rng(seed, 0)
    n1 = sample_box(out, [0.1, 0.015, 0.1], [0.9, 0.14, 0.9], s["h"], 0, rng, 0.2, m["p0"])
    n2 = sample_sphere(out, [0.2, 0.55, 0.5], 0.11, s["h"], 0, rng, 0.2, m["p0"])
    vel = np.zeros((n1 + n2, 3))
    vel[n1:, 0] = 2.6
    vel[n1:, 1] = -3.2
    return _finish(s, out, [m], seed, False, vel, info="Snow: snowball smash")


def scene_plasticine(seed=0):
    s = base(0.013)
    s.update(substeps=8, iterations=6, timeScale=0.4)
    out, rng = _new(), Rng(seed, 0)
    sample_sphere(out, [0.5, 0.4, 0.5], 0.16, s["h"], 0, rng)
    return _finish(s, out, [PLASTICINE(80)], seed, False, info="Plasticine: drop & poke")


def scene_goo(seed=0):
    s = base(0.013)
    s.update(substeps=9, iterations=3, timeScale=0.5)
    out, rng = _new(), Rng(seed, 0)
    sample_sphere(out, [0.5, 0.42, 0.5], 0.17, s["h"], 0, rng)
    return _finish(s, out, [GOO(1.0)], seed, False, info="Goo: Herschel-Bulkley")


def scene_snow_classic(seed=0):
    s = base(0.013)
    s.update(substeps=10, timeScale=0.38, iterations=4)
    out, rng = _new(), Rng(seed, 0)
    sample_sphere(out, [0.35, 0.45, 0.42], 0.1, s["h"], 0, rng, 0.2, 1.0)
    sample_sphere(out, [0.65, 0.5, 0.58], 0.1, s["h"], 0, rng, 0.2, 1.0)
    pos = np.concatenate(out["pos"])
    vel = np.zeros_like(pos)
    vel[:, 0] = np.where(pos[:, 0] < 0.5, 1.5, -1.5)
    return _finish(s, out, [SNOW_CLASSIC()], seed, False, vel, info="Snow: classic (Stomakhin)")


def scene_dam_breach(seed=0):
    s = base(0.019)
    s.update(substeps=6, iterations=3, timeScale=0.45)
    out, rng = _new(), Rng(seed, 0)
    sample_box(out, [0.42, 0.015, 0.05], [0.6, 0.34, 0.95], s["h"], 0, rng)
    sample_box(out, [0.03, 0.015, 0.05], [0.36, 0.5, 0.95], s["h"], 1, rng)
    return _finish(s, out, [SAND(30, 0.004), WATER()], seed, True, info="Dam breach: water + sand")


def scene_elastic(seed=0):
    s = base(0.014)
    s.update(substeps=10, iterations=4, timeScale=0.5)
    out, rng = _new(), Rng(seed, 0)
    sample_box(out, [0.38, 0.3, 0.38], [0.62, 0.54, 0.62], s["h"], 0, rng)
    return _finish(s, out, [ELASTIC()], seed, False, info="Elastic: StVK block")


SCENES = {
    "Sand: column collapse": scene_sand_column,
    "Sand: two blocks collide": scene_two_blocks,
    "Snow: snowball smash": scene_snowball,
    "Plasticine: drop & poke": scene_plasticine,
    "Goo: Herschel-Bulkley": scene_goo,
    "Snow: classic (Stomakhin)": scene_snow_classic,
    "Dam breach: water + sand": scene_dam_breach,
    "Elastic: StVK block": scene_elastic,
}


# ---------------------------------------------------------------- benchmark workloads
def _solver(h, dt, iterations, gravity=-9.81, **kw):
    s = base(h)
    s.update(gravity=[0.0, gravity, 0.0], dt=dt, iterations=iterations)
    s.update(kw)
    return s


def DP_SAND_STIFF(phi=30.0):
    """Drucker-Prager sand of the benchmark columns: E = 3.5e5 Pa, nu = 0.3, rho = 2000."""
    return SAND(phi, 0.0, E=3.5e5, nu=0.3, density=2000.0)


def sand_column(nx, ny, nz, r, domain, seed, phi=30.0, dt=1.0 / 6000.0, iterations=20, jitter=0.1):
    """A seeded-jitter (+-jitter/2 r) DP column centred in x, z on the floor."""
    s = _solver(r, dt, iterations)
    s["domainMin"], s["domainMax"] = [0.0, 0.0, 0.0], list(domain)
    ext = [nx * r, ny * r, nz * r]
    mn = [0.5 * (domain[0] - ext[0]), 0.0, 0.5 * (domain[2] - ext[2])]
    mx = [mn[0] + ext[0], ext[1], mn[2] + ext[2]]
    out, rng = _new(), Rng(seed, 0)
    sample_box(out, mn, mx, r, 0, rng, jitter)
    return _finish(s, out, [DP_SAND_STIFF(phi)], seed, False, info=f"DP sand column {nx}x{ny}x{nz}")


def sand_column_128k(seed=2):
    """3D DP sand column 40x80x40 at r = 0.01 (128,000 particles), dt = 1/6000, 20 iterations."""
    return sand_column(40, 80, 40, 0.01, [2.0, 1.0, 2.0], seed)


def sand_column_1m(seed=1):
    """1M DP sand: 80x160x80 at r = 0.005 (1,024,000 particles), dt = 1/6000, 20 iterations."""
    return sand_column(80, 160, 80, 0.005, [1.5, 1.0, 1.5], seed)


def dam_breach(scale=1.0, seed=3, dt=1.0 / 6000.0, iterations=10):
    """Dam breach at r = 0.005: PBF water (rho0 1000, CFM eps 100, s_corr k 0.1) next to
    a DP sand wall (phi 35), the blocks and the domain scaled by `scale` from the
    1M layout (60x100x100 water, 50x80x100 sand in [0,1.5]x[0,0.75]x[0,0.5])."""
    r = 0.005
    s = _solver(r, dt, iterations, pbfEps=100.0, sCorrK=0.1)
    s["domainMin"], s["domainMax"] = [0.0, 0.0, 0.0], [1.5 * scale, 0.75 * scale, 0.5 * scale]
    out, rng = _new(), Rng(seed, 0)
    nw = [int(round(60 * scale)), int(round(100 * scale)), int(round(100 * scale))]
    ns = [int(round(50 * scale)), int(round(80 * scale)), int(round(100 * scale))]
    x0w, x0s = 0.05 * scale, 0.05 * scale + 60 * scale * r
    sample_box(out, [x0w, 0.0, 0.0], [x0w + nw[0] * r, nw[1] * r, nw[2] * r], r, 1, rng, 0.1)
    sample_box(out, [x0s, 0.0, 0.0], [x0s + ns[0] * r, ns[1] * r, ns[2] * r], r, 0, rng, 0.1)
    mats = [DP_SAND_STIFF(35.0), WATER(1000.0)]
    nW, nS = nw[0] * nw[1] * nw[2], ns[0] * ns[1] * ns[2]
    return _finish(s, out, mats, seed, True, info=f"dam breach {nW} water + {nS} sand")


def dam_breach_1m(seed=3, dt=1.0 / 6000.0, iterations=10):
    """Dam breach: 60x100x100 PBF water next to a 50x80x100 DP sand wall (phi 35), r = 0.005."""
    return dam_breach(1.0, seed, dt, iterations)


def vm_poke(n, r, seed, dt=1.0 / 6000.0, iterations=20, sigY=5e3):
    """A von Mises block of n^3 particles at spacing r (sigmaY 5e3, E 2e5, nu 0.3) a
    little above the floor of the unit box, and a sphere of radius 10 r moving down at
    1 m/s from just above it."""
    s = _solver(r, dt, iterations)
    s["domainMin"], s["domainMax"] = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
    out, rng = _new(), Rng(seed, 0)
    e = n * r
    mn = [0.5 - e / 2, 0.05, 0.5 - e / 2]
    sample_box(out, mn, [mn[0] + e, mn[1] + e, mn[2] + e], r, 0, rng, 0.1)
    s["collider"] = dict(pos=[0.5, mn[1] + e + 14 * r, 0.5], R=10 * r, vel=[0.0, -1.0, 0.0])
    return _finish(s, out, [PLASTICINE(sigY, E=2e5, nu=0.3, density=2000.0)], seed, False,
                   info=f"von Mises block {n}^3 + poke")


def vm_poke_1m(seed=5, dt=1.0 / 6000.0, iterations=20):
    """von Mises block 100^3 at r = 0.005 (sigY 5e3, E 2e5, nu 0.3) with a 0.05 m sphere at 1 m/s."""
    return vm_poke(100, 0.005, seed, dt, iterations)


def ensemble_64(seed=4, count=64, dt=1.0 / 6000.0, iterations=20, r=0.01):
    """Batched ensemble: `count` independent scenes with particle counts log-uniform in
    2k-60k (seeded), each a jittered block dropped on its own floor, its material drawn
    uniformly from Drucker-Prager (phi 20-40 deg), NACC (xi 0-10, beta 0.3), von Mises
    (sigmaY 1e3-1e4 Pa), Herschel-Bulkley (h 0.3-1.0, sigmaY 50-500 Pa, eta 10-100) and
    the snow clamp (theta_c 2.5e-2, theta_s 7.5e-3, xi 10); E 2e5 Pa, nu 0.3, rho 2000.
    Each block stands on its floor, so it compresses and flows from the first substeps.
    Returns the list of scene dicts."""
    pick = Rng(seed, 1000)
    out_scenes = []
    for q in range(count):
        u = pick.take(4)
        n = int(round(math.exp(math.log(2000) + u[0] * (math.log(60000) - math.log(2000)))))
        kind = int(u[1] * 5)
        v = u[2]
        if kind == 0:
            m, st = SAND(20 + 20 * v, 0.0, E=2e5, nu=0.3, density=2000.0), 0.0
        elif kind == 1:
            alpha0 = -0.01
            m = dict(kind=MAT["NACC"], E=2e5, nu=0.3, density=2000.0, p0=alpha0, p1=0.3, p2=10.0 * v, p3=1.85,
                     color=[0.85, 0.9, 1.0])
            st = alpha0
        elif kind == 2:
            m, st = PLASTICINE(1e3 * 10 ** v, E=2e5, nu=0.3, density=2000.0), 0.0
        elif kind == 3:
            m = GOO(0.3 + 0.7 * v, 50 + 450 * u[3], 10 + 90 * v, E=2e5, nu=0.3, density=2000.0)
            st = 0.0
        else:
            m, st = SNOW_CLASSIC(2.5e-2, 7.5e-3, 10.0, E=2e5, nu=0.3, density=2000.0), 1.0
        # a block of about n particles, twice as tall as wide
        nx = max(4, int(round((n / 2.0) ** (1.0 / 3.0))))
        ny = max(4, int(round(n / (nx * nx))))
        ext = [nx * r, ny * r, nx * r]
        dom = [ext[0] * 3.0, ext[1] * 1.3 + 4 * r, ext[2] * 3.0]
        s = _solver(r, dt, iterations)
        s["domainMin"], s["domainMax"] = [0.0, 0.0, 0.0], dom
        mn = [ext[0], 0.0, ext[2]]
        mx = [mn[0] + ext[0], mn[1] + ext[1], mn[2] + ext[2]]
        o, rng = _new(), Rng(seed, q)
        sample_box(o, mn, mx, r, 0, rng, 0.1, state0=st)
        out_scenes.append(_finish(s, o, [m], seed * 1000 + q, False,
                                  info=f"ensemble scene {q}: {m['kind']} x {nx}x{ny}x{nx}"))
    return out_scenes
