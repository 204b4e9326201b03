"""The learned warm start on the GPU against the CPU oracle.

* features and training targets (xpbi_collect_warmstart_pairs) equal the oracle's
  collect mode on the same snapshot, to the f32 conditioning of the constraint;
* a network that predicts rho = omega (plain Jacobi's first step) reproduces plain
  Jacobi bit for bit;
* a learned substep (fixed iteration count, so that both stop together) matches the
  oracle's learned substep to the bar of the plain parity tests, on DP sand and on
  Herschel-Bulkley goo (whose return map takes the network's bracket hint), and at
  full size on the benchmark's 1M DP column and the 1M von Mises poke.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import xpbi
from xpbi import scenes
from xpbi.mlp import init_params
from helpers import mats_record
from test_gpu_parity import compare, make_sim, perturbed, snapshot

pytestmark = pytest.mark.gpu


def _net(seed, rho=1.5, t=0.5):
    """A small random network around lambda0 = rho d0 and HB hint t."""
    rng = np.random.default_rng(seed)
    d = mo.split(init_params(seed).astype(np.float64))
    d["W3"] = d["W3"] * 0.05
    d["b3"] = np.array([rho, t])
    d["b1"] = rng.normal(0, 0.1, 64)
    return mo.join(d).astype(np.float32)


def _settled(name_or_scene, frames=40):
    scene = scenes.SCENES[name_or_scene](seed=1) if isinstance(name_or_scene, str) else name_or_scene
    sim = make_sim(scene)
    sim.step(frames * scene["substeps"], scene["iterations"], scene["hasWater"])
    return scene, sim


def _rows_err(a, b, scale):
    return np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / scale


@pytest.mark.parametrize("name", ["Sand: column collapse", "Goo: Herschel-Bulkley", "Dam breach: water + sand"])
def test_features_and_targets_match_oracle(oracle, name):
    scene, sim = _settled(name)
    I, water = scene["iterations"], scene["hasWater"]
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    x, t, w = sim.collect_warmstart_pairs(I, water)
    mats = mats_record(scene)
    f32, f64 = np.zeros((sim.N, 16), np.float32), np.zeros((sim.N, 16))
    t32, t64 = np.zeros((sim.N, 2), np.float32), np.zeros((sim.N, 2))
    oracle.substep(st32, sim.params, mats, I, water, collect=True, feat=f32, target=t32)
    oracle.substep(st64, sim.params, mats, I, water, collect=True, feat=f64, target=t64)
    active = f32[:, 12] != 0
    print(f"{name}: N={sim.N}, active {active.sum()}", flush=True)
    assert active.sum() > 0.2 * sim.N
    # the same particles carry an active constraint (a handful may sit on the C = 1e-7 cut)
    assert ((x[:, 0] != 0) != active).sum() <= max(3, 1e-3 * sim.N)
    both = active & (x[:, 0] != 0) & (f64[:, 12] != 0)
    scale = np.maximum(np.abs(f64[both, :12]).max(axis=0), 1e-3)
    eg = _rows_err(x[both], f64[both, :12], scale)
    eo = _rows_err(f32[both, :12], f64[both, :12], scale)
    for k in range(12):
        q, qo = np.quantile(eg[:, k], 0.999), np.quantile(eo[:, k], 0.999)
        print(f"  feature {k}: gpu-f64 q999 {q:.2e} max {eg[:, k].max():.2e} | f32-f64 q999 {qo:.2e}", flush=True)
        assert q <= max(1e-4, 3 * qo), k
    # targets: lambda / d0 after the plain solve, and the HB root position
    tv = both & np.isfinite(t64[:, 0])
    er = np.abs(t[tv, 0] - t64[tv, 0]) / np.maximum(np.abs(t64[tv, 0]), 1.0)
    ero = np.abs(t32[tv, 0] - t64[tv, 0]) / np.maximum(np.abs(t64[tv, 0]), 1.0)
    print(f"  target rho: gpu q99 {np.quantile(er, 0.99):.2e} | f32 q99 {np.quantile(ero, 0.99):.2e}", flush=True)
    assert np.quantile(er, 0.99) <= max(1e-3, 3 * np.quantile(ero, 0.99))
    # residual weights (-C - gamma dt dC.v)^2
    ew = np.abs(w[both] - f64[both, 15] ** 2) / np.maximum(f64[both, 15] ** 2, 1e-30)
    ewo = np.abs(f32[both, 15].astype(np.float64) ** 2 - f64[both, 15] ** 2) / np.maximum(f64[both, 15] ** 2, 1e-30)
    print(f"  weight: gpu q999 {np.quantile(ew, 0.999):.2e} | f32 q999 {np.quantile(ewo, 0.999):.2e}", flush=True)
    assert np.quantile(ew, 0.999) <= max(1e-4, 3 * np.quantile(ewo, 0.999))
    assert (w[~active & (x[:, 0] == 0)] == 0).all()
    hb = tv & np.isfinite(t64[:, 1]) & np.isfinite(t[:, 1])
    if hb.any():
        assert np.quantile(np.abs(t[hb, 1] - t64[hb, 1]), 0.99) <= 1e-3
    sim.close()


@pytest.mark.parametrize("scene_name,tol", [(None, 0.0), (None, 1e-3), ("Dam breach: water + sand", 1e-3)])
def test_omega_network_is_plain_jacobi(scene_name, tol):
    """A network predicting rho = omega makes the learned first iteration plain Jacobi's first:
    consecutive substeps agree bit for bit and count the same iterations."""
    if scene_name is None:
        scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=5, iterations=20)
        settle = 230
    else:
        scene = scenes.SCENES[scene_name](seed=1)
        settle = 20 * scene["substeps"]
    water = scene["hasWater"]
    a = make_sim(scene)
    a.step(settle, scene["iterations"], water)
    snap = {k: a.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    b = make_sim(scene)
    for sim in (a, b):
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_tolerance(tol, 4)
    p = np.zeros(5122, np.float32)
    p[5120], p[5121] = 0.5, -1.0  # y = (omega, no HB hint)
    a.set_warmstart(p)
    for _ in range(3):  # consecutive substeps: the previous-lambda feature is exercised too
        a.substep(60, water)
        b.substep(60, water)
        assert a.last_iterations() == b.last_iterations()
        for k in ("POS", "VEL", "F", "AUX", "ID"):
            assert np.array_equal(a.read(k), b.read(k)), k
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["Sand: column collapse", "Goo: Herschel-Bulkley"])
def test_learned_substep_matches_oracle(oracle, name):
    scene, sim = _settled(name)
    net = _net(3)
    I, water = scene["iterations"], scene["hasWater"]
    sim.set_warmstart(net)
    sim.substep(I, water)  # the previous-lambda feature now comes from a learned substep
    st32 = snapshot(sim, oracle)
    # the oracle needs lambda of the last substep: aux.y is in the snapshot (AUX)
    st64 = st32.copy(np.float64)
    variants = [st32] + [perturbed(st32, k) for k in (1, 2)]
    mats = mats_record(scene)
    sim.substep(I, water)
    for st in variants + [st64]:
        oracle.substep(st, sim.params, mats, I, water, warm=net)
    assert np.array_equal(sim.read("PERM"), st32.sortedIdx)
    compare(sim, variants, st64, f"learned {name}")
    sim.close()


def test_learned_substep_headline_column_1m(oracle):
    """The benchmark's learned path at full size: the 1,024,000-particle DP column
    landed (300 substeps), one learned substep so that the previous-lambda feature
    comes from a learned start, then one more (20 iterations, the first from the
    network inside the features pass) against the f32 oracle, its 1-ulp
    perturbations and the f64 oracle, each given the same network."""
    scene = scenes.sand_column_1m()
    sim = make_sim(scene)
    print("headline column: settling 300 substeps", flush=True)
    sim.step(300, 20, False)
    net = _net(5)
    sim.set_warmstart(net)
    sim.substep(20, False)
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    variants = [st32] + [perturbed(st32, k) for k in (1, 2)]
    mats = mats_record(scene)
    sim.substep(20, False)
    for k, st in enumerate(variants + [st64]):
        print(f"oracle run {k + 1} of {len(variants) + 1}", flush=True)
        oracle.substep(st, sim.params, mats, 20, False, warm=net)
    assert np.array_equal(sim.read("PERM"), st32.sortedIdx)
    compare(sim, variants, st64, "learned 1M DP column after 301 substeps")
    sim.close()


def test_learned_substep_vm_poke_1m(oracle):
    """The learned start on the 1M von Mises poke at full size: 400 substeps with the
    sphere moving down at 1 m/s (the previous-lambda feature from learned substeps),
    then one learned substep with the sphere pressing into the block against the f32
    oracle, its 1-ulp perturbations and the f64 oracle, each given the same network."""
    scene = scenes.vm_poke_1m()
    sim = xpbi.Simulation(scene, device=0)
    col, dt, I = scene["collider"], scene["dt"], scene["iterations"]

    def place(k):
        pos = [col["pos"][j] + col["vel"][j] * k * dt for j in range(3)]
        sim.writeParams(xpbi.params_from_scene(scene, collider=dict(pos=pos, R=col["R"], vel=col["vel"])))

    net = _net(7)
    sim.set_warmstart(net)
    steps = 400
    print(f"von Mises poke 1M, learned start: {steps} substeps with the sphere moving", flush=True)
    for k in range(steps):
        place(k)
        sim.substep(I, False)
    place(steps)
    c = np.asarray(col["pos"]) + np.asarray(col["vel"]) * steps * dt
    inside = int((np.linalg.norm(sim.read("POS")[:, :3] - c, axis=1) < col["R"]).sum())
    print(f"particles inside the sphere before the checked substep: {inside}", flush=True)
    assert inside > 100, "the sphere does not press into the block"
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    variants = [st32] + [perturbed(st32, k) for k in (1, 2)]
    mats = mats_record(scene)
    sim.substep(I, False)
    for k, st in enumerate(variants + [st64]):
        print(f"oracle run {k + 1} of {len(variants) + 1}", flush=True)
        oracle.substep(st, sim.params, mats, I, False, warm=net)
    assert np.array_equal(sim.read("PERM"), st32.sortedIdx)
    compare(sim, variants, st64, f"learned 1M von Mises poke after {steps} substeps")
    sim.close()


def test_trained_network_gains():
    """Pairs from 30 substeps of a collapsing DP column, GPU training, then the next 10
    substeps from one snapshot: the learned warm start reaches residual 1e-4 in fewer
    Jacobi iterations (its learned first iteration counted) than plain Jacobi from
    lambda = 0, and ends at the same positions."""
    from xpbi import warmstart
    scene = scenes.sand_column(24, 48, 24, 0.01, [0.8, 0.6, 0.8], seed=11, iterations=20)
    X, T, W = warmstart.collect(scene, 260, 30, keep=0.5)
    W = W / W.mean()
    p, losses = warmstart.train(X, T, W, epochs=30)
    print(f"pairs {len(X)}, loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert losses[-1] < 0.2 * losses[0]
    sim = make_sim(scene)
    sim.step(300, 20, False)
    snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    out = {}
    for mode, net in (("plain", None), ("learned", p)):
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_warmstart(net)
        sim.set_tolerance(1e-4, 4)
        its = []
        for _ in range(10):
            sim.substep(200, False)
            n, r = sim.last_iterations()
            assert r <= 1e-4 and n < 200
            its.append(n)
        pos, ids = sim.read("POS"), sim.read("ID")
        out[mode] = (np.mean(its), pos[np.argsort(ids)])
        print(f"{mode}: iterations {its}", flush=True)
    ratio = out["plain"][0] / out["learned"][0]
    dpos = np.abs(out["plain"][1][:, :3] - out["learned"][1][:, :3]).max()
    print(f"iteration ratio {ratio:.3f}, max position difference {dpos:.2e} m", flush=True)
    assert ratio >= 1.15
    assert dpos <= 1e-4
    sim.close()


@pytest.mark.parametrize("name", ["Sand: column collapse", "Dam breach: water + sand"])
def test_given_lambda0_matches_oracle(oracle, name):
    """xpbi_set_lambda0 against the oracle's given start on the same snapshot: the
    first iteration applies lambda0 of each particle's id (here 1.3 times the
    multipliers of the previous substep), then plain Jacobi."""
    scene, sim = _settled(name)
    I, water = scene["iterations"], scene["hasWater"]
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    lam = np.zeros(sim.N, np.float32)
    lam[sim.read("ID")] = sim.read("AUX")[:, 1]
    assert (lam != 0).mean() > 0.2
    lam0 = (1.3 * lam).astype(np.float32)
    variants = [st32] + [perturbed(st32, k) for k in (1, 2)]
    mats = mats_record(scene)
    sim.set_lambda0(lam0)
    sim.substep(I, water)
    for st in variants + [st64]:
        oracle.substep(st, sim.params, mats, I, water, lam0_by_id=lam0)
    assert np.array_equal(sim.read("PERM"), st32.sortedIdx)
    compare(sim, variants, st64, f"given lambda0 {name}")
    # the given start set lambda of the first iteration: with one iteration only,
    # lambda ends at lambda0 of the id wherever the constraint is active
    snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    sim.substep(1, water)
    ids, lam1 = sim.read("ID"), sim.read("AUX")[:, 1]
    act = lam1 != 0
    assert act.mean() > 0.2 and np.array_equal(lam1[act], lam0[ids[act]])
    sim.set_lambda0(None)
    for k, v in snap.items():
        sim.write(k, v)
    sim.substep(1, water)
    assert not np.array_equal(sim.read("AUX")[:, 1], lam1)
    sim.close()
