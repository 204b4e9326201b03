"""Herschel-Bulkley return map on the GPU against the CPU oracle: the bisection branch.

The reference's HB return map (web/js/shaders/common.js:244-280) solves its flow
rule analytically when the exponent h is 1 and otherwise bisects 18 times over
[sY, s_pre] (common.js:260-271).  The reference Goo scene uses h = 1, so these
tests run it with h = 0.5 and h = 2.0, where every yielding particle bisects:

* the root positions of the first iteration's return map (features row 13, the
  training target of the network's second output) and one plain substep match the
  oracle;
* the learned bracket hint (a network whose outputs are constants (omega, t), so
  that the warm start is plain Jacobi's first step and only the hint differs):
  t near the median root (narrowed 13-step bisection for the particles whose root
  lies within 1/64 of t, the sign check's fallback for the others), t = 0 and
  t = 1 (the bracket clipped at an end), t far from every root (fallback
  everywhere), NaN and t outside [0, 1] (hint ignored).  Each is compared with the
  oracle running the same network, and F with the oracle's plain 18-step run.
"""
import numpy as np
import pytest

import xpbi
from xpbi import scenes
from helpers import F3, mats_record
from test_gpu_parity import check_field, compare, make_sim, one_substep_parity, perturbed, snapshot

pytestmark = pytest.mark.gpu


def _goo(hexp, frames=40):
    scene = scenes.scene_goo(seed=1)
    scene["materials"] = [scenes.GOO(hexp)]
    sim = make_sim(scene)
    sim.step(frames * scene["substeps"], scene["iterations"], False)
    return scene, sim


def _const_net(omega, t):
    """All weights zero: the network outputs (omega, t) for every particle."""
    p = np.zeros(xpbi._abi.MLP_PARAMS, np.float32)
    p[-2], p[-1] = omega, t
    return p


def _inf_net(w3):
    """Finite weights whose output overflows: hidden units 0 and 1 of the second layer
    reach +inf (3e38 from the first layer's bias, times 2 in f32), and output k is
    w3[k][0] inf + w3[k][1] inf.  (1, 1) gives rho = +inf (clamped to 2), (1, -1) NaN."""
    import mlp_oracle as mo
    d = mo.split(np.zeros(xpbi._abi.MLP_PARAMS, np.float64))
    d["b1"][0] = 3e38
    d["W2"][0, 0] = d["W2"][1, 0] = 2.0
    d["W3"][0, :2] = w3[0]
    d["W3"][1, :2] = w3[1]
    return mo.join(d).astype(np.float32)


def _roots(sim, scene, oracle):
    """HB root positions of the next substep's first return map: GPU, f32 and f64 oracle."""
    I = scene["iterations"]
    st32 = snapshot(sim, oracle)
    st64 = st32.copy(np.float64)
    _, t, _ = sim.collect_warmstart_pairs(I, False)
    mats = mats_record(scene)
    out = []
    for st in (st32, st64):
        f = np.zeros((sim.N, 16), st.dtype)
        tt = np.zeros((sim.N, 2), st.dtype)
        oracle.substep(st, sim.params, mats, I, False, collect=True, feat=f, target=tt)
        out.append(tt[:, 1].astype(np.float64))
    return t[:, 1].astype(np.float64), out[0], out[1]


@pytest.mark.parametrize("hexp", [0.5, 2.0])
def test_hb_bisection_matches_oracle(oracle, hexp):
    scene, sim = _goo(hexp)
    g, o32, o64 = _roots(sim, scene, oracle)
    hb = np.isfinite(g) & np.isfinite(o64) & np.isfinite(o32)
    print(f"goo h={hexp}: N={sim.N}, bisecting particles gpu {np.isfinite(g).sum()} "
          f"oracle {np.isfinite(o64).sum()}", flush=True)
    assert hb.sum() > 0.05 * sim.N, "too few particles take the bisection branch"
    # the same particles yield (a few may sit on the s_pre = sY cut)
    assert (np.isfinite(g) != np.isfinite(o64)).sum() <= max(3, 2e-3 * sim.N)
    assert ((g[hb] >= 0) & (g[hb] <= 1)).all()
    e, eo = np.abs(g[hb] - o64[hb]), np.abs(o32[hb] - o64[hb])
    print(f"  root position: gpu-f64 q50 {np.median(e):.2e} q99 {np.quantile(e, 0.99):.2e} | "
          f"f32-f64 q99 {np.quantile(eo, 0.99):.2e}", flush=True)
    # 18 halvings of [0, 1] leave a 2^-18 bracket; f32 inputs move the root by their
    # own conditioning
    assert np.median(e) <= 4 * 2.0 ** -18 + np.median(eo)
    assert np.quantile(e, 0.99) <= max(1e-3, 3 * np.quantile(eo, 0.99))
    one_substep_parity(sim, scene, oracle, f"goo h={hexp}", water=False)
    sim.close()


@pytest.mark.parametrize("hexp", [0.5, 2.0])
def test_hb_bracket_hint_matches_oracle(oracle, hexp):
    scene, sim = _goo(hexp)
    I = scene["iterations"]
    omega = float(sim.params.omega)
    g, _, o64 = _roots(sim, scene, oracle)  # also advances one plain substep
    live = np.isfinite(o64)
    med = float(np.median(o64[live]))
    t_near = round(med * 64) / 64
    near = np.abs(o64[live] - t_near) < 1 / 64
    print(f"goo h={hexp}: median root {med:.3f}; hint {t_near:.4f} brackets the root of "
          f"{near.sum()} of {live.sum()} particles", flush=True)
    assert 0 < near.sum() < live.sum(), "the hint should narrow some brackets and not others"
    far = 0.0 if med > 0.5 else 1.0  # an end of the bracket, far from most roots
    base = snapshot(sim, oracle)
    mats = mats_record(scene)
    # the oracle's plain 18-step substep: the hinted runs reach the same roots to 2^-18
    plain32, plain64 = base.copy(), base.copy(np.float64)
    oracle.substep(plain32, sim.params, mats, I, False)
    oracle.substep(plain64, sim.params, mats, I, False)
    cases = [(f"hint {t}", _const_net(omega, t), True) for t in
             (t_near, far, 1.0 - far, 0.37 if med > 0.7 or med < 0.3 else 0.97, 1.5, -0.5)]
    # NaN hint (rho = +inf, clamped to 2): the full bracket, as with no hint at all
    cases.append(("NaN hint, rho inf", _inf_net([(1, 1), (1, -1)]), False))
    # NaN rho and hint: lambda0 = 0 (plain Jacobi from zero) and the full bracket
    cases.append(("NaN rho and hint", _inf_net([(1, -1), (1, -1)]), False))
    states = {}
    for label, net, vs_plain in cases:
        for k in ("POS", "VEL", "F", "AUX", "ID"):
            sim.write(k, getattr(base, {"POS": "pos", "VEL": "velA", "F": "F", "AUX": "aux", "ID": "id"}[k]))
        sim.set_warmstart(net)
        sim.substep(I, False)
        sim.set_warmstart(None)
        variants = [base.copy()] + [perturbed(base, k) for k in (1, 2)]
        st64 = base.copy(np.float64)
        for st in variants + [st64]:
            oracle.substep(st, sim.params, mats, I, False, warm=net)
        assert np.array_equal(sim.read("PERM"), variants[0].sortedIdx)
        label = f"goo h={hexp}, {label}"
        compare(sim, variants, st64, label)
        if vs_plain:
            check_field("F vs the plain 18-step oracle", F3(sim.read("F")), [F3(plain32.F)], F3(plain64.F), 1.0,
                        1e-4, label)
        states[label] = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX")}
    # a NaN hint is no hint: the run equals rho = 2 with the hint output -1, bit for bit
    for k in ("POS", "VEL", "F", "AUX", "ID"):
        sim.write(k, getattr(base, {"POS": "pos", "VEL": "velA", "F": "F", "AUX": "aux", "ID": "id"}[k]))
    sim.set_warmstart(_const_net(2.0, -1.0))
    sim.substep(I, False)
    sim.set_warmstart(None)
    for k, v in states[f"goo h={hexp}, NaN hint, rho inf"].items():
        assert np.array_equal(sim.read(k), v), k
    sim.close()
