"""The frame driver's host logic (no device): what one frame of the reference's loop
(web/js/main.js:366-415) writes and submits, for the Python driver (xpbi.frame) and
the JavaScript one (js/frame.js), against the loop's own formulas:

* dt = (1/60) timeScale / substeps (main.js:388), unless the scene fixes dt;
* the collider velocity is the finite difference of its position over the frame's
  wall time (at least 1 ms), clamped to +-1.5 m/s per axis (main.js:390-398);
* the wall time is clamped to 1/30 s (main.js:369-371);
* the substeps go out in chunks of 6 (main.js:404-412);
* simTime advances by (1/60) timeScale per frame (main.js:413)."""
import json
import os
import shutil
import subprocess

import pytest

from xpbi import scenes
from xpbi.frame import CHUNK, FrameDriver

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(REPO, "xpbi-webgpu_amd", "js")


class _Recorder:
    """Stands in for a Simulation: records what the driver writes and submits."""

    def __init__(self):
        self.params, self.calls = [], []

    def writeParams(self, p):
        self.params.append(dict(p))

    def graph_step(self, n, I, water):
        self.calls.append(("graph", n, I, water))

    def step(self, n, I, water):
        self.calls.append(("eager", n, I, water))


def _cfg(substeps=14, timeScale=0.42):
    cfg = dict(scenes.SCENES["Plasticine: drop & poke"](seed=1), substeps=substeps, timeScale=timeScale)
    cfg.pop("dt", None)
    return cfg


def test_frame_dt_chunks_and_time():
    cfg = _cfg(substeps=14)
    rec = _Recorder()
    fd = FrameDriver(rec, cfg, graph=True)
    fd.frame(1 / 60)
    fd.frame(1 / 60)
    dt = (1 / 60) * 0.42 / 14
    assert rec.params[0]["dt"] == pytest.approx(dt, rel=1e-15)
    assert [c[1] for c in rec.calls] == [6, 6, 2, 6, 6, 2]
    assert all(c[0] == "graph" and c[2] == cfg["iterations"] and c[3] == cfg["hasWater"] for c in rec.calls)
    assert rec.params[0]["time"] == 0.0 and rec.params[1]["time"] == pytest.approx((1 / 60) * 0.42)
    assert fd.sim_time == pytest.approx(2 * (1 / 60) * 0.42)
    assert fd.frame_dt() == pytest.approx(dt * 14)
    rec2 = _Recorder()
    FrameDriver(rec2, _cfg(substeps=12), graph=False).frame()
    assert rec2.calls == [("eager", 6, cfg["iterations"], cfg["hasWater"])] * 2


def test_collider_velocity_finite_difference_and_clamps():
    cfg = _cfg()
    rec = _Recorder()
    fd = FrameDriver(rec, cfg)
    fd.move_collider([0.01, -0.002, 0.0])
    fd.frame(0.02)  # 0.5 m/s, -0.1 m/s, 0
    fd.move_collider([0.2, 0.0, -0.2])
    fd.frame(0.02)  # 10 m/s -> clamped to 1.5, -10 -> -1.5
    fd.move_collider([0.001, 0.0, 0.0])
    fd.frame(0.5)  # wall time clamped to 1/30: 0.03 m/s
    fd.frame(1e-6)  # no motion; the 1 ms floor keeps the quotient finite
    v = [p["colliderVel"] for p in rec.params]
    assert v[0] == pytest.approx([0.5, -0.1, 0.0])
    assert v[1] == [1.5, 0.0, -1.5]
    assert v[2] == pytest.approx([0.03, 0.0, 0.0])
    assert v[3] == [0.0, 0.0, 0.0]
    assert rec.params[1]["colliderPos"] == pytest.approx([0.5 + 0.21, 0.3 - 0.002, 0.5 - 0.2])


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_js_frame_params_match():
    """js/frame.js frameParams on the same inputs gives the same Params (dt, clamped
    collider velocity, time) as the Python driver."""
    cfg = _cfg(substeps=10, timeScale=0.6)
    script = """
const f = require(%s);
const cfg = %s;
const inp = f.defaultInput();
const out = [];
inp.colliderPos = [0.51, 0.29, 0.5];
out.push(f.frameParams(cfg, inp, 0.02, 0.0));
inp.colliderPos = [0.81, 0.29, 0.3];
out.push(f.frameParams(cfg, inp, 0.02, 0.01));
console.log(JSON.stringify(out.map(p => ({dt: p.dt, colliderVel: p.colliderVel, time: p.time}))));
""" % (json.dumps(os.path.join(JS, "frame.js")),
       json.dumps({k: cfg[k] for k in ("substeps", "timeScale", "domainMin", "domainMax", "gravity", "xsphC",
                                      "omega", "contactOmega", "pbfEps", "sCorrK")}))
    js = json.loads(subprocess.check_output(["node", "-e", script], text=True))
    rec = _Recorder()
    fd = FrameDriver(rec, cfg)
    fd.collider_pos = [0.51, 0.29, 0.5]
    fd.frame(0.02)
    fd.collider_pos = [0.81, 0.29, 0.3]
    fd.sim_time = 0.01
    fd.frame(0.02)
    for a, b in zip(js, rec.params):
        assert a["dt"] == pytest.approx(b["dt"], rel=1e-15)
        assert a["colliderVel"] == pytest.approx(b["colliderVel"], rel=1e-12)
        assert a["time"] == pytest.approx(b["time"])
    assert js[1]["colliderVel"] == [1.5, 0, -1.5]
    assert CHUNK == 6
