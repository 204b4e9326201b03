"""The JavaScript host (sim.js + frame driver over the N-API addon) drives the same
library as the Python host: frames of a reference scene give identical state."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import xpbi
from xpbi import scenes

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
@pytest.mark.parametrize("name", ["Sand: two blocks collide", "Dam breach: water + sand"])
def test_js_frames_match_python(tmp_path, name):
    frames = 12
    out = tmp_path / "js.bin"
    print("js frames", name, flush=True)
    subprocess.run(["node", os.path.join(REPO, "xpbi-webgpu_amd", "js", "run.js"), "frames", name, "1", str(frames),
                    str(out)], check=True, timeout=120)
    scene = scenes.SCENES[name](seed=1)
    sim = xpbi.Simulation(scene, device=0)
    sim.writeParams(xpbi.params_from_scene(scene))
    # frame semantics of main.js:383-415: Params once per frame, substeps in chunks
    for _ in range(frames):
        sim.step(scene["substeps"], scene["iterations"], scene["hasWater"])
    raw = out.read_bytes()
    N = sim.N
    js_pos = np.frombuffer(raw[: 16 * N], np.float32).reshape(N, 4)
    js_id = np.frombuffer(raw[16 * N:], np.uint32)
    assert np.array_equal(js_id, sim.read("ID"))
    assert np.array_equal(js_pos, sim.read("POS"))
    sim.close()
    # the Python frame driver, its chunks replayed from a HIP graph, lands on the same state
    from xpbi.frame import FrameDriver
    sim = xpbi.Simulation(scene, device=0)
    fd = FrameDriver(sim, scene, graph=True)
    for _ in range(frames):
        fd.frame()
    assert np.array_equal(js_id, sim.read("ID"))
    assert np.array_equal(js_pos, sim.read("POS"))
    sim.close()
