"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (no device).

oracle/sanitize_main.cpp, linked with the oracle and built with
-fsanitize=address,undefined -fno-sanitize-recover=all, runs every oracle entry
point on a small water + sand scene whose material table holds one record of
every kind: whole substeps in f32 and f64 with each start (plain, residual loop,
training pairs, learned, given lambda0, staged stops) and the unit entry points
(kernels, SVD including a zero and a rank-one matrix, return maps with and
without hardening, the HB bisection with and without a bracket hint, the grid
hash, the MLP forward).  A sanitizer finding aborts the run; the test requires a
clean exit."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from xpbi import scenes
from xpbi.mlp import init_params
from helpers import mats_record, params_record


def _scene():
    r = 0.02
    s = scenes._solver(r, 1.0 / 3000.0, 10, pbfEps=100.0, sCorrK=1e-4)
    s["domainMin"], s["domainMax"] = [0.0, 0.0, 0.0], [0.5, 0.4, 0.3]
    out, rng = scenes._new(), scenes.Rng(3, 0)
    scenes.sample_box(out, [0.05, 0.0, 0.05], [0.05 + 8 * r, 10 * r, 0.05 + 8 * r], r, 1, rng, 0.2)
    scenes.sample_box(out, [0.25, 0.0, 0.05], [0.25 + 6 * r, 8 * r, 0.05 + 8 * r], r, 0, rng, 0.2)
    mats = [scenes.SAND(30, 0.004), scenes.WATER(1.0), scenes.SNOW_NACC(), scenes.PLASTICINE(),
            scenes.GOO(0.5), scenes.SNOW_CLASSIC(), scenes.ELASTIC()]
    return scenes._finish(s, out, mats, 3, True)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_oracle_entry_points_are_sanitizer_clean(oracle, tmp_path):
    exe = oracle.build_sanitized()
    sc = _scene()
    P, C = params_record(sc)
    M = mats_record(sc)
    st = oracle.initial_state(sc, C)
    for _ in range(3):  # let the blocks move so that constraints and contacts are active
        oracle.substep(st, P, M, 5, True)
    inp = tmp_path / "scene.bin"
    with open(inp, "wb") as f:
        f.write(np.array([st.N, C, 1], np.uint32).tobytes())
        f.write(bytes(P))
        f.write(bytes(M))
        for a in (st.pos, st.velA, st.F, st.aux):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
        f.write(np.ascontiguousarray(st.id, np.uint32).tobytes())
        f.write(np.ascontiguousarray(init_params(1), np.float32).tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                       env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout
    assert "unit entry points done" in r.stdout
    assert r.stdout.count("iterations, position sum") == 12
