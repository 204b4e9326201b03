"""Planted mistakes: each must fail one of the oracle's pins (no device).

Every entry below is a plausible slip in the restatement of the reference's
solver (a dropped term, a wrong sign, index or constant, a shared helper changed
in both of its uses).  The test copies the oracle's sources, plants one mistake,
builds the copy with the oracle's own g++ line, and runs the oracle's pin and
invariant tests (test_oracle_pins.py, test_oracle_substep.py) against it through
the oracle fixture's XPBI_ORACLE_LIB override: each run must fail.  The untouched
copy must pass, so a failure is the planted mistake's doing.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import oracle_py

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.dirname(oracle_py.LIB)

# (name, text in xpbi_oracle.hpp, replacement, occurrences replaced)
MUTATIONS = [
    ("pinv tolerance 0.2 -> 0.02", "const T tol = T(0.2);", "const T tol = T(0.02);", 1),
    ("volume V_b dropped from A", "mscale(outer(g, r), vin[4 * (size_t)b + 3])", "mscale(outer(g, r), T(1))", 1),
    ("alpha~ inverted", "return (T(1) / P.V0) / (P.dt * P.dt);", "return P.V0 * (P.dt * P.dt);", 1),
    ("(1 + gamma) dropped (solveA and features)", "(T(1) + g) * c.denom", "c.denom", 2),
    ("sign of gamma dt dC.v flipped (solveA and features)", "- g * P.dt * c.cdot", "+ g * P.dt * c.cdot", 3),
    ("selfDv dropped", "dv = V(sp[3], sp[7], sp[11]);", "dv = splat<T>(0);", 1),
    ("constraint gather gradient flipped", "Vi * mul(loadM(S.Sp, b), kernelGradW(P, dxp))",
     "Vi * mul(loadM(S.Sp, b), kernelGradW(P, -dxp))", 1),
    ("contact weight m_b / (m_a + m_b) -> 1", "T w = mb / (m.mass + mb);", "T w = T(1);", 1),
    ("wall friction dropped", "T drop = wmin(fric * dvn, vtl);", "T drop = T(0) * wmin(fric * dvn, vtl);", 1),
    ("wall speed from the predicted position", "T vnNew = (bound - xn[k]) / P.dt;",
     "T vnNew = (bound - xq[k]) / P.dt;", 1),
    ("sphere speed cap 0.6 -> 6", "(P.colliderR - dist) / P.dt, T(0.6))", "(P.colliderR - dist) / P.dt, T(6.0))", 1),
    ("PBF CFM epsilon dropped", " + P.pbfEps);", ");", 1),
    ("PBF neighbour multiplier dropped", "(li + lb + sCorr)", "(li + sCorr)", 1),
    ("s_corr exponent 4 -> 2", "wmax(w0, T(1e-12)), T(4))", "wmax(w0, T(1e-12)), T(2))", 1),
    ("XSPH velocity difference flipped", "(xyz(vin, b) - vi) * kernelW", "(vi - xyz(vin, b)) * kernelW", 1),
]


def _build(tmp, tag, mutation):
    d = os.path.join(tmp, tag)
    os.makedirs(d, exist_ok=True)
    for f in ("xpbi_oracle.cpp", "xpbi_oracle.hpp"):
        shutil.copy(os.path.join(ORACLE, f), d)
    if mutation is not None:
        _, old, new, count = mutation
        p = os.path.join(d, "xpbi_oracle.hpp")
        src = open(p).read()
        assert src.count(old) == count, f"{mutation[0]}: '{old}' occurs {src.count(old)} times"
        open(p, "w").write(src.replace(old, new))
    lib = os.path.join(d, "liboracle.so")
    cmd = [a if a != oracle_py.LIB else lib for a in oracle_py.GXX_CMD]
    cmd = [os.path.join(d, "xpbi_oracle.cpp") if a.endswith("xpbi_oracle.cpp") else a for a in cmd]
    subprocess.run(cmd, check=True)
    return lib


def _pins(lib):
    env = dict(os.environ, XPBI_ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(HERE, "test_oracle_pins.py"), os.path.join(HERE, "test_oracle_substep.py")],
                       env=env, capture_output=True, text=True, cwd=os.path.dirname(HERE))
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    return r.returncode, failed


def test_every_planted_mistake_fails_a_pin(tmp_path):
    jobs = [("clean", None)] + [(f"m{k}", m) for k, m in enumerate(MUTATIONS)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = list(ex.map(lambda j: _build(str(tmp_path), *j), jobs))
        results = list(ex.map(_pins, libs))
    rc, failed = results[0]
    assert rc == 0, f"the unmutated copy fails: {failed}"
    survivors = []
    for (tag, m), (rc, failed) in zip(jobs[1:], results[1:]):
        print(f"{m[0]}: {'caught by ' + failed[0].split(' ')[1] if rc else 'NOT CAUGHT'}", flush=True)
        if rc == 0:
            survivors.append(m[0])
    assert not survivors, f"planted mistakes no pin catches: {survivors}"
