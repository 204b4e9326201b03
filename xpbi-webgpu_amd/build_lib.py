"""Build recipe of libxpbi.so (gfx950) and the Node N-API addon, in-tree.

hipcc flags:
  --offload-arch=gfx950                  MI355X only
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE f32 division and sqrt (the grid
                                         hash floor((p-o)/cs) must match bit for bit)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed-FP32 instructions are emitted
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
SOURCES = ["xpbi_sort.hip", "xpbi_solve.hip", "xpbi_mlp.hip", "xpbi_api.hip", "xpbi2d.hip"]
LIB = os.path.join(PKG, "libxpbi.so")
ADDON_SRC = os.path.join(PKG, "js", "addon", "xpbi_napi.cc")
ADDON = os.path.join(PKG, "js", "xpbi_napi.node")
BENCH_SRC = os.path.join(CSRC, "xpbi_bench_main.cpp")
BENCH = os.path.join(PKG, "xpbi_bench")

HIPCC_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
               "-fhip-fp32-correctly-rounded-divide-sqrt",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-value", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return out


def build_lib():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I" + os.path.join(REPO, "include")] + \
        [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    """N-API addon (node_api.h from the system Node headers), linked against libxpbi.so."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")) or not os.path.exists(ADDON_SRC):
        return None
    tmp = ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + inc, "-I" + os.path.join(REPO, "include"),
           ADDON_SRC, "-L" + PKG, "-lxpbi", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    _run(cmd)
    os.replace(tmp, ADDON)
    return ADDON


def build_bench():
    """The native substep driver (C ABI only), linked against libxpbi.so."""
    tmp = BENCH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include"), BENCH_SRC, "-L" + PKG, "-lxpbi",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    _run(cmd)
    os.replace(tmp, BENCH)
    return BENCH


def build():
    build_lib()
    build_addon()
    build_bench()


if __name__ == "__main__":
    build()
    print(LIB)
