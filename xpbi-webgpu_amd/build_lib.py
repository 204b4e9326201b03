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


UBSAN_LIB = os.path.join(PKG, "ubsan", "libxpbi.so")


def build_lib_ubsan():
    """libxpbi.so with UndefinedBehaviorSanitizer on its host code only (GPU code is
    built as usual), the shared sanitizer runtime found by rpath; a finding aborts.
    Load it in place of the library with XPBI_LIB=<this path> (the Python host)."""
    import glob
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    rt = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.ubsan_standalone-x86_64.so")
    if not rt:
        raise RuntimeError("no shared UBSan runtime under /opt/rocm/lib/llvm")
    os.makedirs(os.path.dirname(UBSAN_LIB), exist_ok=True)
    tmp = UBSAN_LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I" + os.path.join(REPO, "include"),
                                   "-Xarch_host", "-fsanitize=undefined",
                                   "-Xarch_host", "-fno-sanitize-recover=undefined", "-shared-libsan",
                                   "-Wl,-rpath," + os.path.dirname(rt[0])] + \
        [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    _run(cmd)
    os.replace(tmp, UBSAN_LIB)
    return UBSAN_LIB


def build():
    build_lib()
    build_addon()
    build_bench()


if __name__ == "__main__":
    if "--ubsan" in sys.argv:
        print(build_lib_ubsan())
    else:
        build()
        print(LIB)
