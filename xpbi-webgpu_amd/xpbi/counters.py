"""Hardware counters of this build's kernels from rocprofv3 --pmc passes.

    python -m xpbi.counters [--regex R] [--copy-bytes B] [--particles N] [--last L]
                            [--alg NAME=BYTES ...] -- PROGRAM [ARGS...]
    python -m xpbi.counters --trace RESULTS_DB [--last-substeps K]

Runs PROGRAM once per counter set under `rocprofv3 --pmc` (every set stays within
the per-block limits of one pass: FETCH_SIZE alone, WRITE_SIZE alone, then SQ and
TA/TCP counters with GRBM_GUI_ACTIVE), keeps the kernels whose names match R, and
prints per kernel the mean over its last L dispatches of:

* us: dispatch duration under the profiler;
* traffic: FETCH_SIZE + WRITE_SIZE in bytes, each scaled by what the counter reads
  for the library's copy kernel (k_copy) of B bytes read and B written, which
  PROGRAM must run (xpbi_bench --copy-mib, run_scene --copy-mib);
* VALU busy: SQ_INSTS_VALU x 4 cycles (a wave64 VALU op on a 16-lane SIMD) over
  1024 SIMDs x the kernel's cycles (GRBM_GUI_ACTIVE / 8 XCDs);
* TA busy: TA_TA_BUSY / 256 CUs over the kernel's cycles;
* TA cycles per load: TA_TA_BUSY / TA_FLAT_READ_WAVEFRONTS;
* lines per load: TCP_TOTAL_CACHE_ACCESSES / SQ_INSTS_VMEM_RD (the cache lines
  one wave-wide load instruction touches);
* L2 requests per load: TCP_TCC_READ_REQ / SQ_INSTS_VMEM_RD.

The second form summarises a `rocprofv3 --kernel-trace` database instead: per
kernel the calls, mean, min and max duration and share of the summed kernel
time, over the whole run and over the last K substeps (counted back from the
end by k_integrate, the last kernel of a substep).  Launches shorter than a fifth
of their kernel's median get a "(short)" line of their own: the solve launches
queued past the end of a residual-driven loop, which return at entry.

Passes of a few substeps profile cleanly; passes over hundreds of substeps have
ended in a segmentation fault inside the profiler's dispatch interception, so a
late state of a scene is profiled by loading a checkpoint (xpbi.run_scene --load).
"""
import argparse
import glob
import os
import re
import shutil
import sqlite3
import subprocess
import sys
import tempfile

SETS = {
    "fetch": ["FETCH_SIZE"],
    "write": ["WRITE_SIZE"],
    "sq": ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "GRBM_GUI_ACTIVE"],
    "ta": ["TA_TA_BUSY", "TA_FLAT_READ_WAVEFRONTS", "TCP_TOTAL_CACHE_ACCESSES", "TCP_TCC_READ_REQ",
           "GRBM_GUI_ACTIVE"],
}
XCDS, CUS, SIMDS = 8, 256, 1024
PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def kernel_name(mangled):
    """xpbi::k_solveB<...>(...) -> k_solveB; other names up to their argument list."""
    m = re.match(r"_ZN4xpbi(\d+)", mangled)
    if m:
        return mangled[m.end():m.end() + int(m.group(1))]
    return mangled.split("(")[0]


def one_pass(counters, program, regex, out_dir, last=10, timeout=180):
    """One rocprofv3 --pmc pass of `program` (argv list).  Returns {kernel: {counter:
    mean over its last `last` dispatches of the chip-wide sum, "us": mean duration,
    "dispatches": count}}, or raises RuntimeError with the tail of the pass's output."""
    os.makedirs(out_dir, exist_ok=True)
    cmd = ["rocprofv3", "--pmc"] + counters + ["--kernel-include-regex", regex, "-d", out_dir, "-o", "run", "--"] + \
        list(program)
    env = dict(os.environ, TMPDIR=out_dir)
    env["PYTHONPATH"] = PKG + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    r = subprocess.run(cmd, cwd=out_dir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout, env=env)
    dbs = glob.glob(os.path.join(out_dir, "**", "*.db"), recursive=True)
    if r.returncode != 0 or not dbs:
        tail = r.stdout.decode(errors="replace")[-2000:]
        raise RuntimeError(f"rocprofv3 pass {counters} ended with {r.returncode}:\n{tail}")
    con = sqlite3.connect(dbs[0])
    cur = con.cursor()
    tabs = {row[0] for row in cur.execute("select name from sqlite_master where type='table'")}

    def table(prefix):
        return [t for t in tabs if t.startswith(prefix)][0]
    pmc, info, disp, sym = (table("rocpd_pmc_event"), table("rocpd_info_pmc"), table("rocpd_kernel_dispatch"),
                            table("rocpd_info_kernel_symbol"))
    symcols = [c[1] for c in cur.execute(f"pragma table_info({sym})")]
    infocols = [c[1] for c in cur.execute(f"pragma table_info({info})")]
    kcol = "kernel_name" if "kernel_name" in symcols else symcols[1]
    ncol = "name" if "name" in infocols else infocols[1]
    rows = cur.execute(f"select d.id, s.{kcol}, i.{ncol}, p.value, d.start, d.end from {pmc} p join {info} i "
                       f"on p.pmc_id = i.id join {disp} d on p.event_id = d.event_id join {sym} s "
                       f"on d.kernel_id = s.id").fetchall()
    con.close()
    per, dur, name = {}, {}, {}
    for did, kn, cn, v, t0, t1 in rows:
        per.setdefault(did, {}).setdefault(cn, 0.0)
        per[did][cn] += v
        dur[did], name[did] = (t1 - t0) / 1e3, kernel_name(kn)
    out = {}
    for k in set(name.values()):
        ids = sorted(i for i in per if name[i] == k)
        tail = ids[-last:]
        agg = {c: sum(per[i].get(c, 0.0) for i in tail) / len(tail) for c in per[tail[0]]}
        agg["us"] = sum(dur[i] for i in tail) / len(tail)
        agg["dispatches"] = len(ids)
        out[k] = agg
    return out


def collect(program, regex, copy_bytes=None, sets=("fetch", "write", "sq", "ta"), last=10, keep=None):
    """All passes of `sets` over `program`; returns (derived per kernel, raw per set).
    With copy_bytes, the regex is widened to k_copy and traffic is calibrated on it."""
    root = keep or tempfile.mkdtemp(prefix="xpbi_pmc_")
    rx = regex + ("|k_copy" if copy_bytes else "")
    raw = {}
    try:
        for s in sets:
            raw[s] = one_pass(SETS[s], program, rx, os.path.join(root, s), last)
    finally:
        if keep is None:
            shutil.rmtree(root, ignore_errors=True)
    return derive(raw, copy_bytes), raw


def derive(raw, copy_bytes=None):
    fs = ws = None
    if copy_bytes and "k_copy" in raw.get("fetch", {}) and "k_copy" in raw.get("write", {}):
        fs = copy_bytes / (raw["fetch"]["k_copy"]["FETCH_SIZE"] * 1024.0)
        ws = copy_bytes / (raw["write"]["k_copy"]["WRITE_SIZE"] * 1024.0)
    kernels = sorted({k for r in raw.values() for k in r} - {"k_copy"})
    out = dict(fetch_scale=fs and round(fs, 4), write_scale=ws and round(ws, 4), kernels={})
    for k in kernels:
        e = {}
        for s in ("sq", "ta", "fetch", "write"):
            if k in raw.get(s, {}):
                e.setdefault("us", round(raw[s][k]["us"], 2))
        if fs and k in raw["fetch"] and k in raw["write"]:
            e["traffic_bytes"] = round(raw["fetch"][k]["FETCH_SIZE"] * 1024.0 * fs +
                                       raw["write"][k]["WRITE_SIZE"] * 1024.0 * ws)
            e["fetch_bytes"] = round(raw["fetch"][k]["FETCH_SIZE"] * 1024.0 * fs)
        q, t = raw.get("sq", {}).get(k), raw.get("ta", {}).get(k)
        if q:
            cyc = q["GRBM_GUI_ACTIVE"] / XCDS
            e["waves"] = round(q["SQ_WAVES"])
            e["valu_per_wave"] = round(q["SQ_INSTS_VALU"] / max(q["SQ_WAVES"], 1.0), 1)
            e["loads_per_wave"] = round(q["SQ_INSTS_VMEM_RD"] / max(q["SQ_WAVES"], 1.0), 1)
            e["valu_busy"] = round(q["SQ_INSTS_VALU"] * 4.0 / SIMDS / cyc, 3)
        if t:
            cyc = t["GRBM_GUI_ACTIVE"] / XCDS
            e["ta_busy"] = round(t["TA_TA_BUSY"] / CUS / cyc, 3)
            e["ta_cycles_per_load"] = round(t["TA_TA_BUSY"] / max(t["TA_FLAT_READ_WAVEFRONTS"], 1.0), 1)
        if q and t:
            vm = max(q["SQ_INSTS_VMEM_RD"], 1.0)
            e["lines_per_load"] = round(t["TCP_TOTAL_CACHE_ACCESSES"] / vm, 1)
            e["l2_requests_per_load"] = round(t["TCP_TCC_READ_REQ"] / vm, 2)
        out["kernels"][k] = e
    return out


def table_text(derived, particles=None, alg=None):
    """The derived figures as a text table (traffic per particle and as a multiple of
    the algorithmic bytes when `particles` and `alg` {kernel: bytes per particle} are given)."""
    cols = ["us", "traffic B/p", "x alg", "VALU busy", "TA busy", "TA cyc/load", "lines/load", "L2 req/load",
            "VALU/wave", "loads/wave"]
    lines = ["kernel".ljust(14) + "".join(c.rjust(13) for c in cols)]
    for k, e in derived["kernels"].items():
        tb = e.get("traffic_bytes")
        bpp = tb / particles if tb is not None and particles else None
        a = (alg or {}).get(k[2:] if k.startswith("k_") else k)
        vals = [e.get("us"), bpp, bpp / a if bpp is not None and a else None, e.get("valu_busy"), e.get("ta_busy"),
                e.get("ta_cycles_per_load"), e.get("lines_per_load"), e.get("l2_requests_per_load"),
                e.get("valu_per_wave"), e.get("loads_per_wave")]
        lines.append(k.ljust(14) + "".join(("-" if v is None else f"{v:.3g}" if abs(v) < 10 else f"{v:.1f}")
                                           .rjust(13) for v in vals))
    return "\n".join(lines)


def trace_table(db, last_substeps=0):
    """Text summary of a rocprofv3 kernel-trace database (see the module docstring)."""
    con = sqlite3.connect(db)
    cur = con.cursor()
    tabs = {r[0] for r in cur.execute("select name from sqlite_master where type='table'")}
    disp = [x for x in tabs if x.startswith("rocpd_kernel_dispatch")][0]
    sym = [x for x in tabs if x.startswith("rocpd_info_kernel_symbol")][0]
    cols = [r[1] for r in cur.execute(f"pragma table_info({sym})")]
    ncol = "kernel_name" if "kernel_name" in cols else cols[1]
    rows = cur.execute(f"select d.start, d.end, s.{ncol} from {disp} d join {sym} s on d.kernel_id = s.id "
                       f"order by d.start").fetchall()
    con.close()
    names = [kernel_name(r[2]) for r in rows]
    ends = [i for i, n in enumerate(names) if n == "k_integrate"]
    start = ends[-last_substeps - 1] + 1 if last_substeps and len(ends) > last_substeps else 0

    def agg(lo):
        a = {}
        for (t0, t1, _), n in zip(rows[lo:], names[lo:]):
            a.setdefault(n, []).append((t1 - t0) / 1e3)
        return a
    every, window = agg(0), agg(start)
    # launches far shorter than their kernel's median (a solve launch after the
    # residual test has stopped the loop returns at entry) get a line of their own
    for k in list(window):
        v = sorted(window[k])
        med = v[len(v) // 2]
        short = [x for x in window[k] if x < 0.2 * med]
        if short:
            window[k] = [x for x in window[k] if x >= 0.2 * med]
            window[k + " (short)"] = short
            e = every[k]
            every[k + " (short)"] = [x for x in e if x < 0.2 * med]
            every[k] = [x for x in e if x >= 0.2 * med]
    tot = sum(sum(v) for v in window.values())
    wall = (rows[-1][1] - rows[start][0]) / 1e3 if rows else 0.0
    out = [f"dispatches: {len(rows)}; window: {len(rows) - start} dispatches from {start}"
           + (f" (last {last_substeps} substeps)" if start else "") + f", {wall:.1f} us wall, {tot:.1f} us in kernels",
           f"{'kernel':30s} {'calls':>7s} {'avg_us(all)':>11s} | {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} "
           f"{'max_us':>9s} {'share':>6s}"]
    for k, v in sorted(window.items(), key=lambda kv: -sum(kv[1])):
        e = every[k]
        out.append(f"{k:30s} {len(e):7d} {sum(e) / len(e):11.2f} | {len(v):6d} {sum(v) / len(v):9.2f} "
                   f"{min(v):9.2f} {max(v):9.2f} {sum(v) / tot:6.3f}")
    return "\n".join(out)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "--trace":
        ap = argparse.ArgumentParser()
        ap.add_argument("--trace", required=True)
        ap.add_argument("--last-substeps", type=int, default=0)
        a = ap.parse_args(argv)
        print(trace_table(a.trace, a.last_substeps))
        return 0
    if "--" not in argv:
        print(__doc__, file=sys.stderr)
        return 2
    cut = argv.index("--")
    ap = argparse.ArgumentParser()
    ap.add_argument("--regex", default="k_")
    ap.add_argument("--copy-bytes", type=int, default=None)
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--last", type=int, default=10)
    ap.add_argument("--alg", action="append", default=[], help="NAME=BYTES algorithmic bytes per particle")
    ap.add_argument("--sets", default="fetch,write,sq,ta")
    a = ap.parse_args(argv[:cut])
    program = argv[cut + 1:]
    alg = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in a.alg}
    d, _ = collect(program, a.regex, a.copy_bytes, tuple(a.sets.split(",")), a.last)
    print(f"# fetch scale {d['fetch_scale']}, write scale {d['write_scale']} (bytes per counted byte, from k_copy)")
    print(table_text(d, a.particles, alg))
    return 0


if __name__ == "__main__":
    sys.exit(main())
