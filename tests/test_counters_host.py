"""Host logic of xpbi.counters (no profiler, no GPU): kernel names from mangled
symbols, the copy-kernel calibration and the derived busy fractions on made-up
counter sums whose answers are known by construction."""
from xpbi import counters


def test_kernel_name():
    assert counters.kernel_name("_ZN4xpbi8k_solveBILi4EEEvPKfS2_") == "k_solveB"
    assert counters.kernel_name("__amd_rocclr_fillBufferAligned(void*)") == "__amd_rocclr_fillBufferAligned"


def _raw():
    cyc = 2.0e5  # kernel cycles per XCD
    copy_b = 256 << 20
    return {
        # the copy reads copy_b and writes copy_b; the fetch counter reads half of it
        "fetch": {"k_copy": {"FETCH_SIZE": copy_b / 2 / 1024.0, "us": 70.0},
                  "k_solveB": {"FETCH_SIZE": 100e6 / 2 / 1024.0, "us": 131.0}},
        "write": {"k_copy": {"WRITE_SIZE": copy_b / 1024.0, "us": 70.0},
                  "k_solveB": {"WRITE_SIZE": 50e6 / 1024.0, "us": 131.0}},
        "sq": {"k_solveB": {"SQ_WAVES": 16000.0, "SQ_INSTS_VALU": 0.5 * 1024 * cyc / 4.0,
                            "SQ_INSTS_VMEM_RD": 1.0e6, "GRBM_GUI_ACTIVE": 8 * cyc, "us": 131.0}},
        "ta": {"k_solveB": {"TA_TA_BUSY": 0.75 * 256 * cyc, "TA_FLAT_READ_WAVEFRONTS": 2.0e6,
                            "TCP_TOTAL_CACHE_ACCESSES": 1.6e7, "TCP_TCC_READ_REQ": 3.0e6,
                            "GRBM_GUI_ACTIVE": 8 * cyc, "us": 131.0}},
    }


def test_derive_known_answers():
    d = counters.derive(_raw(), copy_bytes=256 << 20)
    assert abs(d["fetch_scale"] - 2.0) < 1e-9 and abs(d["write_scale"] - 1.0) < 1e-9
    e = d["kernels"]["k_solveB"]
    assert abs(e["traffic_bytes"] - 150e6) < 2
    assert e["valu_busy"] == 0.5 and e["ta_busy"] == 0.75
    assert e["lines_per_load"] == 16.0 and e["l2_requests_per_load"] == 3.0
    assert e["ta_cycles_per_load"] == round(0.75 * 256 * 2.0e5 / 2.0e6, 1)
    assert "k_copy" not in d["kernels"]
    t = counters.table_text(d, particles=1_000_000, alg={"solveB": 96})
    row = [l for l in t.splitlines() if l.startswith("k_solveB")][0].split()
    assert row[2] == "150.0" and row[3] == "1.56"


def test_derive_without_calibration():
    raw = _raw()
    del raw["fetch"]["k_copy"]
    d = counters.derive(raw, copy_bytes=256 << 20)
    assert d["fetch_scale"] is None and "traffic_bytes" not in d["kernels"]["k_solveB"]
    assert d["kernels"]["k_solveB"]["ta_busy"] == 0.75
