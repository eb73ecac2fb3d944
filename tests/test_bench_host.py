"""CPU: bench.py's counter arithmetic on synthetic profiles (no device).

The roofline entry and the calibration record are built from rocprofv3 counter
readings; these tests feed known readings and check the derived figures (busy
fractions, scale factors, probe rates, the bound chosen, JSON serialisation)."""
import importlib.util
import json
import os

import pytest

_spec = importlib.util.spec_from_file_location(
    "fbct_bench", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)

G = 8 * 1.0e6  # GRBM_GUI_ACTIVE: 8 XCDs x 1e6 GPU cycles


def _probe_profile():
    """Probe readings: the FMA probe issues a VALU instruction every quad-cycle on every
    SIMD (2^30 wave instructions), the LDS probe moves 2^37 bytes, the copy 512 MiB."""
    vec4 = bench.PROBE_COPY_VEC4
    return {
        "probe.valu": {"kernels": ["probe_valu_kernel"], "ns": {"a": 2.0e6},
                       "counters": {"GRBM_GUI_ACTIVE": G, "SQ_ACTIVE_INST_VALU": G * 32.0, "SQ_ACTIVE_INST_VALU2": 0.0,
                                    "SQ_LDS_IDX_ACTIVE": 0.0, "SQ_INSTS_VALU_FLOPS_FP32": 2.0 * 2 ** 30}},
        "probe.lds": {"kernels": ["probe_lds_kernel"], "ns": {"a": 1.0e6},
                      "counters": {"GRBM_GUI_ACTIVE": G, "SQ_ACTIVE_INST_VALU": 0.0, "SQ_LDS_IDX_ACTIVE": G * 16.0,
                                   "SQ_INSTS_LDS_LOAD_BANDWIDTH": G * 16.0 * 4.0}},
        "probe.copy": {"kernels": ["probe_copy_kernel"], "ns": {"a": 1.0e5},
                       "counters": {"GRBM_GUI_ACTIVE": G, "SQ_ACTIVE_INST_VALU": 0.0, "SQ_LDS_IDX_ACTIVE": 0.0,
                                    "FETCH_SIZE": 16.0 * vec4 / 1024.0 / 2.0, "WRITE_SIZE": 16.0 * vec4 / 1024.0}},
    }


def test_calibration_factors_and_probe_rates():
    cal = bench._calibration(_probe_profile())
    assert cal["fetch_factor"] == pytest.approx(2.0) and cal["write_factor"] == pytest.approx(1.0)
    assert cal["probe_valu_busy"] == pytest.approx(1.0)
    assert cal["probe_lds_busy"] == pytest.approx(0.5)
    assert cal["lds_bytes_per_active_cycle"] == pytest.approx(256.0)
    assert cal["flops_per_count"] == pytest.approx(64.0)
    peaks = cal["measured_peaks"]
    assert peaks["fp32_fma_tflops"] == pytest.approx(2.0 ** 37 / 2.0e6 / 1e3, rel=1e-3)
    assert peaks["lds_read_tbs"] == pytest.approx(2.0 ** 37 / 1.0e6 / 1e3, rel=1e-3)
    assert peaks["hbm_copy_gbs"] == pytest.approx(32.0 * bench.PROBE_COPY_VEC4 / 1.0e5, rel=1e-3)
    out = bench.calibration_summary(cal)
    json.dumps(out)  # serialisable, nested probe rates kept
    assert out["measured_peaks"] == peaks


def test_roofline_bound_and_busy():
    prof = _probe_profile()
    # one dispatch of 2 ms: VALU issued 80% of quad-cycles, LDS active 40% of cycles,
    # 1 GB fetched (after scaling), 2 wave lanes of 64 active on average
    prof["op"] = {"kernels": ["void fbct::k<4, unsigned short>"], "ns": {"a": 2.0e6},
                  "counters": {"GRBM_GUI_ACTIVE": G, "SQ_ACTIVE_INST_VALU": G * 32.0 * 0.8,
                               "SQ_ACTIVE_INST_VALU2": 0.0, "SQ_LDS_IDX_ACTIVE": G * 32.0 * 0.4,
                               "SQ_LDS_BANK_CONFLICT": G * 32.0 * 0.1, "FETCH_SIZE": 1.0e9 / 1024.0 / 2.0,
                               "WRITE_SIZE": 0.0, "TCC_HIT": 3.0, "TCC_MISS": 1.0,
                               "SQ_THREAD_CYCLES_VALU": 32.0 * 100.0, "SQ_ACTIVE_INST_VALU_b": 100.0}}
    cal = bench._calibration(prof)
    r = bench.roofline("k", 2.0, "op", prof, cal, flops=1.0e12, lds_bytes=2.0e12, model="m")
    assert r["kernel"] == "k<4, bf16>"
    assert r["bound"] == "valu_issue" and r["unit"] == "TFLOP/s"
    assert r["achieved"] == pytest.approx(500.0)
    assert r["busy"]["valu_issue"] == pytest.approx(0.8) and r["busy"]["lds_pipe"] == pytest.approx(0.4)
    assert r["busy"]["hbm"] == pytest.approx(1.0e9 / 2.0e-3 / 1e9 / bench.HBM_PEAK_GBS)
    assert r["l2_hit_rate"] == pytest.approx(0.75)
    assert r["lds_bank_conflict_per_active"] == pytest.approx(0.25)
    assert r["valu_lane_utilisation"] == pytest.approx(0.5)
    json.dumps(r)


def test_roofline_tie_prefers_the_priced_resource():
    prof = _probe_profile()
    prof["f"] = {"kernels": ["ramp_gemm_kernel"], "ns": {"a": 1.0e6},
                 "counters": {"GRBM_GUI_ACTIVE": G, "SQ_ACTIVE_INST_VALU": 0.0, "SQ_LDS_IDX_ACTIVE": G * 32.0 * 0.25,
                              "SQ_INSTS_MFMA": 1.0, "SQ_VALU_MFMA_BUSY_CYCLES": G / 8.0 * 1024.0 * 0.25,
                              "FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0}}
    cal = bench._calibration(prof)
    r = bench.roofline("f", 1.0, "f", prof, cal, flops=1.0e12)
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s"


def test_kernel_names_lose_their_argument_lists():
    disp = {0: {"kernel": "void fbct::(anonymous namespace)::probe_marker_kernel(float*)", "grid": 4 * 64,
                "ns": 1.0, "counters": {}},
            1: {"kernel": "void fbct::(anonymous namespace)::fp_kernel<4, 8, unsigned short, unsigned short>"
                          "(fbct::(anonymous namespace)::FpArgs, int const*, int)", "grid": 256, "ns": 3.0,
                "counters": {"SQ_WAVES": 4.0}}}
    ops = bench._group_by_marker(disp, ["a", "b", "c", "op"])
    assert ops["op"]["kernels"] == ["void fbct::fp_kernel<4, 8, unsigned short, unsigned short>"]
    assert ops["op"]["ns"] == 3.0 and ops["op"]["counters"] == {"SQ_WAVES": 4.0}


def test_roofline_without_profile_says_why():
    r = bench.roofline("k", 1.0, "op", {"error": "disabled (--no-counters)"}, None, flops=1.0)
    assert r["bound"] is None and "disabled" in r["note"]
