"""bench.py command-line contract (no GPU needed: these paths exit before
the library is touched)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_unknown_workload_is_rejected():
    r = _run("--workload", "no_such_workload")
    assert r.returncode != 0
    assert "invalid choice" in r.stderr


def test_int8_modes_price_each_site_on_its_engine():
    """In the int8 modes a call site runs either the int8 emulation (priced
    at 19 or 13 int8 ops per algorithmic FLOP against the 5 POP/s int8 peak)
    or the FP64 kernel (3/4 against 78.6 TFLOP/s): the autotuner's pick when
    it ran, else the built-in sites (P0 contraction, Sigma P.Y1)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    assert b.site_engine("i8x4", "sig_py", None) == (13.0, 5000.0, "int8 emulation")
    assert b.site_engine("i8x5", "p0_contract", None) == (19.0, 5000.0, "int8 emulation")
    assert b.site_engine("i8x4", "sig_vg", None) == (0.75, 78.6, "fp64 kernel")
    tune = {"chosen": {"sig_py": "bk8_w3", "sig_vg": "int8"}}
    assert b.site_engine("i8x5", "sig_py", tune) == (0.75, 78.6, "fp64 kernel")
    assert b.site_engine("i8x5", "sig_vg", tune) == (19.0, 5000.0, "int8 emulation")
    assert b.site_engine("fp64", "sig_py", tune) == (0.75, 78.6, "fp64")
    # five slices picked in a four-slice plan are priced at 19 ops
    assert b.site_engine("i8x4", "sig_vg", {"chosen": {"sig_vg": "int8x5"}}) == (19.0, 5000.0, "int8 emulation")
    r = b.mode_roofline("i8x4", {"sig_py": (2000.0, 2.0e14, 10), "p0_gv": (10.0, 1.0e12, 10)})
    assert r["engine"] == "int8 emulation" and abs(r["frac_mfma"] - 100.0 * 13 / 5000.0) < 1e-9


def test_mode_roofline_counts_matrix_core_flops():
    """Per-mode roofline of the dominant site: matrix-core FLOPs executed per
    algorithmic FLOP are 3/4 (FP64 Gauss), 1 (FP32) and 3 (split-bf16 x3)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    sites = {"sig_py": (2000.0, 1.0e14, 10), "p0_gv": (10.0, 1.0e12, 10)}
    for mode, factor, peak in (("fp64", 0.75, 78.6), ("fp32", 1.0, 157.3), ("bf16x3", 3.0, 2500.0)):
        r = b.mode_roofline(mode, sites)
        assert r["kernel"] == "gemm site sig_py"
        assert abs(r["achieved_algorithmic"] - 50.0) < 1e-9
        assert abs(r["achieved_mfma"] - 50.0 * factor) < 1e-9
        assert r["peak"] == peak and abs(r["frac_mfma"] - round(50.0 * factor / peak, 4)) < 1e-12
