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


def test_graph_replay_needs_resident_integrals():
    """A captured iteration replays fixed device addresses, so --graph is
    refused for integrals streamed from host memory (the workload's default)."""
    r = _run("--graph")
    assert r.returncode != 0
    assert "--integrals device" in r.stderr


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
