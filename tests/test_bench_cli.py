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
