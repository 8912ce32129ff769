"""bench.py's command-line contract, without a GPU: one GPU only (more exits with
a message before anything touches a device), and the documented switches."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_switches_are_documented():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-single", "--no-batch", "--no-precision",
                 "--no-gemm", "--gemm-study", "--sweep-study", "--pmc-pass"):
        assert flag in r.stdout, flag
