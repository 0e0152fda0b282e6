"""The measurement script's command-line contract, on the CPU: one GPU only
(any other --gpus value exits with status 2 and says so), and every workload
named in --help."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr and r.stdout == ""


def test_help_names_the_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("leads", "uneven", "small", "warmstart", "gw", "scba", "corrconv", "ephoton", "vendor-zgemm",
              "copy"):
        assert w in r.stdout, w
    for flag in ("--steps", "--warmup", "--gpus", "--no-mixed", "--no-cpu"):
        assert flag in r.stdout, flag
