"""bench.py's command-line contract on the CPU: one GPU only (a request for
more exits with a message before touching a device), and the options the
profiles' commands use exist."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_profile_options_exist():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--workload", "--shape", "--shape-mixed",
                "--ldos-only", "--ldos-shape", "--converge-only", "--copy-probe-mib",
                "--no-cpu", "--no-converge", "--no-ldos"):
        assert opt in r.stdout, opt
