"""bench.py's command line on the CPU: one GPU only, and the flags the round-end run relies on."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_before_any_device_work():
    for n in ("2", "8"):
        r = _run("--gpus", n, "--steps", "1", "--warmup", "0")
        assert r.returncode == 2, r.stderr
        assert "exactly one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_contract_flags():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--workload", "--lattice", "--storage", "--streaming",
                 "--wss-every", "--ref-cadence-steps", "--learned", "--no-cpu"):
        assert flag in r.stdout, flag
