"""bench.py's command-line contract, on the CPU: more than one GPU is
refused with a message and exit status 2 before anything touches a device,
and the help text names every option the measurement commands use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "4", "--steps", "3")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "4" in r.stderr
    assert r.stdout == ""


def test_options_of_the_measurement_commands():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--gpus", "--config", "--share", "--no-cpu-baseline", "--no-extra",
                "--no-adjoint", "--traffic-probe", "--eager", "--blocks-per-cu", "--no-autotune"):
        assert opt in r.stdout, opt
