"""bench.py's command line on the CPU: one GPU only (a request for more exits
with a JSON error before any device is touched), and the help text lists the
kernel switches the measurements in the design notes use."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    for mode in ("vmult", "solve"):
        r = _run("--gpus", "2", "--mode", mode)
        assert r.returncode == 2
        line = json.loads(r.stdout.strip().splitlines()[-1])
        assert "one GPU" in line["error"]


def test_help_lists_the_kernel_switches():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--dg-kernel", "--patch-cells", "--cells-per-wave", "--precision", "--degree", "--kind"):
        assert flag in r.stdout
