"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message before anything touches a device."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""          # no JSON line


def test_bench_help_lists_the_options():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu-baseline", "--thresholds", "--grad-thresholds"):
        assert opt in r.stdout, opt
