"""bench.py's command line (no GPU): one GPU only, and the flags the profiles use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_profile_flags_are_accepted():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--sections", "--skip-cpu", "--lattice-tiles", "--target-z", "--source-points",
                 "--source-jitter", "--tile", "--no-near", "--no-series", "--no-rows",
                 "--hologram-steps", "--mc-passes", "--calibrate-copy", "--repeats"):
        assert flag in r.stdout, flag
