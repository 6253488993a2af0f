"""bench.py's command line on the host: the refusals that come before any
device work (more than one GPU, a malformed fixed launch form)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")]
                          + list(args), capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    p = _run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""


def test_malformed_launch_form_is_refused():
    for bad in ("mixed=8,8,4", "fp32=8,0,0,1", "mixed"):
        p = _run("--set-tuning", bad)
        assert p.returncode == 2, bad
        assert "--set-tuning takes" in p.stderr
        assert p.stdout == ""
