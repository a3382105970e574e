"""bench.py's command line on the CPU: one GPU only, and its help text."""
import os
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _bench(*args, env=None):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120,
                          env={**os.environ, **(env or {})})


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    r = _bench("--gpus", "1", env={"WORLD_SIZE": "4"})
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_help_lists_the_forms():
    r = _bench("--help")
    assert r.returncode == 0
    for form in ("precision", "batching", "serving", "spmm"):
        assert form in r.stdout
