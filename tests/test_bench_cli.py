"""bench.py's command line without a GPU: more than one GPU is refused with
a message (the library has no multi-GPU path), and the help lists the
sections it can skip."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_help_lists_the_sections():
    r = _bench("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--no-closure", "--no-batch", "--no-cpu", "--closure-only"):
        assert flag in r.stdout
