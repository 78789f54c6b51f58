"""The benchmark script's command-line contract (CPU): it compiles, and a
request for more than one GPU exits with status 2 and says why, before
touching a GPU."""
import os
import py_compile
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_compiles():
    py_compile.compile(os.path.join(ROOT, "bench.py"), doraise=True)


def test_bench_rejects_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
