"""bench.py's command-line contract that needs no GPU: more than one GPU is refused with a
message and a non-zero exit, as is a multi-rank launch."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          env=e, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2"])
    assert r.returncode != 0 and r.stdout == ""
    assert "one GPU" in r.stderr


def test_multi_rank_launch_is_refused():
    r = _run([], env={"WORLD_SIZE": "2"})
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_workloads_listed():
    r = _run(["--help"])
    assert r.returncode == 0
    for w in ("serve", "dam", "batch", "mixed256", "obstacles", "frame", "reference64", "agents", "tune"):
        assert w in r.stdout
    assert "--grids" in r.stdout and "--grids-out" in r.stdout
