"""bench.py's command line without a GPU: more than one GPU, or a
multi-rank launch, is refused with exit code 2 and a message, before any
device is touched."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, BENCH] + args, capture_output=True, text=True,
                          timeout=120, env=e, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2"])
    assert r.returncode == 2
    assert "only one GPU is measured" in r.stderr
    assert r.stdout == ""


def test_multi_rank_launch_is_refused():
    r = _run([], env={"WORLD_SIZE": "2", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
