"""bench.py's command-line contract that holds without a device: more than
one GPU is refused with a message and exit status 2 (the single-GPU path is
the one implemented and measured), from the flag or from a launcher's
WORLD_SIZE."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, env=e)


def test_more_than_one_gpu_is_refused():
    for args, env in ((["--gpus", "2"], None), (["--gpus", "8", "--steps", "1"], None),
                      ([], {"WORLD_SIZE": "2"})):
        p = _run(args, env)
        assert p.returncode == 2, (args, env, p.stderr)
        msg = json.loads(p.stdout.strip().splitlines()[-1])
        assert "one GPU" in msg["error"]


def test_help_lists_the_workloads():
    p = _run(["--help"])
    assert p.returncode == 0
    for w in ("hd", "uneven", "shadowfit", "orbits", "rk4", "offline", "shader", "dense", "fit"):
        assert w in p.stdout
