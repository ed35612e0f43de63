"""bench.py's command line without a GPU: it parses, --help works, and more than one GPU is refused with a
message (this build runs on one GPU)."""
import json
import os
import subprocess
import sys

from conftest import ROOT


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_help():
    r = run("--help")
    assert r.returncode == 0 and "--steps" in r.stdout


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
