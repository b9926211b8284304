"""bench.py's command line on the CPU: more than one GPU is refused with a message (there is no multi-GPU path), and
the defaults give a run of a few steps with the mixed COCG policy the measurements are quoted for."""
import os
import subprocess
import sys

import bench as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for argv in (["--gpus", "2"], ["--gpus", "8", "--steps", "1"]):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, capture_output=True, text=True,
                           timeout=120, cwd=ROOT)
        assert r.returncode == 2, r.stderr
        assert "one GPU only" in r.stderr
        assert r.stdout == ""   # no JSON line


def test_defaults(monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = B.parse()
    assert (a.gpus, a.steps, a.warmup) == (1, 4, 1)
    assert a.method == "cocg" and a.mode == "bench"
    # 0 and -1 select the policy's defaults: a budget of 1200 and the progress check at 2/5 of it
    assert a.fp32_maxit == 0 and a.fp32_check == -1
    assert B.fp32_policy(a) == (1200, 480)
    a.fp32_maxit, a.fp32_check = 2000, 0
    assert B.fp32_policy(a) == (2000, 0)
    a.method, a.fp32_maxit = "tfqmr", 0
    assert B.fp32_policy(a) == (3000, 0)
