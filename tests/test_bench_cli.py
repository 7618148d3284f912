"""bench.py's command line on the CPU: more than one GPU is refused before
anything touches a device, and the training-threshold option parses into the
stepper's keyword arguments.  CPU only."""
import argparse
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_learned_tau_option():
    import bench
    assert bench.tau_args(argparse.Namespace(learned_tau="")) == {}
    assert bench.tau_args(argparse.Namespace(learned_tau="6e-3,1.6e-2")) == {"tau_skip": 6e-3, "tau_step": 1.6e-2}
    saved = sys.argv
    try:
        sys.argv = ["bench.py", "--learned-tau", "8e-3,2.4e-2", "--steps", "2"]
        args = bench.parse()
    finally:
        sys.argv = saved
    assert args.learned_tau == "8e-3,2.4e-2" and args.steps == 2 and args.gpus == 1
