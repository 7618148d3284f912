"""closed_form_steps (voreen_amd/csrc/step_run.h), the fast-forward through
empty space that the march kernels use, against the reference loop's
sequential t += step (cpuraycaster.cpp:315): 200,000 random runs checked on
the host by tests/native/closed_form_check.cpp, compiled from the same
header the kernels include.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("hipcc not found")


def test_closed_form_steps_equal_sequential_addition(tmp_path):
    exe = str(tmp_path / "closed_form_check")
    subprocess.run([_hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "voreen_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "closed_form_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, cases, jumps = r.stdout.split()
    assert ok == "ok" and int(cases) == 200000 and int(jumps) > 100000
