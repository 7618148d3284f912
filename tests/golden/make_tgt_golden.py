"""Writes tests/golden/tgt_ray_samples.txt: the output of oracle/_ref/tgt_probe,
the harness that evaluates CPURaycaster's ray setup, sample recurrence,
texture-to-voxel transform, clamp and rounding with the reference's own
header-only tgt, and the projection * view matrices of tgt::Camera for a set
of cameras and window sizes (built by `make -C oracle ref` from the reference's
tgt sources where the reference tree is present).  Run from the repository root: python tests/golden/make_tgt_golden.py"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
PROBE = os.path.join(ROOT, "oracle", "_ref", "tgt_probe")
OUT = os.path.join(ROOT, "tests", "golden", "tgt_ray_samples.txt")

if __name__ == "__main__":
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    text = subprocess.run([PROBE], check=True, capture_output=True, text=True).stdout
    with open(OUT, "w") as f:
        f.write(text)
    print(f"wrote {OUT} ({text.count(chr(10))} lines)")
