"""Build libhipdvr.so (HIP kernels for gfx950 + the C-ABI) in-tree.

    python -m voreen_amd.build          # or: python voreen_amd/build.py

Every source under voreen_amd/csrc is compiled by hipcc into one object per
file, then linked into voreen_amd/libhipdvr.so.  Flags:
  --offload-arch=gfx950             MI355X (CDNA4) only
  -ffp-contract=off                 no implicit FMA: the march restates the
                                    reference's float expression order
  -target-feature -packed-fp32-ops  no packed FP32 (v_pk_*_f32) instructions
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhipdvr.so")
BUILD_DIR = os.path.join(HERE, "_build")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD_DIR, src + ".o")
    cmd = [hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


MODULE = os.path.join(HERE, "hipraycaster")
MODULE_OUT = os.path.join(MODULE, "libhipraycaster_shim.so")


def build_module(verbose=True):
    """HIPRaycaster processor + module compiled against the Voreen shim, with
    the test driver (host code only, g++), linked to libhipdvr.so."""
    srcs = [os.path.join(MODULE, p) for p in ("shim/shim.cpp", "hipraycastermodule.cpp",
                                               "processors/hipraycaster.cpp", "test/driver.cpp")]
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-I", os.path.join(MODULE, "shim"),
           "-I", os.path.join(HERE, "..", "include"), "-I", MODULE, *srcs,
           "-L", HERE, "-lhipdvr", "-Wl,-rpath,$ORIGIN/..", "-o", MODULE_OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"module build failed:\n{r.stderr}")
    os.replace(MODULE_OUT + ".tmp", MODULE_OUT)
    if verbose:
        print(f"built {MODULE_OUT}")
    return MODULE_OUT


def _stale(srcs):
    """Sources whose object is missing or older than the source or than any
    header (headers are shared, so a changed header rebuilds everything)."""
    hdr_dirs = (CSRC, os.path.join(HERE, "..", "include"))
    newest_hdr = max(os.path.getmtime(os.path.join(d, f)) for d in hdr_dirs for f in os.listdir(d)
                     if f.endswith(".h"))
    out = []
    for s in srcs:
        obj = os.path.join(BUILD_DIR, s + ".o")
        if (not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(os.path.join(CSRC, s))
                or os.path.getmtime(obj) < newest_hdr):
            out.append(s)
    return out


def build(verbose=True, incremental=False):
    """Compile every source (incremental=True: only those whose object is
    stale) and link libhipdvr.so and the HIPRaycaster module."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    todo = _stale(srcs) if incremental else srcs
    if todo:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(BUILD_DIR, s + ".o") for s in srcs]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "--offload-arch=gfx950", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs,
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    build_module(verbose)
    return OUT


if __name__ == "__main__":
    try:
        build(incremental="--incremental" in sys.argv)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
