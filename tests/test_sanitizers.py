"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Two harnesses are built with g++ -fsanitize=address,undefined
-fno-sanitize-recover=all, so that any finding aborts with a report:
  - tests/native/host_sanitize_check.cpp with io.cpp (the .vvd parser) and
    tf.cpp (TF tables, pre-integration tables): the fixture .vvd files,
    truncated and corrupted copies of them, random and degenerate TF keys;
  - tests/native/module_sanitize_check.cpp with the HIPRaycaster processor,
    the Voreen shim and the test driver: every property with valid and
    invalid values, volumes of both voxel types, images, the step rules and
    process() (its error path without a GPU).
The device code is not sanitized (no sanitizer runs on the GPU here); the
harnesses call only host code.  The leak checker is off: the HIP runtime
holds allocations until exit.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def _gxx():
    g = shutil.which("g++")
    if not g:
        pytest.skip("g++ not found")
    return g


def test_vvd_parser_and_tf_builders_are_clean(tmp_path):
    exe = str(tmp_path / "host_check")
    csrc = os.path.join(ROOT, "voreen_amd", "csrc")
    # io.cpp's volume loaders reference device-side helpers the harness never
    # calls; they are left unresolved
    subprocess.run([_gxx(), *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "host_sanitize_check.cpp"), os.path.join(csrc, "tf.cpp"),
                    os.path.join(csrc, "io.cpp"), "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
                    "-Wl,--unresolved-symbols=ignore-in-object-files", "-o", exe], check=True, capture_output=True)
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "voreen"), str(scratch)], capture_output=True,
                       text=True, timeout=300, env=ENV)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]


def test_processor_and_shim_are_clean(tmp_path):
    from voreen_amd import hipdvr as H
    H.load()   # libhipdvr.so is built
    exe = str(tmp_path / "module_check")
    mod = os.path.join(ROOT, "voreen_amd", "hipraycaster")
    lib_dir = os.path.join(ROOT, "voreen_amd")
    subprocess.run([_gxx(), *SAN, "-I", os.path.join(mod, "shim"), "-I", os.path.join(ROOT, "include"), "-I", mod,
                    os.path.join(ROOT, "tests", "native", "module_sanitize_check.cpp"),
                    os.path.join(mod, "shim", "shim.cpp"), os.path.join(mod, "hipraycastermodule.cpp"),
                    os.path.join(mod, "processors", "hipraycaster.cpp"), os.path.join(mod, "test", "driver.cpp"),
                    "-L", lib_dir, "-lhipdvr", f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
