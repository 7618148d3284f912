"""The oracle's pins catch the mistakes they are meant to catch.

Each case copies oracle/dvr_oracle.c to a temporary directory, plants one
plausible mistake by a text substitution (a transposed pre-integration
table, swapped classify indices, the float 0.95 early-termination literal,
gamma applied as its reciprocal), builds the copy with the oracle's own
flags, and checks that the pin of that function fails on it while the same
pin passes on the real oracle.  CPU only; nothing here touches the GPU or
the product library.
"""
import os
import shutil
import subprocess
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import test_oracle as T  # noqa: E402

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

MUTANTS = {
    "preint_table_transposed": (
        "float* dst = table + 4 * ((size_t)sb * n + sf);",
        "float* dst = table + 4 * ((size_t)sf * n + sb);",
        T.pin_preint_orientation),
    "classify_indices_swapped": (
        "const float* c = P->preint_table + 4 * ((size_t)i2 * N + i1);",
        "const float* c = P->preint_table + 4 * ((size_t)i1 * N + i2);",
        T.pin_preint_classify_order),
    "preint_ert_float_literal": (
        "((double)a < 0.95)",
        "(a < 0.95f)",
        T.pin_preint_early_termination),
    "gamma_reciprocal": (
        "return (gamma != 1.f) ? powf(v, gamma) : v;",
        "return (gamma != 1.f) ? powf(v, 1.f / gamma) : v;",
        T.pin_tf_gamma_ramp),
}


def _build_mutant(tmp, old, new):
    src = open(os.path.join(ORACLE_DIR, "dvr_oracle.c")).read()
    assert src.count(old) >= 1, old
    open(os.path.join(tmp, "dvr_oracle.c"), "w").write(src.replace(old, new, 1))
    for f in ("dvr_oracle.h", "mlp_oracle.c"):
        shutil.copy(os.path.join(ORACLE_DIR, f), tmp)
    out = os.path.join(tmp, "libdvr_oracle_mutant.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-shared", "-o", out, os.path.join(tmp, "dvr_oracle.c"), os.path.join(tmp, "mlp_oracle.c"),
                    "-lm"], check=True)
    return out


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(oracle, tmp_path, name):
    old, new, pin = MUTANTS[name]
    pin(oracle)                       # passes on the oracle itself
    lib = _build_mutant(str(tmp_path), old, new)
    with oracle.variant(lib):
        with pytest.raises(AssertionError):
            pin(oracle)
