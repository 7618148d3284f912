"""hipdvr_synth_turbulence_device_noalloc (include/hipdvr_synth.h), the entry
point that allocates nothing, against hipdvr_synth_turbulence_device: the
same volume bit for bit and the same empty fraction, on two calls in a row,
the second with a smaller volume (the shared scratch keeps the first call's
histogram, which the second must clear)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_turbulence_equals_the_allocating_entry_point(ctx):
    from voreen_amd import hipdvr as H
    cases = [((160, 144, 128), 0x4A11, 0.4), ((67, 41, 29), 0x4A12, 0.25)]
    for dims, seed, frac in cases:
        print(f"[case] no-alloc turbulence {dims} seed {seed:#x} empty {frac}", flush=True)
        n = dims[0] * dims[1] * dims[2]
        a = ctx.alloc(n * 2)
        b = ctx.alloc(n * 2)
        try:
            _, fa = ctx.synth_turbulence(seed, dims, frac, out=a)
            _, fb = ctx.synth_turbulence(seed, dims, frac, out=b, noalloc=True)
            va = a.download((dims[2], dims[1], dims[0]), np.uint16)
            vb = b.download((dims[2], dims[1], dims[0]), np.uint16)
        finally:
            a.free()
            b.free()
        assert np.array_equal(va, vb)
        assert fa == fb and abs(float(np.mean(vb == 0)) - fb) < 1e-12
        assert abs(fb - frac) < 0.02 and vb.max() > 0


def test_noalloc_turbulence_rejects_bad_arguments(ctx):
    from voreen_amd import hipdvr as H
    lib = H.load()
    frac = C.c_double()
    print("[case] no-alloc turbulence argument checks", flush=True)
    assert lib.hipdvr_synth_turbulence_device_noalloc(ctx.h, 1, 0, 4, 4, 0.4, None, C.byref(frac)) == H.ERR_INVALID
    buf = ctx.alloc(64 * 2)
    try:
        assert lib.hipdvr_synth_turbulence_device_noalloc(ctx.h, 1, 4, 4, 4, 1.5, buf.ptr, None) == H.ERR_INVALID
    finally:
        buf.free()
    assert lib.hipdvr_synth_turbulence_device_noalloc(None, 1, 4, 4, 4, 0.4, None, None) == H.ERR_INVALID
