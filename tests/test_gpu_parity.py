"""GPU parity: the HIP ray march (through the C-ABI) against the CPU oracle.

Tolerance for the plain path (no shading): max abs RGBA error <= 1e-3 against
the FP32 CPURaycaster restatement at equal step size; the kernels are written
to be bit-exact and the tests also report the exact-pixel fraction.  Phong has
no CPURaycaster counterpart: it is checked against the FP32 restatement of the
GLSL Phong terms (max abs <= 1e-3 on all but ERT-threshold pixels) and the
FP64 oracle (PSNR).
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import fixtures as F  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-3


@pytest.fixture(params=["reference", "bricked", "cell", "slab"])
def kernel(request):
    from voreen_amd import hipdvr as H
    return {"reference": H.KERNEL_REFERENCE, "bricked": H.KERNEL_BRICKED,
            "cell": H.KERNEL_CELL, "slab": H.KERNEL_SLAB}[request.param]


def filters_for(kernel):
    """Filter modes a kernel serves (the cell records hold trilinear
    footprints; the slab kernel reconstructs trilinear samples only)."""
    from voreen_amd import hipdvr as H
    return (1,) if kernel in (H.KERNEL_CELL, H.KERNEL_SLAB) else (0, 1)


def say(msg):
    print(f"[case] {msg}", flush=True)


def render_both(ctx, oracle, vol, tf, domain, en, ex, step, kernel=None, rwm=(1.0, 0.0),
                spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0), filter_=1, shading=0,
                classification=0, preint=None, camera_pos=(0.0, 0.0, 3.5), fp64=False,
                compositing=0, gamma=0.0, iso_value=0.5):
    from voreen_amd import hipdvr as H
    kernel = H.KERNEL_REFERENCE if kernel is None else kernel
    v = ctx.volume(vol, rwm=rwm, spacing=spacing, offset=offset)
    t = ctx.transfunc(tf, domain)
    if preint is not None:
        t.set_preint(preint)
    p = H.default_params(step, filter_=filter_, classification=classification, shading=shading,
                         kernel=kernel, camera_pos=camera_pos, compositing=compositing, gamma=gamma,
                         iso_value=iso_value)
    try:
        gpu, st = ctx.render(v, t, en, ex, p)
    finally:
        v.close()
        t.close()
    light = dict(camera=camera_pos)
    cpu, n = oracle.render(vol, tf, domain, en, ex, step, filter_=filter_,
                           classification=classification, preint=preint, shading=shading,
                           light=light, rwm=rwm, spacing=spacing, offset=offset,
                           compositing=compositing, gamma=gamma, iso_value=iso_value)
    out = dict(gpu=gpu, cpu=cpu, n=n, stats=st)
    if fp64:
        out["fp64"], _ = oracle.render(vol, tf, domain, en, ex, step, filter_=filter_,
                                       classification=classification, preint=preint,
                                       shading=shading, light=light, rwm=rwm, spacing=spacing,
                                       offset=offset, fp64=True)
    return out


def report(name, r, oracle):
    d = np.abs(r["gpu"] - r["cpu"])
    exact = float(np.mean(np.all(r["gpu"] == r["cpu"], axis=2)))
    msg = (f"{name}: max_abs={d.max():.3g} exact_pixels={exact:.4f} "
           f"psnr_vs_fp32={oracle.psnr(r['gpu'], r['cpu']):.1f}")
    if "fp64" in r:
        msg += f" psnr_vs_fp64={oracle.psnr(r['gpu'], r['fp64']):.1f}"
    print(msg, flush=True)
    return d


def c0_inputs(oracle, size=256, view=0, nviews=1):
    from voreen_amd import workloads as W
    vol = W.make_volume("c0")
    keys, tf = W.make_tf("c0")
    cam = W.orbit_camera(view, nviews)
    llf, urb = W.world_box((128, 128, 128))
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, cam.fovy_deg, cam.near_dist, size, size,
                               llf=llf, urb=urb)
    return vol, tf, en, ex, W.step_for((128, 128, 128), 2.0)


def test_loaded_library_was_built_from_these_sources(ctx):
    """On the GPU box the prebuilt libhipdvr.so is loaded as it travelled:
    its embedded digest must equal the digest of the sources beside it."""
    from voreen_amd import build, hipdvr as H
    assert H.source_digest() == build.source_digest(), "libhipdvr.so is stale: run python -m voreen_amd.build"


def test_entry_exit_device_matches_oracle(ctx, oracle):
    from voreen_amd import hipdvr as H, workloads as W
    cases = [(W.orbit_camera(3, 36), 256, 256, False), (W.orbit_camera(7, 36), 97, 61, True)]
    inside = H.Camera()
    inside.eye[:] = [0.1, 0.05, -0.1]
    inside.focus[:] = [0.0, 0.0, -1.0]
    inside.up[:] = [0.0, 1.0, 0.0]
    inside.fovy_deg, inside.near_dist = 50.0, 0.05
    cases.append((inside, 64, 48, False))
    # portrait windows (fovy spans the width) and a fovy clamped to 175 degrees
    cases.append((W.orbit_camera(11, 36), 61, 97, False))
    wide = W.orbit_camera(5, 36, radius=0.4, fovy=179.0, near=0.01)   # the eye inside the volume
    cases.append((wide, 48, 80, True))
    llf, urb = W.world_box((128, 96, 114), (0.015625,) * 3)
    for cam, w, h, q in cases:
        say(f"entry_exit {w}x{h} quantize={q}")
        de, dx = ctx.alloc(w * h * 16), ctx.alloc(w * h * 16)
        ctx.entry_exit_device(cam, llf, urb, w, h, de, dx, quantize=q)
        ge = de.download((h, w, 4), np.float32)
        gx = dx.download((h, w, 4), np.float32)
        de.free()
        dx.free()
        ce, cx = oracle.entry_exit(cam.eye, cam.focus, cam.up, cam.fovy_deg, cam.near_dist, w, h,
                                   llf=llf, urb=urb, quantize=q)
        assert np.array_equal(ge, ce) and np.array_equal(gx, cx)
        assert 0.0 < np.mean(ge[..., 3] > 0) <= 1.0


@pytest.mark.parametrize("size", [(512, 128), (560, 136), (1040, 272), (500, 500), (250, 90)])
def test_tile_orders_write_every_pixel(ctx, oracle, size):
    """Frames of at least one chip region (32 x 8 tiles of 16x16 pixels) are
    marched in per-XCD tile regions, smaller ones in 2x2-tile blocks per XCD.
    Whole, partial and ragged regions (sizes not multiples of 16): the output
    starts as NaN and every pixel, background included, must come out equal
    to the oracle, in the cell and bricked kernels."""
    from voreen_amd import hipdvr as H, workloads as W
    w, h = size
    dims = (64, 64, 64)
    vol = W.make_volume("c0", dims=dims)
    _, tf = W.make_tf("c0")
    cam = W.orbit_camera(5, 36, radius=0.9)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, cam.fovy_deg, cam.near_dist, w, h, llf=llf, urb=urb)
    step = W.step_for(dims, 1.0)
    cpu, n = oracle.render(vol, tf, (0.0, 1.0), en, ex, step)
    assert np.mean(cpu[..., 3] > 0) > 0.5
    v = ctx.volume(vol)
    t = ctx.transfunc(tf)
    de, dx, do = ctx.alloc(w * h * 16), ctx.alloc(w * h * 16), ctx.alloc(w * h * 16)
    try:
        de.upload(en)
        dx.upload(ex)
        for kernel in (H.KERNEL_CELL, H.KERNEL_BRICKED, H.KERNEL_SLAB):
            say(f"tile orders {w}x{h} kernel {kernel}")
            do.upload(np.full((h, w, 4), np.nan, np.float32))
            st = ctx.render_device(v, t, de, dx, w, h, H.default_params(step, kernel=kernel), do, stats=True)
            gpu = do.download((h, w, 4), np.float32)
            assert not np.isnan(gpu).any(), f"{w}x{h} kernel {kernel}: pixels not written"
            assert np.array_equal(gpu, cpu), f"{w}x{h} kernel {kernel}"
            assert st["samples"] == n
    finally:
        for b in (de, dx, do):
            b.free()
        v.close()
        t.close()


def test_host_buffer_render_keeps_staging_across_sizes(ctx, oracle):
    """hipdvr_render (host buffers, the processor's path) keeps its device and
    pinned staging between frames and grows it for a larger frame: frames of
    256^2, 64x40, 300x280 and 256^2 again equal the device-buffer render."""
    from voreen_amd import hipdvr as H, workloads as W
    vol = W.make_volume("c0")
    _, tf = W.make_tf("c0")
    llf, urb = W.world_box((128, 128, 128))
    v = ctx.volume(vol)
    t = ctx.transfunc(tf)
    try:
        for k, (w, h) in enumerate(((256, 256), (64, 40), (300, 280), (256, 256))):
            say(f"host-buffer render {w}x{h}")
            cam = W.orbit_camera(3 + k, 36)
            en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, w, h, llf=llf, urb=urb)
            p = H.default_params(W.step_for((128, 128, 128), 2.0))
            img, st = ctx.render(v, t, en, ex, p)
            de, dx, do = (ctx.alloc(w * h * 16) for _ in range(3))
            try:
                de.upload(en)
                dx.upload(ex)
                st2 = ctx.render_device(v, t, de, dx, w, h, p, do, stats=True)
                dev = do.download((h, w, 4), np.float32)
            finally:
                for b in (de, dx, do):
                    b.free()
            assert np.array_equal(img, dev) and st["samples"] == st2["samples"]
            cpu, n = oracle.render(vol, tf, (0.0, 1.0), en, ex, p.step)
            assert np.array_equal(img, cpu) and st["samples"] == n
    finally:
        v.close()
        t.close()


def test_c0_full_frame(ctx, oracle, kernel):
    say(f"c0 128^3 u8 256^2 step 1/256 linear kernel {kernel}")
    vol, tf, en, ex, step = c0_inputs(oracle)
    r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, step, kernel=kernel, fp64=True)
    d = report("c0", r, oracle)
    assert d.max() <= TOL
    assert r["stats"]["samples"] == r["n"]
    assert r["stats"]["rays"] == int(np.sum(~np.all((en == 0) & (ex == 0), axis=2)))
    assert oracle.psnr(r["gpu"], r["fp64"]) > 50.0


def test_c0_orbit_views_nearest_and_linear(ctx, oracle, kernel):
    for view in (5, 13, 29):
        for filt in filters_for(kernel):
            say(f"c0 view {view}/36 filter {filt} 160^2 kernel {kernel}")
            vol, tf, en, ex, step = c0_inputs(oracle, size=160, view=view, nviews=36)
            r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, step, filter_=filt,
                            kernel=kernel)
            assert report(f"c0 v{view} f{filt}", r, oracle).max() <= TOL
            assert r["stats"]["samples"] == r["n"]


def test_nucleon_regression_workspace_setup(ctx, oracle, kernel):
    # modules/base/test/render/cpuraycaster.vws: nucleon, camera :96-103,
    # canvas 200x200, samplingRate 2, linear filter, TF domain [0, 249]
    # (real-world units, so rwm scale 255), 12-key TF of :228-270 (== nucleon.tfi).
    v = F.load_vvd("nucleon")
    t = F.load_tfi("nucleon")
    tf = oracle.tf_build(t["keys"], t["width"], t["gamma"], t["threshold"], t["alpha_mode"])
    sp = np.array(v["spacing"], np.float32)
    llf = np.array(v["offset"], np.float32) - np.float32(0.5) * sp
    urb = llf + np.array(v["dims"], np.float32) * sp
    en, ex = oracle.entry_exit((1.76227701, 1.01016998, 1.77568305),
                               (-0.023203, 0.019314, 0.012643),
                               (-0.74486113, 0.22378904, 0.62857014), 45.0, 0.1, 200, 200,
                               llf=llf, urb=urb, quantize=True)
    step = float(np.float32(1.0) / (np.float32(41) * np.float32(2.0)))
    for filt in sorted(filters_for(kernel), reverse=True):
        say(f"nucleon 41^3 u8 200^2 filter {filt} kernel {kernel}")
        r = render_both(ctx, oracle, v["data"], tf, t["domain"], en, ex, step, rwm=(255.0, 0.0),
                        filter_=filt, spacing=tuple(sp), offset=v["offset"], kernel=kernel)
        assert report(f"nucleon f{filt}", r, oracle).max() <= TOL
        assert r["stats"]["samples"] == r["n"]
        assert np.mean(r["gpu"][..., 3] > 0) > 0.05


def test_walnut_uint16_noncubic(ctx, oracle, kernel):
    from voreen_amd import workloads as W
    v = F.load_vvd("walnut")
    t = F.load_tfi("walnut")
    tf = oracle.tf_build(t["keys"], t["width"], t["gamma"], t["threshold"], t["alpha_mode"])
    llf, urb = W.world_box(v["dims"], v["spacing"])
    for view in (0, 9):
        say(f"walnut 128x96x114 u16 view {view} 192x144 kernel {kernel}")
        cam = W.orbit_camera(view, 24, radius=2.5)
        en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 192, 144, llf=llf, urb=urb)
        step = W.step_for(v["dims"], 2.0)
        r = render_both(ctx, oracle, v["data"], tf, t["domain"], en, ex, step, rwm=(65535.0, 0.0),
                        kernel=kernel)
        assert report(f"walnut v{view}", r, oracle).max() <= TOL
        assert r["stats"]["samples"] == r["n"]
        assert np.mean(r["gpu"][..., 3] > 0) > 0.02


def test_phong_on_shells(ctx, oracle, kernel):
    from voreen_amd import workloads as W
    dims = (96, 96, 96)
    vol = W.make_volume("c1", dims=dims)
    keys, _ = W.make_tf("c1")
    tf = oracle.tf_build(keys, 4096)
    cam = W.orbit_camera(4, 36)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 128, 128, llf=llf, urb=urb)
    sp = (1.0 / 96,) * 3
    off = tuple(float(llf[i] + 0.5 / 96) for i in range(3))
    say(f"phong shells 96^3 u16 128^2 kernel {kernel}")
    r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, W.step_for(dims, 2.0), shading=1,
                    spacing=sp, offset=off, camera_pos=tuple(cam.eye), kernel=kernel, fp64=True)
    d = report("phong", r, oracle).max(axis=2)
    assert np.mean(d > TOL) < 0.002
    assert oracle.psnr(r["gpu"], r["fp64"]) > 45.0
    assert r["stats"]["shaded"] > 0


SHADING_MODELS = list(range(2, 10))


@pytest.mark.parametrize("model", SHADING_MODELS)
def test_shading_models(ctx, oracle, model):
    """The shading models of mod_shading.frag (phong parts, toon, Cook-Torrance,
    Oren-Nayar, Ward) on the reference kernel (both filters) and on the cell
    kernel / automatic choice, against the FP32 oracle, same tolerance as Phong."""
    from voreen_amd import hipdvr as H, workloads as W
    dims = (64, 64, 64)
    vol = W.make_volume("c1", dims=dims)
    keys, _ = W.make_tf("c1")
    tf = oracle.tf_build(keys, 4096)
    cam = W.orbit_camera(9, 36)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 96, 96, llf=llf, urb=urb)
    sp = (1.0 / 64,) * 3
    off = tuple(float(llf[i] + 0.5 / 64) for i in range(3))
    name = H.SHADING_KEYS[model]
    for kernel, filt in ((H.KERNEL_REFERENCE, 0), (H.KERNEL_REFERENCE, 1), (H.KERNEL_CELL, 1),
                         (H.KERNEL_SLAB, 1), (H.KERNEL_AUTO, 1), (H.KERNEL_AUTO, 0)):
        say(f"shading {name} kernel {kernel} filter {filt} 64^3 u16 96^2")
        r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, W.step_for(dims, 2.0), shading=model,
                        filter_=filt, spacing=sp, offset=off, camera_pos=tuple(cam.eye), kernel=kernel)
        # relative beyond 1: Ward's specular term grows without bound at grazing
        # light or view (its denominator sqrt(L.N * V.N) tends to 0)
        d = (report(f"{name} k{kernel} f{filt}", r, oracle) / np.maximum(1.0, np.abs(r["cpu"]))).max(axis=2)
        assert np.mean(d > TOL) < 0.002
        assert r["stats"]["samples"] == r["n"] and r["stats"]["shaded"] > 0
        assert np.mean(r["gpu"][..., 3] > 0) > 0.05
    # the bricked kernel serves Phong only
    v = ctx.volume(vol)
    t = ctx.transfunc(tf)
    try:
        with pytest.raises(H.HipDvrError) as e:
            ctx.render(v, t, en, ex, H.default_params(0.01, kernel=H.KERNEL_BRICKED, shading=model))
        assert e.value.code == H.ERR_UNSUPPORTED
    finally:
        v.close()
        t.close()


def test_shading_model_with_isosurface_compositing(ctx, oracle):
    from voreen_amd import hipdvr as H
    vol, tf, en, ex, step = c0_inputs(oracle, size=96, view=3, nviews=36)
    for model in (H.SHADE_TOON, H.SHADE_WARD):
        say(f"iso compositing with {H.SHADING_KEYS[model]}")
        r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, step, shading=model,
                        compositing=H.COMPOSITE_ISO, iso_value=0.3)
        d = report(f"iso {model}", r, oracle).max(axis=2)
        assert np.mean(d > TOL) < 0.002
        assert np.mean(r["gpu"][..., 3] > 0) > 0.02


@pytest.mark.parametrize("integral", [False, True])
def test_preint_table_on_the_device(ctx, oracle, integral):
    """hipdvr_tf_preint_device against the oracle's table (the device's pow/exp
    are within an ulp of libm's) and the render with it."""
    from voreen_amd import hipdvr as H, workloads as W
    vol, tf, en, ex, step = c0_inputs(oracle, size=128)
    for keys, n in ((W.make_tf("c0")[0], 256), (W.random_keys(8, 7), 1024)):
        say(f"device preint integral={integral} n={n}")
        ref = oracle.preint_build(keys, n, step, use_integral=integral)
        t = ctx.transfunc(tf)
        v = ctx.volume(vol)
        try:
            t.preint_device(keys, n, step, use_integral=integral)
            img, st = ctx.render(v, t, en, ex, H.default_params(step, classification=H.CLASS_PREINT))
            t.set_preint(ref)
            img_ref, _ = ctx.render(v, t, en, ex, H.default_params(step, classification=H.CLASS_PREINT))
        finally:
            t.close()
            v.close()
        cpu, _ = oracle.render(vol, tf, (0.0, 1.0), en, ex, step, classification=1, preint=ref)
        print(f"n={n}: render max {np.abs(img - cpu).max():.3g}", flush=True)
        assert np.array_equal(img_ref, cpu)
        assert np.abs(img - cpu).max() <= TOL


@pytest.mark.parametrize("integral", [False, True])
def test_preintegrated_classification(ctx, oracle, integral):
    vol, tf, en, ex, step = c0_inputs(oracle, size=128)
    from voreen_amd import workloads as W
    keys, _ = W.make_tf("c0")
    table = oracle.preint_build(keys, 256, step, use_integral=integral)
    say(f"preint integral={integral} 128^2")
    r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, step, classification=1, preint=table)
    assert report(f"preint {integral}", r, oracle).max() <= TOL


def test_edge_cases(ctx, oracle, kernel):
    from voreen_amd import workloads as W
    rng = np.random.default_rng(5)
    keys = W.random_keys(99, 5)
    tf256 = oracle.tf_build(keys, 256)
    # degenerate rays, background, ragged image
    h, w = 23, 37
    en = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    ex = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    en[..., 3] = ex[..., 3] = 1.0
    en[0:3] = 0.0
    ex[0:3] = 0.0                     # background rows
    ex[5, :] = en[5, :]               # degenerate rays
    vol = rng.integers(0, 256, (9, 11, 13), dtype=np.uint8)
    say("edge: ragged 37x23 with background and degenerate rays")
    r = render_both(ctx, oracle, vol, tf256, (0.0, 1.0), en, ex, 1 / 26, kernel=kernel)
    assert report("ragged", r, oracle).max() <= TOL
    assert np.all(r["gpu"][0:3] == 0.0)
    # tiny volumes and a 1x1 image
    for shape in ((1, 1, 1), (5, 3, 2), (2, 1, 7)):
        v = rng.integers(0, 65536, shape, dtype=np.uint16)
        e1 = np.array([[[0.1, 0.2, 0.3, 1.0]]], np.float32)
        x1 = np.array([[[0.9, 0.7, 0.6, 1.0]]], np.float32)
        say(f"edge: volume {shape} 1x1")
        r = render_both(ctx, oracle, v, oracle.tf_build(keys, 4096), (0.0, 1.0), e1, x1, 1 / 14,
                        kernel=kernel)
        assert report(f"tiny {shape}", r, oracle).max() <= TOL
        assert r["stats"]["samples"] == r["n"]
    # loop cap, TF width 2, degenerate domain
    e1 = np.array([[[0.0, 0.5, 0.5, 1.0]]], np.float32)
    x1 = np.array([[[1.0, 0.5, 0.5, 1.0]]], np.float32)
    zero_tf = np.zeros((256, 4), np.uint8)
    say("edge: loop cap")
    r = render_both(ctx, oracle, vol, zero_tf, (0.0, 1.0), e1, x1, 1e-6, kernel=kernel)
    assert r["stats"]["samples"] == 255 * 255 == r["n"]
    say("edge: TF width 2, domain lo == hi")
    tf2 = np.array([[255, 0, 0, 30], [0, 255, 0, 200]], np.uint8)
    for dom in ((0.0, 1.0), (0.4, 0.4), (0.2, 0.3)):
        r = render_both(ctx, oracle, vol, tf2, dom, en, ex, 1 / 40, kernel=kernel)
        assert report(f"tf2 {dom}", r, oracle).max() <= TOL


def test_axis_aligned_rays_through_empty_bricks(ctx, oracle, kernel):
    """Rays along +-x, +-y, +-z across runs of empty bricks: exactly on an
    axis, with direction components of 1e-35 (below the 1e-30 under which an
    axis bounds no empty run), 1e-6 off the axis, and slightly oblique.  The
    empty-space bounds may skip only samples the oracle also finds
    transparent: every kernel equals the oracle bit for bit."""
    rng = np.random.default_rng(17)
    vol = np.zeros((72, 80, 96), np.uint16)             # z, y, x
    for z0, y0, x0 in ((0, 10, 20), (30, 50, 60), (46, 5, 70)):
        vol[z0:z0 + 24, y0:y0 + 28, x0:x0 + 22] = rng.integers(40000, 65536, (24, 28, 22), dtype=np.uint16)
    tf = np.zeros((4096, 4), np.uint8)
    tf[1500:] = (220, 140, 60, 90)                      # intensity 0: transparent, empty bricks
    us = np.linspace(0.03, 0.97, 16, dtype=np.float32)
    rows_en, rows_ex = [], []
    for axis in range(3):
        for sign in (1, -1):
            lo, hi = (0.0, 1.0) if sign > 0 else (1.0, 0.0)
            others = [i for i in range(3) if i != axis]
            for kind in ("exact", "denormal", "near", "oblique"):
                en = np.zeros((16, 4), np.float32)
                ex = np.zeros((16, 4), np.float32)
                en[:, 3] = ex[:, 3] = 1.0
                en[:, axis], ex[:, axis] = lo, hi
                for j, o in enumerate(others):
                    if kind == "denormal":
                        en[:, o], ex[:, o] = 0.0, 1e-35     # on the volume's face, |dir| ~ 1e-35
                    else:
                        d = {"exact": 0.0, "near": 1e-6, "oblique": 3e-3 * (1 + j)}[kind]
                        en[:, o] = us if j == 0 else us[::-1]
                        ex[:, o] = en[:, o] + np.float32(d)
                rows_en.append(en)
                rows_ex.append(ex)
    en = np.ascontiguousarray(np.stack(rows_en))
    ex = np.ascontiguousarray(np.stack(rows_ex))
    say(f"axis-aligned rays through empty bricks, kernel {kernel}")
    r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, 1 / 192, kernel=kernel)
    report("axis-aligned", r, oracle)
    assert np.array_equal(r["gpu"], r["cpu"])
    assert r["stats"]["samples"] == r["n"]
    assert np.count_nonzero(r["cpu"][..., 3]) > 0        # some rays meet the blobs


MODES = [("mip", 1, 0.0), ("mop", 2, 0.0), ("mida", 3, 0.0), ("mida", 3, -0.5), ("mida", 3, 0.6),
         ("iso", 4, 0.0), ("fhp", 5, 0.0), ("fhn", 6, 0.0)]


@pytest.mark.parametrize("mode,code,gamma", MODES, ids=[f"{m}{g}" for m, _, g in MODES])
def test_compositing_modes(ctx, oracle, mode, code, gamma):
    """MIP/MOP/MIDA/ISO/FHP/FHN (mod_compositing.frag) against the oracle on the
    blob volume and the walnut fixture, both filters, with and without Phong."""
    from voreen_amd import workloads as W
    vol, tf, en, ex, step = c0_inputs(oracle, size=96, view=7, nviews=36)
    for filt in (0, 1):
        for shading in (0, 1):
            say(f"compositing {mode} gamma {gamma} filter {filt} shading {shading}")
            r = render_both(ctx, oracle, vol, tf, (0.0, 1.0), en, ex, step, filter_=filt, shading=shading,
                            compositing=code, gamma=gamma, iso_value=0.3)
            d = report(f"{mode} f{filt} s{shading}", r, oracle).max(axis=2)
            assert np.mean(d > TOL) < 0.002 if shading else d.max() <= TOL
            assert r["stats"]["samples"] == r["n"]
    v = F.load_vvd("walnut")
    t = F.load_tfi("walnut")
    wtf = oracle.tf_build(t["keys"], t["width"], t["gamma"], t["threshold"], t["alpha_mode"])
    llf, urb = W.world_box(v["dims"], v["spacing"])
    cam = W.orbit_camera(5, 24, radius=2.5)
    wen, wex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 96, 72, llf=llf, urb=urb)
    say(f"compositing {mode} walnut")
    # MIDA weighs the real-world intensity directly (beta = 1 - delta*(1+gamma),
    # mod_compositing.frag:116-119): it is meaningful on normalized values, so
    # its walnut case uses the identity mapping and the domain in [0, 1]
    rwm, dom = ((1.0, 0.0), tuple(d / 65535.0 for d in t["domain"])) if mode == "mida" else \
        ((65535.0, 0.0), t["domain"])
    r = render_both(ctx, oracle, v["data"], wtf, dom, wen, wex, W.step_for(v["dims"], 2.0),
                    rwm=rwm, compositing=code, gamma=gamma, iso_value=0.3)
    assert report(f"{mode} walnut", r, oracle).max() <= TOL
    assert np.mean(r["gpu"][..., 3] > 0) > 0.02


def test_compositing_modes_reject_skipping_kernels(ctx):
    from voreen_amd import hipdvr as H
    v = ctx.volume(np.zeros((4, 4, 4), np.uint8))
    t = ctx.transfunc(np.zeros((256, 4), np.uint8))
    en = np.zeros((2, 2, 4), np.float32)
    say("compositing modes: invalid combinations")
    try:
        for k in (H.KERNEL_BRICKED, H.KERNEL_CELL, H.KERNEL_SLAB):
            with pytest.raises(H.HipDvrError) as e:
                ctx.render(v, t, en, en, H.default_params(0.1, kernel=k, compositing=H.COMPOSITE_MIP))
            assert e.value.code == H.ERR_UNSUPPORTED
        with pytest.raises(H.HipDvrError) as e:
            ctx.render(v, t, en, en, H.default_params(0.1, compositing=9))
        assert e.value.code == H.ERR_UNSUPPORTED
    finally:
        v.close()
        t.close()


def test_invalid_arguments_return_codes(ctx):
    from voreen_amd import hipdvr as H
    vol = np.zeros((4, 4, 4), np.uint8)
    v = ctx.volume(vol)
    t = ctx.transfunc(np.zeros((256, 4), np.uint8))
    en = np.zeros((2, 2, 4), np.float32)
    say("invalid arguments")
    try:
        with pytest.raises(H.HipDvrError) as e:
            ctx.render(v, t, en, en, H.default_params(0.0))
        assert e.value.code == H.ERR_INVALID
        with pytest.raises(H.HipDvrError) as e:
            ctx.render(v, t, en, en, H.default_params(0.1, classification=H.CLASS_PREINT))
        assert e.value.code == H.ERR_INVALID
        p = H.default_params(0.1)
        p.compositing = 42
        with pytest.raises(H.HipDvrError) as e:
            ctx.render(v, t, en, en, p)
        assert e.value.code == H.ERR_UNSUPPORTED
        with pytest.raises(H.HipDvrError):
            ctx.transfunc(np.zeros((5000, 4), np.uint8))
        with pytest.raises(H.HipDvrError):
            ctx.transfunc(np.zeros((256, 4), np.uint8), domain=(1.0, 0.0))
    finally:
        v.close()
        t.close()


def test_bricked_equals_reference_and_skips(ctx, oracle):
    """The bricked and cell kernels must produce the reference kernel's image
    bit for bit (the cell kernel also when shaded) while reconstructing fewer
    samples; a TF change must refresh the brick occupancy (same volume,
    second TF)."""
    from voreen_amd import hipdvr as H, workloads as W
    dims = (160, 144, 128)
    vol = W.make_volume("c1", dims=dims)
    keys, _ = W.make_tf("c1")
    cam = W.orbit_camera(11, 36)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 200, 176, llf=llf, urb=urb)
    step = W.step_for(dims, 2.0)
    v = ctx.volume(vol)
    try:
        for seed in (0, 1):
            tfk = keys if seed == 0 else W.random_keys(1234, 6)
            tf = oracle.tf_build(tfk, 4096)
            t = ctx.transfunc(tf)
            outs = {}
            runs = [("ref", H.KERNEL_REFERENCE, 0, 0), ("brick", H.KERNEL_BRICKED, 0, 0)]
            runs += [(f"cell{b}", H.KERNEL_CELL, b, 1) for b in (1, 2, 4, 8)]
            runs += [("cellfast", H.KERNEL_CELL, 4, 0)]
            for name, k, batch, exact in runs:
                for shading in (0, 1):
                    say(f"bricked-vs-reference tf{seed} {name} shading {shading}")
                    p = H.default_params(step, shading=shading, kernel=k, camera_pos=tuple(cam.eye),
                                         batch=batch, exact_shading=exact)
                    outs[(name, shading)] = ctx.render(v, t, en, ex, p)
            t.close()
            for shading in (0, 1):
                a, sa = outs[("ref", shading)]
                for name in ("cell1", "cell2", "cell4", "cell8"):
                    c, sc = outs[(name, shading)]
                    assert np.array_equal(a, c), f"{name} tf{seed} shading {shading}"
                    assert sa["samples"] == sc["samples"] and sa["shaded"] == sc["shaded"]
                    assert sc["fetched"] < sa["fetched"] or seed == 1
                c, sc = outs[("cellfast", shading)]
                if shading == 0:
                    assert np.array_equal(a, c)
                else:
                    d = np.abs(a - c).max(axis=2)
                    print(f"tf{seed} phong cell-fast-vs-reference max {d.max():.3g}", flush=True)
                    assert np.mean(d > TOL) < 0.002
                b, sb = outs[("brick", shading)]
                if shading == 0:
                    assert np.array_equal(a, b), f"tf{seed} shading {shading}"
                    assert sa["shaded"] == sb["shaded"]
                else:
                    # the bricked kernel evaluates the central differences from
                    # plane bilinears: equal up to float rounding
                    d = np.abs(a - b).max(axis=2)
                    print(f"tf{seed} phong bricked-vs-reference max {d.max():.3g}", flush=True)
                    assert np.mean(d > TOL) < 0.002
                assert sa["samples"] == sb["samples"]
                print(f"tf{seed} shading {shading}: fetched {sb['fetched']} of {sb['samples']}",
                      flush=True)
                if seed == 0:
                    assert sb["fetched"] < 0.8 * sb["samples"]
            cpu, n = oracle.render(vol, tf, (0.0, 1.0), en, ex, step)
            assert np.abs(outs[("brick", 0)][0] - cpu).max() <= TOL
    finally:
        v.close()


def test_bricked_all_transparent_fetches_nothing(ctx, oracle):
    from voreen_amd import hipdvr as H, workloads as W
    vol = W.make_volume("c0", dims=(64, 64, 64))
    tf = np.zeros((256, 4), np.uint8)
    tf[:, :3] = 200
    en, ex = oracle.entry_exit((0.5, 0.6, 3.0), (0.5, 0.5, 0.5), (0, 1, 0), 30.0, 0.1, 64, 64)
    say("bricked all-transparent TF")
    v, t = ctx.volume(vol), ctx.transfunc(tf)
    try:
        img, st = ctx.render(v, t, en, ex, H.default_params(1 / 128, kernel=H.KERNEL_BRICKED))
    finally:
        v.close()
        t.close()
    assert st["fetched"] == 0 and st["samples"] > 0 and np.all(img == 0)


@pytest.mark.parametrize("brick_log2", [2, 3, 4, 5])
def test_bricked_brick_sizes(ctx, oracle, brick_log2):
    from voreen_amd import hipdvr as H, workloads as W
    vol = W.make_volume("c0", dims=(70, 50, 90))
    keys, tf = W.make_tf("c0")
    cam = W.orbit_camera(20, 36)
    llf, urb = W.world_box((70, 50, 90))
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 120, 90, llf=llf, urb=urb)
    step = W.step_for((70, 50, 90), 2.0)
    say(f"bricked brick_log2 {brick_log2}")
    v, t = ctx.volume(vol, brick_log2=brick_log2), ctx.transfunc(tf)
    try:
        img, st = ctx.render(v, t, en, ex, H.default_params(step, kernel=H.KERNEL_BRICKED))
    finally:
        v.close()
        t.close()
    cpu, n = oracle.render(vol, tf, (0.0, 1.0), en, ex, step)
    assert np.abs(img - cpu).max() <= TOL and st["samples"] == n


@pytest.mark.parametrize("name", ["nucleon", "walnut"])
def test_vvd_load_and_volume_minmax(ctx, oracle, name):
    """A volume loaded from the .vvd renders like the same voxels uploaded from
    memory, and its min/max match the fixture's VolumeMinMax derived data."""
    from voreen_amd import hipdvr as H, workloads as W
    ref = F.load_vvd(name)
    say(f"vvd load {name}")
    v = ctx.load_vvd(os.path.join(F.GOLDEN, name + ".vvd"))
    vm = ctx.volume(ref["data"])
    tf = oracle.tf_build(W.random_keys(3, 5), 256 if ref["data"].dtype == np.uint8 else 4096)
    t = ctx.transfunc(tf)
    try:
        mn = v.minmax()
        assert abs(mn[0] - ref["derived"]["min"]) <= 1e-7 and abs(mn[1] - ref["derived"]["max"]) <= 1e-7
        assert mn[0] == np.float32(ref["data"].min()) / np.float32(255 if ref["data"].dtype == np.uint8 else 65535)
        llf, urb = W.world_box(ref["dims"], ref["spacing"])
        cam = W.orbit_camera(3, 12)
        en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 64, 64, llf=llf, urb=urb)
        p = H.default_params(W.step_for(ref["dims"], 2.0))
        a, _ = ctx.render(v, t, en, ex, p)
        b, _ = ctx.render(vm, t, en, ex, p)
        assert np.array_equal(a, b) and np.mean(a[..., 3] > 0) > 0.01
    finally:
        v.close()
        vm.close()
        t.close()


GRADIENTS = [(1, "forward"), (2, "none"), (3, "filtered"), (4, "sobel")]


@pytest.mark.parametrize("mode,name", GRADIENTS, ids=[n for _, n in GRADIENTS])
def test_gradient_modes(ctx, oracle, mode, name):
    """The "gradient" property's modes (mod_gradients.frag) under Phong and
    FHN on the reference kernel, both filters, against the oracle; the
    skipping kernels refuse them."""
    from voreen_amd import hipdvr as H, workloads as W
    dims = (48, 48, 48)
    vol = W.make_volume("c1", dims=dims)
    keys, _ = W.make_tf("c1")
    tf = oracle.tf_build(keys, 4096)
    cam = W.orbit_camera(11, 36)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 80, 80, llf=llf, urb=urb)
    sp = (1.0 / 48,) * 3
    off = tuple(float(llf[i] + 0.5 / 48) for i in range(3))
    step = W.step_for(dims, 2.0)
    for kernel, filt, shading, comp in ((H.KERNEL_REFERENCE, 0, 1, 0), (H.KERNEL_REFERENCE, 1, 1, 0),
                                        (H.KERNEL_AUTO, 1, 8, 0), (H.KERNEL_AUTO, 1, 0, H.COMPOSITE_FHN)):
        say(f"gradient {name} kernel {kernel} filter {filt} shading {shading} compositing {comp}")
        v = ctx.volume(vol, spacing=sp, offset=off)
        t = ctx.transfunc(tf)
        try:
            gpu, st = ctx.render(v, t, en, ex, H.default_params(step, filter_=filt, shading=shading, kernel=kernel,
                                                               camera_pos=tuple(cam.eye), compositing=comp,
                                                               gradient=mode))
        finally:
            v.close()
            t.close()
        cpu, n = oracle.render(vol, tf, (0.0, 1.0), en, ex, step, filter_=filt, shading=shading,
                               light=dict(camera=tuple(cam.eye)), spacing=sp, offset=off, compositing=comp,
                               gradient=mode)
        d = np.abs(gpu - cpu).max(axis=2)
        print(f"{name} k{kernel} f{filt} s{shading} c{comp}: max_abs={d.max():.3g} "
              f"exact={np.mean(np.all(gpu == cpu, axis=2)):.4f}", flush=True)
        assert np.mean(d > TOL) < 0.002
        assert st["samples"] == n
        assert np.mean(gpu[..., 3] > 0) > 0.05
    v = ctx.volume(vol)
    t = ctx.transfunc(tf)
    try:
        for k in (H.KERNEL_BRICKED, H.KERNEL_CELL, H.KERNEL_SLAB):
            with pytest.raises(H.HipDvrError) as e:
                ctx.render(v, t, en, ex, H.default_params(step, kernel=k, shading=1, gradient=mode))
            assert e.value.code == H.ERR_UNSUPPORTED
        # unshaded DVR never evaluates a gradient: every kernel accepts any mode
        img, _ = ctx.render(v, t, en, ex, H.default_params(step, kernel=H.KERNEL_CELL, gradient=mode))
        cpu, _ = oracle.render(vol, tf, (0.0, 1.0), en, ex, step)
        assert np.abs(img - cpu).max() <= TOL
    finally:
        v.close()
        t.close()


def tf2d_inputs(oracle, gdtype=np.float32, channels=3):
    """Shells volume, a central-difference gradient volume of it and a 2D TF
    that makes boundaries (large gradient magnitude) opaque."""
    from voreen_amd import workloads as W
    dims = (48, 40, 36)
    vol = W.make_volume("c1", dims=dims)
    f = vol.astype(np.float64) / 65535.0
    g = np.stack([np.gradient(f, axis=2), np.gradient(f, axis=1), np.gradient(f, axis=0)], axis=3)
    g = np.abs(g) * 4.0
    if channels == 4:
        g = np.concatenate([g, np.full(g.shape[:3] + (1,), 0.7)], axis=3)
    gmax = {np.float32: 1.0, np.uint8: 255.0, np.uint16: 65535.0}[gdtype]
    grad = g.astype(np.float32) if gdtype == np.float32 else np.clip(np.round(g * gmax), 0, gmax).astype(gdtype)
    W2, H2 = 64, 32
    xi = np.linspace(0, 1, W2)[None, :]
    yg = np.linspace(0, 1, H2)[:, None]
    tf2d = np.zeros((H2, W2, 4), np.float32)
    tf2d[..., 0] = xi
    tf2d[..., 1] = 1.0 - xi * 0.5
    tf2d[..., 2] = yg
    tf2d[..., 3] = np.clip(yg * 1.5 - 0.1, 0, 1) * (xi > 0.15)
    cam = W.orbit_camera(2, 36)
    llf, urb = W.world_box(dims)
    en, ex = oracle.entry_exit(cam.eye, cam.focus, cam.up, 30.0, 0.1, 72, 64, llf=llf, urb=urb)
    sp = tuple(1.0 / max(dims) for _ in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    return vol, grad, tf2d, ((0.0, 1.0), (0.0, 0.4)), en, ex, W.step_for(dims, 2.0), sp, off, tuple(cam.eye)


@pytest.mark.parametrize("gdtype,channels", [(np.float32, 3), (np.uint8, 4), (np.uint16, 3)],
                         ids=["f32x3", "u8x4", "u16x3"])
def test_tf2d_classification(ctx, oracle, gdtype, channels):
    """The CPURaycaster 2D TF path (intensity x gradient-volume magnitude)
    against the oracle: bit-exact unshaded, Phong tolerance shaded."""
    from voreen_amd import hipdvr as H
    vol, grad, tf2d, dom, en, ex, step, sp, off, eye = tf2d_inputs(oracle, gdtype, channels)
    for filt in (0, 1):
        for shading in (0, 1):
            say(f"tf2d grad {np.dtype(gdtype).name}x{channels} filter {filt} shading {shading}")
            v = ctx.volume(vol, spacing=sp, offset=off)
            v.set_gradient(grad)
            t = ctx.transfunc2d(tf2d, dom)
            try:
                gpu, st = ctx.render(v, t, en, ex, H.default_params(step, filter_=filt, shading=shading,
                                                                   classification=H.CLASS_TF2D,
                                                                   camera_pos=eye))
                # a second step size rebuilds the opacity-corrected table
                gpu2, _ = ctx.render(v, t, en, ex, H.default_params(step * 0.5, filter_=filt,
                                                                    classification=H.CLASS_TF2D))
            finally:
                v.close()
                t.close()
            cpu, n = oracle.render(vol, None, (0.0, 1.0), en, ex, step, filter_=filt, shading=shading,
                                   classification=oracle.CLASS_TF2D, grad_volume=grad, tf2d=tf2d,
                                   tf2d_domain=dom, light=dict(camera=eye), spacing=sp, offset=off)
            cpu2, _ = oracle.render(vol, None, (0.0, 1.0), en, ex, step * 0.5, filter_=filt,
                                    classification=oracle.CLASS_TF2D, grad_volume=grad, tf2d=tf2d,
                                    tf2d_domain=dom, spacing=sp, offset=off)
            d = np.abs(gpu - cpu).max(axis=2)
            print(f"tf2d f{filt} s{shading}: max_abs={d.max():.3g} "
                  f"exact={np.mean(np.all(gpu == cpu, axis=2)):.4f}", flush=True)
            if shading:
                assert np.mean(d > TOL) < 0.002
            else:
                assert d.max() <= TOL
            assert np.abs(gpu2 - cpu2).max() <= TOL
            assert st["samples"] == n
            assert np.mean(gpu[..., 3] > 0) > 0.05


def test_tf2d_rejections(ctx, oracle):
    from voreen_amd import hipdvr as H
    vol, grad, tf2d, dom, en, ex, step, sp, off, eye = tf2d_inputs(oracle)
    say("tf2d: invalid combinations")
    v = ctx.volume(vol)
    t2 = ctx.transfunc2d(tf2d, dom)
    t1 = ctx.transfunc(np.zeros((256, 4), np.uint8))
    try:
        cases = [(t2, dict(), H.ERR_INVALID),                                   # 2D TF, 1D classification
                 (t1, dict(classification=H.CLASS_TF2D), H.ERR_INVALID),        # 1D TF, 2D classification
                 (t2, dict(classification=H.CLASS_TF2D), H.ERR_INVALID)]        # no gradient volume yet
        for t, kw, code in cases:
            with pytest.raises(H.HipDvrError) as e:
                ctx.render(v, t, en, ex, H.default_params(step, **kw))
            assert e.value.code == code
        v.set_gradient(grad)
        for kw in (dict(kernel=H.KERNEL_CELL), dict(kernel=H.KERNEL_BRICKED), dict(kernel=H.KERNEL_SLAB),
                   dict(compositing=H.COMPOSITE_MIP)):
            with pytest.raises(H.HipDvrError) as e:
                ctx.render(v, t2, en, ex, H.default_params(step, classification=H.CLASS_TF2D, **kw))
            assert e.value.code == H.ERR_UNSUPPORTED
        with pytest.raises(ValueError):
            v.set_gradient(grad[:, :, :-1])
    finally:
        v.close()
        t1.close()
        t2.close()
