"""Volumes whose cell records exceed 4 GiB (64-bit record addressing in the
per-frame, batched and learned kernels) and the device turbulence generator
of the large-volume workload.

A 640 x 640 x 656 uint16 turbulence volume (537 MB; 4.3 GB of cell records)
is made on the device.  The cell kernel equals the reference kernel bit for
bit (unshaded and with exact Phong), the batched launch equals the single
frames, the learned baseline matches the plain image, and a subset of rays
(every 8th pixel) equals the CPU oracle on the downloaded volume.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-3
DIMS = (640, 640, 656)          # x, y, z


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def big(ctx):
    from voreen_amd import hipdvr as H, workloads as W
    say(f"turbulence {DIMS} on the device")
    buf, frac = ctx.synth_turbulence(0xC4, DIMS, 0.4)
    llf, urb = W.world_box(DIMS)
    sp = tuple(float((urb[i] - llf[i]) / DIMS[i]) for i in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    v = ctx.volume_device(buf, DIMS, H.U16, spacing=sp, offset=off)
    keys, tfa = W.make_tf("c4")
    t = ctx.transfunc(tfa)
    w = h = 128
    cam = W.orbit_camera(3, 36)
    de, dx, do, dr, dl = (ctx.alloc(w * h * 16) for _ in range(5))
    ctx.entry_exit_device(cam, llf, urb, w, h, de, dx)
    yield dict(buf=buf, frac=frac, v=v, t=t, tfa=tfa, keys=keys, de=de, dx=dx, do=do, dr=dr, dl=dl, w=w, h=h,
               cam=cam, step=W.step_for(DIMS, 2.0), sp=sp, off=off)
    for b in (de, dx, do, dr, dl, buf):
        b.free()
    v.close()
    t.close()


def render(ctx, s, kernel, shading=0, exact=0, out="do"):
    from voreen_amd import hipdvr as H
    p = H.default_params(s["step"], kernel=kernel, shading=shading, exact_shading=exact,
                         camera_pos=tuple(s["cam"].eye))
    st = ctx.render_device(s["v"], s["t"], s["de"], s["dx"], s["w"], s["h"], p, s[out], stats=True)
    return s[out].download((s["h"], s["w"], 4), np.float32), st


def test_turbulence_generator(ctx, big):
    """Empty fraction at the 4096-bin boundary nearest 40%, values of the rest
    in [1, 65535], and the same seed gives the same volume."""
    assert abs(big["frac"] - 0.4) < 0.01
    v = big["buf"].download((DIMS[2], DIMS[1], DIMS[0]), np.uint16)
    assert abs(float(np.mean(v == 0)) - big["frac"]) < 1e-12
    assert v.max() > 65535 - 65535 // 256      # the top histogram bin maps onto the top of the range
    buf2, frac2 = ctx.synth_turbulence(0xC4, (96, 80, 72), 0.4)
    buf3, frac3 = ctx.synth_turbulence(0xC4, (96, 80, 72), 0.4)
    try:
        a = buf2.download((72, 80, 96), np.uint16)
        b = buf3.download((72, 80, 96), np.uint16)
        assert np.array_equal(a, b) and frac2 == frac3
    finally:
        buf2.free()
        buf3.free()


def test_wide_cell_kernel_equals_reference(ctx, big):
    from voreen_amd import hipdvr as H
    for shading, exact in ((0, 0), (1, 1)):
        say(f"wide cell vs reference shading {shading}")
        ref, sr = render(ctx, big, H.KERNEL_REFERENCE, shading, exact, out="dr")
        img, sc = render(ctx, big, H.KERNEL_CELL, shading, exact)
        print(f"shading {shading}: max_abs={np.abs(img - ref).max():.3g} coverage={np.mean(ref[..., 3] > 0):.3f} "
              f"samples {sr['samples']} fetched {sc['fetched']}", flush=True)
        assert np.array_equal(img, ref)
        assert sc["samples"] == sr["samples"] and sc["fetched"] < sr["fetched"]
        assert np.mean(ref[..., 3] > 0) > 0.2


def test_wide_rays_equal_the_oracle(ctx, big, oracle):
    from voreen_amd import hipdvr as H
    say("wide: every 8th pixel against the oracle")
    img, _ = render(ctx, big, H.KERNEL_AUTO)
    vol = big["buf"].download((DIMS[2], DIMS[1], DIMS[0]), np.uint16)
    en = big["de"].download((big["h"], big["w"], 4), np.float32)[::8, ::8]
    ex = big["dx"].download((big["h"], big["w"], 4), np.float32)[::8, ::8]
    cpu, n = oracle.render(vol, big["tfa"], (0.0, 1.0), en, ex, big["step"], spacing=big["sp"], offset=big["off"])
    d = np.abs(img[::8, ::8] - cpu)
    print(f"oracle subset: max_abs={d.max():.3g} rays={int(np.sum(cpu[..., 3] > 0))}", flush=True)
    assert d.max() <= TOL and np.mean(cpu[..., 3] > 0) > 0.2


def test_wide_batch_equals_single_frames(ctx, big):
    from voreen_amd import hipdvr as H, workloads as W
    say("wide batch: the large volume and a small one in one launch")
    ref, _ = render(ctx, big, H.KERNEL_CELL, 1, 0, out="dr")
    small = H.synth_blobs(5, (40, 56, 48), H.U16)
    vs = ctx.volume(small)
    ts = ctx.transfunc(H.tf_from_keys(W.random_keys(9, 5), 4096))
    w2 = h2 = 64
    cam2 = W.orbit_camera(7, 36)
    llf, urb = W.world_box((40, 56, 48))
    de2, dx2, do2, dr2 = (ctx.alloc(w2 * h2 * 16) for _ in range(4))
    try:
        ctx.entry_exit_device(cam2, llf, urb, w2, h2, de2, dx2)
        p1 = H.default_params(big["step"], shading=1, camera_pos=tuple(big["cam"].eye))
        p2 = H.default_params(W.step_for((40, 56, 48), 2.0), shading=1, camera_pos=tuple(cam2.eye))
        ctx.render_device(vs, ts, de2, dx2, w2, h2, H.default_params(p2.step, shading=1, kernel=H.KERNEL_CELL,
                                                                     camera_pos=tuple(cam2.eye)), dr2)
        b = H.Batch(ctx, [(big["v"], big["t"], big["de"], big["dx"], big["do"], big["w"], big["h"], p1),
                          (vs, ts, de2, dx2, do2, w2, h2, p2)])
        try:
            for tile, bl, k in ((0, 4, 2), (2, 5, 4)):
                b.set_config(tile, bl, k)
                b.launch()
                ctx.synchronize()
                assert np.array_equal(big["do"].download((big["h"], big["w"], 4), np.float32), ref)
                assert np.array_equal(do2.download((h2, w2, 4), np.float32), dr2.download((h2, w2, 4), np.float32))
        finally:
            b.close()
    finally:
        for x in (de2, dx2, do2, dr2):
            x.free()
        vs.close()
        ts.close()


def test_wide_learned_baseline_matches_plain(ctx, big):
    from voreen_amd import hipdvr as H
    say("wide learned baseline")
    plain, _ = render(ctx, big, H.KERNEL_AUTO, 1, 0, out="dr")
    p = H.default_params(big["step"], shading=1, camera_pos=tuple(big["cam"].eye))
    l = H.Learned(ctx, big["v"])
    try:
        l.prepare(big["t"], p, None)
        ls = l.render_device(big["t"], big["de"], big["dx"], big["w"], big["h"], p, big["dl"], stats=True)
        img = big["dl"].download((big["h"], big["w"], 4), np.float32)
    finally:
        l.close()
    mse = float(np.mean((img.astype(np.float64) - plain) ** 2))
    q = float("inf") if mse == 0 else 10 * np.log10(1 / mse)
    print(f"learned baseline: psnr {q:.1f} fell_back {ls['fell_back']}", flush=True)
    assert q > 60.0 and not ls["fell_back"]
