"""Learned stepping end to end on the GPU: the no-network table (every brick
at the base step) renders the plain frame itself, a network trained on plain
traces (the library's own kernels) keeps >= 45 dB while reconstructing fewer
samples than the plain march, a constant double-step network marches a wide
frame completely, and bad weights trip the sparse judge so the frame falls
back to the plain image exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1 / mse)


@pytest.fixture(scope="module")
def scene(ctx):
    from voreen_amd import hipdvr as H, workloads as W
    dims = (128, 128, 128)
    vol = W.make_volume("c1", dims=dims)
    llf, urb = W.world_box(dims)
    sp = tuple(float((urb[i] - llf[i]) / dims[i]) for i in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    v = ctx.volume(vol, spacing=sp, offset=off)
    keys, _ = W.make_tf("c1")
    t = ctx.transfunc(H.tf_from_keys(keys, 4096))
    w = h = 256
    cam = W.orbit_camera(5, 36)
    de, dx, do, dl = (ctx.alloc(w * h * 16) for _ in range(4))
    ctx.entry_exit_device(cam, llf, urb, w, h, de, dx)
    step = W.step_for(dims, 2.0)
    yield dict(v=v, t=t, de=de, dx=dx, do=do, dl=dl, w=w, h=h, step=step, cam=cam, dims=dims)
    for b in (de, dx, do, dl):
        b.free()
    v.close()
    t.close()


def _plain(ctx, s, shading):
    from voreen_amd import hipdvr as H
    p = H.default_params(s["step"], shading=shading, camera_pos=tuple(s["cam"].eye))
    st = ctx.render_device(s["v"], s["t"], s["de"], s["dx"], s["w"], s["h"], p, s["do"], stats=True)
    return s["do"].download((s["h"], s["w"], 4), np.float32), st, p


def test_learned_baseline_matches_plain(ctx, scene):
    from voreen_amd import hipdvr as H
    print("[case] learned baseline (no network)", flush=True)
    plain, st, p = _plain(ctx, scene, 0)
    l = H.Learned(ctx, scene["v"])
    try:
        l.prepare(scene["t"], p, None)
        ls = l.render_device(scene["t"], scene["de"], scene["dx"], scene["w"], scene["h"], p, scene["dl"],
                             stats=True)
        img = scene["dl"].download((scene["h"], scene["w"], 4), np.float32)
    finally:
        l.close()
    q = psnr(img, plain)
    print(f"baseline: psnr {q:.1f} dB sparse {ls['sparse_psnr']:.1f} samples {ls['samples']} "
          f"(plain fetched {st['fetched']}) fell_back {ls['fell_back']}", flush=True)
    # a table that keeps every occupied brick at the base step cannot gain:
    # the plain frame itself, without the judge
    assert np.array_equal(img, plain) and not ls["fell_back"] and ls["sparse_psnr"] == 999.0
    assert ls["samples"] == st["fetched"]


def test_learned_march_wide_frame(ctx, scene):
    """A 560x136 frame spans more than one chip region of tiles, so the learned
    march takes its tiles in per-XCD regions (partial ones at the edges).  A
    network with constant outputs (no skip, step multiple 2 in every brick)
    makes the learned march run; with the floor at 20 dB its own image is
    kept: every pixel written, close to the plain frame, fewer samples."""
    from voreen_amd import hipdvr as H, workloads as W
    print("[case] learned march 560x136, constant double-step network", flush=True)
    w, h = 560, 136
    llf, urb = W.world_box(scene["dims"])
    bufs = [ctx.alloc(w * h * 16) for _ in range(4)]
    de, dx, do, dl = bufs
    l = H.Learned(ctx, scene["v"])
    m = H.Mlp(ctx, seed=5)
    try:
        prm = m.params()
        prm[-130:-2] = 0.0          # last layer's weights: outputs = its biases
        prm[-2], prm[-1] = -50.0, 1.0   # no skip; log2 step multiple 1
        m.set_params(prm)
        ctx.entry_exit_device(W.orbit_camera(8, 36, radius=1.6), llf, urb, w, h, de, dx)
        p = H.default_params(scene["step"], camera_pos=tuple(scene["cam"].eye))
        st = ctx.render_device(scene["v"], scene["t"], de, dx, w, h, p, do, stats=True)
        plain = do.download((h, w, 4), np.float32)
        l.prepare(scene["t"], p, m)
        dl.upload(np.full((h, w, 4), np.nan, np.float32))
        ls = l.render_device(scene["t"], de, dx, w, h, p, dl, opts=H.learned_opts(psnr_floor=20.0), stats=True)
        img = dl.download((h, w, 4), np.float32)
    finally:
        l.close()
        m.close()
        for b in bufs:
            b.free()
    assert not np.isnan(img).any()
    assert np.mean(plain[..., 3] > 0) > 0.2
    q = psnr(img, plain)
    print(f"double step: psnr {q:.1f} dB, samples {ls['samples']} (plain {st['fetched']}), "
          f"fell_back {ls['fell_back']}", flush=True)
    assert not ls["fell_back"] and q > 30.0 and 0 < ls["samples"] < st["fetched"]


def test_trained_network_gains_over_baseline(ctx, scene):
    from voreen_amd import hipdvr as H, stepper
    print("[case] train stepping network (24 TFs x 8 views) and render", flush=True)
    # thresholds below the library's defaults: a small network trained on few
    # traces of a 128^3 volume, checked against the 45 dB floor
    mlp, info = stepper.train(ctx, scene["v"], scene["dims"], n_tfs=24, n_views=8, trace_size=128,
                              epochs=6, tau_step=4e-3, tau_skip=2e-3)
    print(f"training: {info}", flush=True)
    try:
        assert info["rows"] > 1000 and info["loss_last"] < info["loss_first"]
        results = {}
        for shading in (0, 1):
            plain, st, p = _plain(ctx, scene, shading)
            l = H.Learned(ctx, scene["v"])
            try:
                for name, net in (("baseline", None), ("learned", mlp)):
                    l.prepare(scene["t"], p, net)
                    ls = l.render_device(scene["t"], scene["de"], scene["dx"], scene["w"], scene["h"], p,
                                         scene["dl"], stats=True)
                    img = scene["dl"].download((scene["h"], scene["w"], 4), np.float32)
                    results[(shading, name)] = (psnr(img, plain), ls)
                    print(f"shading {shading} {name}: psnr {psnr(img, plain):.1f} dB, "
                          f"samples {ls['samples']}, sparse {ls['sparse_psnr']:.1f}, "
                          f"fell_back {ls['fell_back']}", flush=True)
            finally:
                l.close()
        for shading in (0, 1):
            qb, lb = results[(shading, "baseline")]
            ql, ll = results[(shading, "learned")]
            if not ll["fell_back"]:
                assert ql >= 45.0
                assert ll["samples"] < lb["samples"]
        assert any(not results[(s, "learned")][1]["fell_back"] for s in (0, 1))
    finally:
        mlp.close()


def test_bad_weights_fall_back_to_plain(ctx, scene):
    from voreen_amd import hipdvr as H
    print("[case] learned fallback with a skip-everything network", flush=True)
    plain, st, p = _plain(ctx, scene, 0)
    m = H.Mlp(ctx, seed=3)
    l = H.Learned(ctx, scene["v"])
    try:
        w = m.params()
        w[-2] = 50.0      # b3[0]: skip logit for every brick
        m.set_params(w)
        l.prepare(scene["t"], p, m)
        ls = l.render_device(scene["t"], scene["de"], scene["dx"], scene["w"], scene["h"], p, scene["dl"],
                             stats=True)
        img = scene["dl"].download((scene["h"], scene["w"], 4), np.float32)
    finally:
        l.close()
        m.close()
    print(f"fallback: sparse psnr {ls['sparse_psnr']:.1f} fell_back {ls['fell_back']}", flush=True)
    assert ls["fell_back"] == 1 and ls["sparse_psnr"] < 45.0
    assert np.array_equal(img, plain)


@pytest.mark.parametrize("shading", [0, 1])
def test_sparse_remarch_equals_plain(ctx, scene, shading):
    """The judge's segmented plain re-march equals the plain frame at the
    sparse pixels up to the float rounding of the segment joins."""
    from voreen_amd import hipdvr as H
    print(f"[case] sparse re-march shading {shading}", flush=True)
    plain, st, p = _plain(ctx, scene, shading)
    l = H.Learned(ctx, scene["v"])
    gw, gh = H.learned_sparse_dims(scene["w"], scene["h"])
    ds = ctx.alloc(gw * gh * 16)
    small = ctx.alloc(gw * gh * 16 - 16)
    try:
        # one pixel short of the grid: refused before any launch
        with pytest.raises(H.HipDvrError):
            l.sparse_remarch(scene["t"], scene["de"], scene["dx"], scene["w"], scene["h"], p, small)
        sw, sh = l.sparse_remarch(scene["t"], scene["de"], scene["dx"], scene["w"], scene["h"], p, ds)
        assert (sw, sh) == (gw, gh)
        sparse = ds.download((sh, sw, 4), np.float32)
    finally:
        small.free()
        ds.free()
        l.close()
    # the judge's pixels: blocks of 8 in every 64 along each axis, from 28
    offs = min(28, scene["w"] - 1, scene["h"] - 1)
    xs = [(g // 8) * 64 + offs + g % 8 for g in range(sw)]
    ys = [(g // 8) * 64 + offs + g % 8 for g in range(sh)]
    assert xs[-1] < scene["w"] and ys[-1] < scene["h"] and len(set(xs)) == sw
    ref = plain[np.ix_(ys, xs)]
    assert ref.shape == sparse.shape == (32, 32, 4)
    d = np.abs(sparse - ref).max(axis=2)
    print(f"sparse vs plain: max_abs={d.max():.3g} exact={np.mean(d == 0):.4f} "
          f"coverage={np.mean(ref[..., 3] > 0):.3f}", flush=True)
    assert np.mean(d > 1e-5) < 0.01 and d.max() < 0.06
    assert np.mean(ref[..., 3] > 0) > 0.2
