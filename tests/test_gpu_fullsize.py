"""The benchmark workloads at their full sizes against the oracle.

C1: 512^3 uint16 shells, 1024^2.  Unshaded frames of two orbit views equal
    the FP32 oracle over every pixel (bit for bit, equal sample counts);
    with Phong, every 16th pixel in x and y is compared with the FP32 and
    FP64 oracles (default and exact shading).
C2: the 64 uneven volumes at 512^2 in one batched launch (graph replay):
    every job equals its single frame bit for bit, and every 8th pixel of
    every 8th job equals the oracle.
C3: the stepping network trained on 200 TFs x 16 views of plain traces;
    every one of the 36 learned C1 frames stays >= 45 dB against the plain
    frame without falling back, and the trained network reconstructs fewer
    samples and runs faster than the method without it.
C4: 2048x2048x1024 uint16 (8 GiB) with 8 views of 3840x2160 and Phong in one
    batched launch: equal to the 8 single frames bit for bit; every 32nd
    pixel of an unshaded view equals the oracle on the downloaded volume;
    learned views >= 45 dB against the plain ones.
"""
import ctypes as C
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-3


def say(msg):
    print(f"[case] {msg}", flush=True)


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 999.0 if mse == 0 else 10 * np.log10(1 / mse)


def linear_ptr(ctx, v):
    p = C.c_void_p()
    ctx.check(ctx.lib.hipdvr_volume_device_ptr(v.h, C.byref(p)), "hipdvr_volume_device_ptr")
    return p


# ---------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def c1(ctx):
    from voreen_amd import hipdvr as H, workloads as W
    cfg = W.CONFIGS["c1"]
    dims = cfg["dims"]
    w, h = cfg["image"]
    say("c1 volume 512^3 u16")
    vol = W.make_volume("c1")
    keys, tf = W.make_tf("c1")
    llf, urb = W.world_box(dims)
    sp = tuple(float((urb[i] - llf[i]) / dims[i]) for i in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    v = ctx.volume(vol, spacing=sp, offset=off)
    t = ctx.transfunc(tf)
    de, dx, do = (ctx.alloc(w * h * 16) for _ in range(3))
    yield dict(vol=vol, keys=keys, tf=tf, v=v, t=t, de=de, dx=dx, do=do, w=w, h=h, llf=llf, urb=urb, sp=sp,
               off=off, step=W.step_for(dims, cfg["rate"]), nviews=cfg["views"])
    for b in (de, dx, do):
        b.free()
    v.close()
    t.close()


def c1_view(ctx, c, view):
    from voreen_amd import workloads as W
    cam = W.orbit_camera(view, c["nviews"])
    ctx.entry_exit_device(cam, c["llf"], c["urb"], c["w"], c["h"], c["de"], c["dx"])
    en = c["de"].download((c["h"], c["w"], 4), np.float32)
    ex = c["dx"].download((c["h"], c["w"], 4), np.float32)
    return cam, en, ex


@pytest.mark.parametrize("view,kernel", [(0, "auto"), (18, "auto"), (0, "slab"), (18, "cell")])
def test_c1_unshaded_full_frame_equals_oracle(ctx, oracle, c1, view, kernel):
    from voreen_amd import hipdvr as H
    say(f"c1 512^3 u16 1024^2 view {view} unshaded, every pixel, kernel {kernel}")
    cam, en, ex = c1_view(ctx, c1, view)
    k = {"auto": H.KERNEL_AUTO, "slab": H.KERNEL_SLAB, "cell": H.KERNEL_CELL}[kernel]
    p = H.default_params(c1["step"], camera_pos=tuple(cam.eye), kernel=k)
    st = ctx.render_device(c1["v"], c1["t"], c1["de"], c1["dx"], c1["w"], c1["h"], p, c1["do"], stats=True)
    gpu = c1["do"].download((c1["h"], c1["w"], 4), np.float32)
    t0 = time.perf_counter()
    cpu, n = oracle.render(c1["vol"], c1["tf"], (0.0, 1.0), en, ex, c1["step"], spacing=c1["sp"], offset=c1["off"])
    d = np.abs(gpu - cpu)
    print(f"c1 view {view}: max_abs={d.max():.3g} exact={np.mean(np.all(gpu == cpu, axis=2)):.6f} "
          f"samples {st['samples']} (oracle {n}, {time.perf_counter() - t0:.1f} s) "
          f"coverage={np.mean(cpu[..., 3] > 0):.3f}", flush=True)
    assert d.max() <= TOL
    assert np.array_equal(gpu, cpu)
    assert st["samples"] == n
    assert np.mean(cpu[..., 3] > 0) > 0.15


@pytest.mark.parametrize("exact,kernel", [(0, "auto"), (1, "auto"), (0, "slab"), (1, "slab")])
def test_c1_phong_every_16th_pixel(ctx, oracle, c1, exact, kernel):
    from voreen_amd import hipdvr as H
    say(f"c1 512^3 u16 1024^2 view 4 phong exact_shading={exact}, every 16th pixel, kernel {kernel}")
    cam, en, ex = c1_view(ctx, c1, 4)
    k = {"auto": H.KERNEL_AUTO, "slab": H.KERNEL_SLAB}[kernel]
    p = H.default_params(c1["step"], shading=H.SHADE_PHONG, exact_shading=exact, camera_pos=tuple(cam.eye),
                         kernel=k)
    st = ctx.render_device(c1["v"], c1["t"], c1["de"], c1["dx"], c1["w"], c1["h"], p, c1["do"], stats=True)
    gpu = c1["do"].download((c1["h"], c1["w"], 4), np.float32)[::16, ::16]
    sub_e = np.ascontiguousarray(en[::16, ::16])
    sub_x = np.ascontiguousarray(ex[::16, ::16])
    kw = dict(spacing=c1["sp"], offset=c1["off"], shading=1, light=dict(camera=tuple(cam.eye)))
    cpu, n = oracle.render(c1["vol"], c1["tf"], (0.0, 1.0), sub_e, sub_x, c1["step"], **kw)
    f64, _ = oracle.render(c1["vol"], c1["tf"], (0.0, 1.0), sub_e, sub_x, c1["step"], fp64=True, **kw)
    d = np.abs(gpu - cpu).max(axis=2)
    print(f"c1 phong exact={exact}: max_abs={d.max():.3g} above_tol={np.mean(d > TOL):.5f} "
          f"psnr_fp32={psnr(gpu, cpu):.1f} psnr_fp64={psnr(gpu, f64):.1f} shaded={st['shaded']}", flush=True)
    assert np.mean(d > TOL) < 0.002
    assert psnr(gpu, f64) > 45.0
    assert st["shaded"] > 0 and np.mean(cpu[..., 3] > 0) > 0.15


# ---------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------

def test_c2_64_volumes_batched_equals_single_frames_and_oracle(ctx, oracle):
    from voreen_amd import hipdvr as H, workloads as W
    say("c2 64 uneven volumes 512^2: batched vs single frames vs oracle")
    specs = W.batch_specs(seed=0xC2, n=64)
    jobs, keep = [], []
    try:
        for s in specs:
            vol = H.synth_blobs(s["vol_seed"], s["dims"], s["dtype"])
            width = 256 if s["dtype"] == H.U8 else 4096
            tfa = H.tf_from_keys(s["keys"], width)
            v = ctx.volume(vol)
            t = ctx.transfunc(tfa)
            w, h = s["image"]
            de, dx, do, dr = (ctx.alloc(w * h * 16) for _ in range(4))
            keep += [v, t, de, dx, do, dr]
            llf, urb = W.world_box(s["dims"])
            ctx.entry_exit_device(s["camera"], llf, urb, w, h, de, dx)
            p = H.default_params(W.step_for(s["dims"], 2.0), camera_pos=tuple(s["camera"].eye))
            jobs.append(dict(vol=vol, tfa=tfa, v=v, t=t, de=de, dx=dx, do=do, dr=dr, w=w, h=h, p=p))
        b = H.Batch(ctx, [(j["v"], j["t"], j["de"], j["dx"], j["do"], j["w"], j["h"], j["p"]) for j in jobs])
        keep.append(b)
        cfg, ms = b.autotune(reps=1)
        print(f"c2 autotuned {cfg} {ms:.3f} ms", flush=True)
        b.capture()
        for j in jobs:
            j["do"].upload(np.full((j["h"], j["w"], 4), -1.0, np.float32))
        ctx.synchronize()
        b.launch()
        ctx.synchronize()
        total = 0
        for i, j in enumerate(jobs):
            p = H.default_params(j["p"].step, kernel=H.KERNEL_CELL, camera_pos=tuple(j["p"].camera_pos))
            st = ctx.render_device(j["v"], j["t"], j["de"], j["dx"], j["w"], j["h"], p, j["dr"], stats=True)
            total += st["samples"]
            single = j["dr"].download((j["h"], j["w"], 4), np.float32)
            batched = j["do"].download((j["h"], j["w"], 4), np.float32)
            assert np.array_equal(batched, single), f"job {i}"
            if i % 8 == 0:
                en = j["de"].download((j["h"], j["w"], 4), np.float32)[::8, ::8]
                ex = j["dx"].download((j["h"], j["w"], 4), np.float32)[::8, ::8]
                cpu, _ = oracle.render(j["vol"], j["tfa"], (0.0, 1.0), np.ascontiguousarray(en),
                                       np.ascontiguousarray(ex), j["p"].step)
                assert np.array_equal(batched[::8, ::8], cpu), f"job {i} vs oracle"
        st = b.launch(stats=True)
        assert st["samples"] == total
    finally:
        for r in reversed(keep):
            (r.free if hasattr(r, "free") else r.close)()


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------

def test_c3_learned_36_views_at_45_db_and_gain(ctx, c1):
    from voreen_amd import hipdvr as H, stepper, workloads as W
    say("c3 learned stepping on c1: training on 200 TFs x 16 views")
    mlp, info = stepper.train(ctx, c1["v"], (512, 512, 512), n_tfs=200, n_views=16)
    print(f"c3 trained: {info}", flush=True)
    lr = H.Learned(ctx, c1["v"])
    dl = ctx.alloc(c1["w"] * c1["h"] * 16)
    BATCHED_VIEWS = (0, 6, 12, 18, 24, 30)
    single = {}
    try:
        psnrs, learned_samples, base_samples, plain_samples = [], 0, 0, 0
        t_learned = t_base = t_plain = 0.0
        for view in range(c1["nviews"]):
            cam = W.orbit_camera(view, c1["nviews"])
            ctx.entry_exit_device(cam, c1["llf"], c1["urb"], c1["w"], c1["h"], c1["de"], c1["dx"])
            p = H.default_params(c1["step"], shading=H.SHADE_PHONG, camera_pos=tuple(cam.eye))
            if view == 0:
                lr.prepare(c1["t"], p, mlp)
            ctx.synchronize()
            t0 = time.perf_counter()
            stp = ctx.render_device(c1["v"], c1["t"], c1["de"], c1["dx"], c1["w"], c1["h"], p, c1["do"], stats=True)
            t_plain += time.perf_counter() - t0
            t0 = time.perf_counter()
            ls = lr.render_device(c1["t"], c1["de"], c1["dx"], c1["w"], c1["h"], p, dl, stats=True)
            t_learned += time.perf_counter() - t0
            plain = c1["do"].download((c1["h"], c1["w"], 4), np.float32)
            learned = dl.download((c1["h"], c1["w"], 4), np.float32)
            psnrs.append(psnr(learned, plain))
            assert ls["fell_back"] == 0, f"view {view} fell back (sparse {ls['sparse_psnr']:.1f} dB)"
            learned_samples += ls["samples"]
            plain_samples += stp["fetched"]
            if view in BATCHED_VIEWS:
                single[view] = learned
        # six of the views in one batched launch with learned stepping: the
        # same images as their single-view learned frames
        say("c3 batched learned launch of six views")
        bufs = {v_: [ctx.alloc(c1["w"] * c1["h"] * 16) for _ in range(3)] for v_ in BATCHED_VIEWS}
        b = None
        try:
            jobs = []
            for v_ in BATCHED_VIEWS:
                cam = W.orbit_camera(v_, c1["nviews"])
                de, dx, do = bufs[v_]
                ctx.entry_exit_device(cam, c1["llf"], c1["urb"], c1["w"], c1["h"], de, dx)
                jobs.append((c1["v"], c1["t"], de, dx, do, c1["w"], c1["h"],
                             H.default_params(c1["step"], shading=H.SHADE_PHONG, camera_pos=tuple(cam.eye))))
            b = H.Batch(ctx, jobs)
            b.set_learned([lr] * len(jobs))
            bst = b.launch_learned(stats=True)
            for v_, st_ in zip(BATCHED_VIEWS, bst):
                img = bufs[v_][2].download((c1["h"], c1["w"], 4), np.float32)
                assert st_["fell_back"] == 0 and np.array_equal(img, single[v_]), f"batched view {v_}"
        finally:
            if b is not None:
                b.close()
            for bb in bufs.values():
                for x in bb:
                    x.free()
        # the method without the network: occupancy skipping only, base step
        # (a table that cannot gain: the plain frame itself)
        lr.prepare(c1["t"], p, None)
        for view in range(c1["nviews"]):
            cam = W.orbit_camera(view, c1["nviews"])
            ctx.entry_exit_device(cam, c1["llf"], c1["urb"], c1["w"], c1["h"], c1["de"], c1["dx"])
            p = H.default_params(c1["step"], shading=H.SHADE_PHONG, camera_pos=tuple(cam.eye))
            ctx.synchronize()
            t0 = time.perf_counter()
            bs = lr.render_device(c1["t"], c1["de"], c1["dx"], c1["w"], c1["h"], p, dl, stats=True)
            t_base += time.perf_counter() - t0
            base_samples += bs["samples"]
        print(f"c3: psnr min {min(psnrs):.2f} mean {np.mean(psnrs):.2f} dB; reconstructed learned "
              f"{learned_samples / 36:.0f} / no network {base_samples / 36:.0f} / plain {plain_samples / 36:.0f} "
              f"per frame; time learned {t_learned:.3f} s, no network {t_base:.3f} s, plain {t_plain:.3f} s "
              "(36 counted frames each)", flush=True)
        assert min(psnrs) >= 45.0
        assert learned_samples < 0.5 * base_samples
        assert t_learned < t_base
    finally:
        dl.free()
        lr.close()
        mlp.close()


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------

def test_c4_8gib_batched_views_single_frames_oracle_and_learned(ctx, oracle):
    from voreen_amd import hipdvr as H, stepper, workloads as W
    cfg = W.CONFIGS["c4"]
    dims = cfg["dims"]
    wimg, himg = cfg["image"]
    nviews = cfg["views"]
    say("c4 2048x2048x1024 u16 on the device, 8 views of 3840x2160, phong")
    buf, empty = W.make_volume_device(ctx, "c4")
    llf, urb = W.world_box(dims)
    sp = tuple(float((urb[i] - llf[i]) / dims[i]) for i in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    v = ctx.volume_device(buf, dims, H.U16, spacing=sp, offset=off)
    buf.free()
    keys, tfa = W.make_tf("c4")
    t = ctx.transfunc(tfa)
    step = W.step_for(dims, cfg["rate"])
    img = wimg * himg * 16
    cams = [W.orbit_camera(i, nviews) for i in range(nviews)]
    bufs = [tuple(ctx.alloc(img) for _ in range(4)) for _ in cams]
    res = [v, t] + [x for bset in bufs for x in bset]
    try:
        for cam, (de, dx, _, _) in zip(cams, bufs):
            ctx.entry_exit_device(cam, llf, urb, wimg, himg, de, dx)
        params = [H.default_params(step, shading=H.SHADE_PHONG, camera_pos=tuple(c.eye)) for c in cams]
        singles = []
        for (de, dx, do, _), p in zip(bufs, params):
            ctx.render_device(v, t, de, dx, wimg, himg, p, do)
            singles.append(do.download((himg, wimg, 4), np.float32))
        say("c4 batched launch of the 8 views")
        b = H.Batch(ctx, [(v, t, de, dx, ob, wimg, himg, p) for (de, dx, _, ob), p in zip(bufs, params)])
        res.append(b)
        b.launch()
        ctx.synchronize()
        for i, ((_, _, _, ob), ref) in enumerate(zip(bufs, singles)):
            assert np.array_equal(ob.download((himg, wimg, 4), np.float32), ref), f"view {i}"

        say("c4 view 0 unshaded: every 32nd pixel against the oracle on the downloaded volume")
        de, dx, do, _ = bufs[0]
        ctx.render_device(v, t, de, dx, wimg, himg, H.default_params(step, camera_pos=tuple(cams[0].eye)), do)
        gpu = do.download((himg, wimg, 4), np.float32)[::32, ::32]
        en = np.ascontiguousarray(de.download((himg, wimg, 4), np.float32)[::32, ::32])
        ex = np.ascontiguousarray(dx.download((himg, wimg, 4), np.float32)[::32, ::32])
        vol = np.empty((dims[2], dims[1], dims[0]), np.uint16)
        ctx.check(ctx.lib.hipdvr_memcpy_d2h(ctx.h, vol.ctypes.data_as(C.c_void_p), linear_ptr(ctx, v), vol.nbytes),
                  "hipdvr_memcpy_d2h")
        assert abs(float(np.mean(vol[::7, ::7, ::7] == 0)) - empty) < 0.01
        cpu, n = oracle.render(vol, tfa, (0.0, 1.0), en, ex, step, spacing=sp, offset=off)
        del vol
        print(f"c4 oracle subset: {int(np.sum(cpu[..., 3] > 0))} rays, max_abs {np.abs(gpu - cpu).max():.3g}",
              flush=True)
        assert np.array_equal(gpu, cpu)
        assert np.mean(cpu[..., 3] > 0) > 0.1

        say("c4 learned views (network trained on a 512^3 volume of the same generator)")
        tdims = (512, 512, 512)
        tb, _ = W.make_volume_device(ctx, "c4", dims=tdims)
        tl, tu = W.world_box(tdims)
        tsp = tuple(float((tu[i] - tl[i]) / tdims[i]) for i in range(3))
        tv = ctx.volume_device(tb, tdims, H.U16, spacing=tsp,
                               offset=tuple(float(tl[i] + 0.5 * tsp[i]) for i in range(3)))
        tb.free()
        mlp, info = stepper.train(ctx, tv, tdims, n_tfs=200, n_views=16)
        tv.close()
        res.append(mlp)
        lr = H.Learned(ctx, v)
        res.append(lr)
        lr.prepare(t, params[0], mlp)
        for i, ((de, dx, _, ob), p) in enumerate(zip(bufs, params)):
            ls = lr.render_device(t, de, dx, wimg, himg, p, ob, stats=True)
            q = psnr(ob.download((himg, wimg, 4), np.float32), singles[i])
            print(f"c4 learned view {i}: {q:.2f} dB, reconstructed {ls['samples']}, fell back {ls['fell_back']}",
                  flush=True)
            assert q >= 45.0
    finally:
        for r in reversed(res):
            (r.free if hasattr(r, "free") else r.close)()
