"""Batched persistent launch (include/hipdvr_batch.h): every job's image equals
its single-frame render bit for bit, for every configuration, through a
captured graph replayed between eager work, and after autotuning."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def jobs(ctx):
    from voreen_amd import hipdvr as H, workloads as W
    specs = W.batch_specs(seed=77, n=7, lo=24, hi=96, image=80)
    specs[3]["image"] = (130, 57)        # ragged image
    out = []
    for s in specs:
        vol = H.synth_blobs(s["vol_seed"], s["dims"], s["dtype"])
        v = ctx.volume(vol)
        t = ctx.transfunc(H.tf_from_keys(s["keys"], 256 if s["dtype"] == H.U8 else 4096))
        w, h = s["image"]
        de, dx, do, dr = (ctx.alloc(w * h * 16) for _ in range(4))
        llf, urb = W.world_box(s["dims"])
        ctx.entry_exit_device(s["camera"], llf, urb, w, h, de, dx)
        p = H.default_params(W.step_for(s["dims"], 2.0), camera_pos=tuple(s["camera"].eye))
        out.append(dict(vol=vol, v=v, t=t, de=de, dx=dx, do=do, dr=dr, w=w, h=h, p=p, spec=s))
    yield out
    for j in out:
        for b in ("de", "dx", "do", "dr"):
            j[b].free()
        j["v"].close()
        j["t"].close()


def singles(ctx, jobs, kernel):
    from voreen_amd import hipdvr as H
    ref = []
    total = 0
    for j in jobs:
        p = H.default_params(j["p"].step, kernel=kernel, camera_pos=tuple(j["p"].camera_pos))
        st = ctx.render_device(j["v"], j["t"], j["de"], j["dx"], j["w"], j["h"], p, j["dr"], stats=True)
        total += st["samples"]
        ref.append(j["dr"].download((j["h"], j["w"], 4), np.float32))
    return ref, total


def outputs(jobs):
    return [j["do"].download((j["h"], j["w"], 4), np.float32) for j in jobs]


def clear(ctx, jobs):
    for j in jobs:
        j["do"].upload(np.full((j["h"], j["w"], 4), -1.0, np.float32))


def test_batch_equals_single_frames_for_every_config(ctx, jobs, oracle):
    from voreen_amd import hipdvr as H
    ref, total = singles(ctx, jobs, H.KERNEL_CELL)
    b = H.Batch(ctx, [(j["v"], j["t"], j["de"], j["dx"], j["do"], j["w"], j["h"], j["p"]) for j in jobs])
    try:
        for tile in (0, 1, 2):
            for bl in (3, 4, 5):
                for k in (2, 4):
                    say(f"batch tile {tile} brick_log2 {bl} batch {k}")
                    b.set_config(tile, bl, k)
                    clear(ctx, jobs)
                    st = b.launch(stats=True)
                    for i, (o, r) in enumerate(zip(outputs(jobs), ref)):
                        assert np.array_equal(o, r), f"job {i} tile {tile} bl {bl} k {k}"
                    assert st["samples"] == total
        # resident workgroups per CU: the same images whatever the residency
        for r in (1, 3, 8):
            say(f"batch resident {r}")
            b.set_config(0, 4, 2, r)
            clear(ctx, jobs)
            st = b.launch(stats=True)
            for i, (o, ref_i) in enumerate(zip(outputs(jobs), ref)):
                assert np.array_equal(o, ref_i), f"job {i} resident {r}"
            assert st["samples"] == total
        for r in (-1, 9):
            with pytest.raises(H.HipDvrError) as e:
                b.set_config(0, 4, 2, r)
            assert e.value.code == H.ERR_INVALID
    finally:
        b.close()
    # one job against the oracle as well
    j = jobs[3]
    en = j["de"].download((j["h"], j["w"], 4), np.float32)
    ex = j["dx"].download((j["h"], j["w"], 4), np.float32)
    tf = oracle.tf_build(j["spec"]["keys"], 256 if j["spec"]["dtype"] == H.U8 else 4096)
    cpu, _ = oracle.render(j["vol"], tf, (0.0, 1.0), en, ex, j["p"].step)
    assert np.abs(ref[3] - cpu).max() <= 1e-3


def test_batch_graph_replays_and_autotune(ctx, jobs):
    from voreen_amd import hipdvr as H
    ref, _ = singles(ctx, jobs, H.KERNEL_CELL)
    b = H.Batch(ctx, [(j["v"], j["t"], j["de"], j["dx"], j["do"], j["w"], j["h"], j["p"]) for j in jobs])
    try:
        say("batch autotune")
        cfg, ms = b.autotune(reps=2)
        print(f"autotuned {cfg} {ms:.3f} ms", flush=True)
        assert cfg == b.config() and ms > 0
        say("batch graph capture and replays")
        b.capture()
        for rep in range(3):
            clear(ctx, jobs)
            ctx.synchronize()
            b.launch()
            ctx.synchronize()
            # eager work between replays
            singles(ctx, jobs[:2], H.KERNEL_REFERENCE)
            for i, (o, r) in enumerate(zip(outputs(jobs), ref)):
                assert np.array_equal(o, r), f"replay {rep} job {i}"
    finally:
        b.close()


def test_batch_rejects_unsupported_jobs(ctx, jobs):
    from voreen_amd import hipdvr as H
    j = jobs[0]
    p = H.default_params(j["p"].step, filter_=H.FILTER_NEAREST)
    say("batch invalid jobs")
    with pytest.raises(H.HipDvrError) as e:
        H.Batch(ctx, [(j["v"], j["t"], j["de"], j["dx"], j["do"], j["w"], j["h"], p)])
    assert e.value.code == H.ERR_UNSUPPORTED
    with pytest.raises(H.HipDvrError) as e:
        H.Batch(ctx, [])
    assert e.value.code == H.ERR_INVALID
