#!/usr/bin/env python3
"""Benchmark of the DVR ray march on one MI355X.

Workload (default `--config c1`): 512^3 uint16 nested-shell volume, 6-key TF
(4096 entries), Phong with central-difference gradients, 1024^2 image, step
1/1024 (samplingRate 2), 36 orbit views.  One bench step = one orbit (36
frames).  Volume, TF and all entry/exit images are resident in HBM before the
timed region.  The step's views are rendered twice: as 36 per-frame march
launches on the library's stream (what HIPRaycaster::process issues; the
`per_frame` figures) and as one persistent batched launch of all 36 views
(hipdvr_batch, autotuned, replayed from a captured graph), whose images equal
the per-frame ones bit for bit.  `value` is the batched rate when a step has
more than one view (C0's single view: the per-frame launch).

Prints one JSON line: frames/s (value), Gsamples/s, the roofline of the timed
kernel (algorithmic bytes per launch / HIP-event launch time; HBM traffic and
L2 hit rate from rocprofv3 counter passes of the same kernel), the per-frame
launches, the host-buffer path, learned stepping, and the CPU baseline (the
FP32 CPURaycaster restatement, no shading, on the host's cores, all rays of
view 0).

    python bench.py [--steps K] [--warmup W] [--config c1|c0|c2|c4] [--kernel auto|reference|bricked|cell]
                    [--batch 1|2|4|8]

`--config c2` and `--config c4` are the batched workloads (64 uneven
volumes; the 8 GiB stress volume with 8 4K views), each printing its own line.
"""
import argparse
import csv
import ctypes as C
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions per second in a model of one wave64 instruction
# every 4 clocks per 16-lane SIMD (256 CUs x 4 SIMDs x 2.4 GHz).  A reference
# rate, not a ceiling: the chip sustains more on some instruction mixes (the
# unshaded 512^3 march measures 1.13 of it), so roofline.valu_issue_frac is
# the launch's VALU instruction rate relative to this model.
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2", "c4"],
                    help="c2: 64 uneven volumes at 512^2 in one batched persistent launch; "
                         "c4: 8 GiB turbulence volume, 8 views of 3840x2160 in one batched launch")
    ap.add_argument("--jobs", type=int, default=64, help="c2: number of volumes")
    ap.add_argument("--no-autotune", action="store_true", help="c2: keep the default configuration")
    ap.add_argument("--kernel", default="auto", choices=["auto", "reference", "bricked", "cell", "slab"])
    ap.add_argument("--brick-log2", type=int, default=4, choices=[2, 3, 4, 5],
                    help="empty-space brick size of the volume (2^n voxels)")
    ap.add_argument("--batch", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="cell kernel: samples fetched together per ray (0 = library default)")
    ap.add_argument("--views", type=int, default=0, help="override the number of orbit views")
    ap.add_argument("--shading", default="default", choices=["default", "none", "phong"],
                    help="override the workload's shading (changes the workload name)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-stepping leg")
    ap.add_argument("--learned-tfs", type=int, default=200, help="random TFs traced for training")
    ap.add_argument("--learned-views", type=int, default=16, help="orbit views traced per TF")
    ap.add_argument("--learned-tau", default="",
                    help="SKIP,STEP error thresholds of the training labels (default: the library's 3e-3,8e-3)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="also run the plain copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (hipdvr_render) leg")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter passes (HBM traffic, L2 hit rate) of this run")
    ap.add_argument("--no-graph", action="store_true", help="c2, c4: launch the batch eagerly (no captured graph)")
    ap.add_argument("--c2-config", default="",
                    help="c2: tile,brick_log2,batch,resident instead of autotuning (the counter passes' child)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the leg that renders all views of a step in one batched launch")
    return ap.parse_args()


def measured_traffic(config, shading):
    """HBM bytes per launch of the march kernel measured with rocprofv3
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) for this workload, from the
    committed profile summary; None when there is none."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and not line.startswith("#") and f[0] == config and f[1] == shading:
            return {"bytes": int(f[2]), "source": f[3]}
    return None


MARCH_KERNELS = {"auto": "march_cell_kernel", "cell": "march_cell_kernel", "slab": "march_slab_kernel",
                 "bricked": "march_fast_kernel", "reference": "march_ref_kernel"}


def _read_counters(outdir):
    """(kernel name, dispatch id, counter, value) rows of a rocprofv3 counter
    collection CSV under outdir."""
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection*.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                try:
                    rows.append((r["Kernel_Name"], int(r.get("Dispatch_Id") or 0), r["Counter_Name"],
                                 float(r["Counter_Value"])))
                except (KeyError, ValueError):
                    continue
    return rows


def pmc_leg(args, views=4, c4=False, c2=None):
    """This run's HBM traffic, L2 hit rate and VALU instruction count of the
    march kernel: rocprofv3 counter passes (FETCH_SIZE; WRITE_SIZE; TCC_HIT +
    TCC_MISS; SQ_INSTS_VALU) over a child run
    of this bench on `views` views of the same workload (c4: one eager launch
    of the batched kernel over all views; c2: one eager launch of the batch
    of volumes in the configuration `c2` the parent tuned), each byte counter
    scaled by what it reads for a plain 1 GiB copy in the same child.  Runs
    before this process touches the GPU (c2: after it has released it); None
    when rocprofv3 is missing or a pass fails."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    common = ["--steps", "1", "--warmup", "0", "--no-cpu", "--no-learned", "--no-pmc", "--copy-calibration", "1024"]
    from voreen_amd import workloads as W   # host-side tables only; no GPU
    nviews = args.views or W.CONFIGS.get(args.config, {}).get("views", 1)
    if c2 is not None:
        # the batch of 64 volumes in the configuration the parent tuned,
        # launched eagerly once (the child's last batch_kernel dispatch)
        kname, views, limit = "batch_kernel<", 1, 300
        child = [sys.executable, os.path.abspath(__file__), "--config", "c2", "--jobs", str(args.jobs),
                 "--no-graph", "--c2-config", ",".join(str(x) for x in c2), *common]
    elif c4:
        kname, views, limit = "batch_kernel<", 1, 420
        child = [sys.executable, os.path.abspath(__file__), "--config", "c4", "--brick-log2", str(args.brick_log2),
                 "--no-graph", *common]
    elif batched_mode(args, nviews):
        # the timed launch: all views in one batched launch, launched eagerly
        # after the autotuner's launches (the last batch_kernel dispatch)
        kname, views, limit = "batch_kernel<", 1, 300
        child = [sys.executable, os.path.abspath(__file__), "--config", args.config, "--shading", args.shading,
                 "--brick-log2", str(args.brick_log2), "--views", str(nviews), "--no-host", "--no-graph", *common]
    else:
        kname, limit = MARCH_KERNELS[args.kernel], 120
        child = [sys.executable, os.path.abspath(__file__), "--config", args.config, "--kernel", args.kernel,
                 "--shading", args.shading, "--brick-log2", str(args.brick_log2), "--batch", str(args.batch),
                 "--views", str(views), "--no-host", "--no-batched", *common]
    passes = {"fetch": ["FETCH_SIZE"], "write": ["WRITE_SIZE"], "tcc": ["TCC_HIT_sum", "TCC_MISS_sum"],
              "valu": ["SQ_INSTS_VALU"]}
    got = {}
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    for name, counters in passes.items():
        outdir = tempfile.mkdtemp(prefix=f"hipdvr_pmc_{name}_")
        cmd = ["timeout", "-s", "KILL", str(limit), exe, "--pmc", *counters, "--output-format", "csv", "-d", outdir,
               "-o", "run", "--", *child]
        print(f"[pmc] pass {' '.join(counters)} ...", file=sys.stderr, flush=True)
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, timeout=limit + 30)
        except (subprocess.TimeoutExpired, OSError):
            shutil.rmtree(outdir, ignore_errors=True)
            return None
        rows = _read_counters(outdir)
        shutil.rmtree(outdir, ignore_errors=True)
        if not rows:
            return None
        got[name] = rows
        print(f"[pmc] {' '.join(counters)}: {len(rows)} counter rows", file=sys.stderr, flush=True)

    def per_dispatch(rows, counter, match):
        d = {}
        for k, disp, c, v in rows:
            if c.startswith(counter) and match in k:
                d[disp] = d.get(disp, 0.0) + v
        return [d[i] for i in sorted(d)]

    # the timed launches are the last `views` dispatches of the march kernel
    fetch = per_dispatch(got["fetch"], "FETCH_SIZE", kname)[-views:]
    write = per_dispatch(got["write"], "WRITE_SIZE", kname)[-views:]
    hit = per_dispatch(got["tcc"], "TCC_HIT", kname)[-views:]
    miss = per_dispatch(got["tcc"], "TCC_MISS", kname)[-views:]
    valu = per_dispatch(got["valu"], "SQ_INSTS_VALU", kname)[-views:]
    cf = per_dispatch(got["fetch"], "FETCH_SIZE", "copy16_kernel")
    cw = per_dispatch(got["write"], "WRITE_SIZE", "copy16_kernel")
    if not (fetch and write and hit and miss and cf and cw):
        return None
    gib_kib = 1024.0 * 1024.0
    f_fac, w_fac = gib_kib / cf[-1], gib_kib / cw[-1]
    traffic = (np.mean(fetch) * f_fac + np.mean(write) * w_fac) * 1024.0
    what = (f"one eager batched launch of {args.jobs} volumes" if c2 is not None else
            "one eager batched launch of all views" if kname.startswith("batch") else
            f"views 0..{views - 1} of the orbit")
    return {"traffic_bytes": int(traffic), "l2_hit_rate": round(float(np.sum(hit) / (np.sum(hit) + np.sum(miss))), 4),
            "valu_insts": float(np.mean(valu)) if valu else None,
            "fetch_factor": round(f_fac, 3), "write_factor": round(w_fac, 3),
            "source": f"rocprofv3 --pmc in this run: {what}, FETCH_SIZE/WRITE_SIZE "
                      "(KiB per dispatch) scaled by a 1 GiB copy kernel, TCC_HIT/TCC_MISS summed over channels"}


def cpu_model():
    """The host CPU as /proc/cpuinfo names it: the model-name string of the
    first processor, with its family/model/stepping numbers."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if info:
                    break
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    return f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, " \
           f"stepping {info.get('stepping', '?')})"


def host_cores():
    """(threads for the baseline, logical CPUs this process may run on, CPUs
    online).  The baseline uses every CPU the process may run on, capped by
    OMP_NUM_THREADS when the environment sets it (a shared GPU box sets it to
    the CPU share of one GPU)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    online = os.cpu_count() or affinity
    omp = int(os.environ.get("OMP_NUM_THREADS") or 0)
    use = min(affinity, omp) if omp > 0 else affinity
    return use, affinity, online


def cpu_baseline(vol, keys, tf, en, ex, step, samples_per_frame, target_s):
    """Time the oracle's CPURaycaster restatement (no shading) on a band of
    rows at the image centre (the whole frame when it fits the time budget):
    one warm-up, then the median of 3..9 timed runs within about target_s;
    frames/s = measured samples/s over the plain samples of a frame."""
    from oracle import oracle as O
    cores, affinity, online = host_cores()
    h = en.shape[0]
    rows = 8
    while True:
        y0 = max(0, h // 2 - rows // 2)
        band_e = np.ascontiguousarray(en[y0:y0 + rows])
        band_x = np.ascontiguousarray(ex[y0:y0 + rows])
        t0 = time.perf_counter()
        _, n = O.render(vol, tf, (0.0, 1.0), band_e, band_x, step, nthreads=cores)
        dt = time.perf_counter() - t0
        if dt * 8 >= target_s or rows >= h:
            break
        rows = min(h, max(rows * 2, int(rows * target_s / 8 / max(dt, 1e-3))))
    runs = max(3, min(9, int(target_s / max(dt, 1e-3))))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, n = O.render(vol, tf, (0.0, 1.0), band_e, band_x, step, nthreads=cores)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    sps = n / dt
    return {
        "value": round(sps / samples_per_frame, 5),
        "unit": "frames/s",
        "cores": cores,
        "cpus_affinity": affinity,
        "cpus_online": online,
        "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
        "cpu_model": cpu_model(),
        "kind": "port",
        "gsamples_per_s": round(sps / 1e9, 5),
        "sample": (f"rows {y0}..{y0 + rows - 1} of view 0 ({rows}x{en.shape[1]} rays, {n} samples), no "
                   f"shading; median of {runs} runs after one warm-up ({dt:.2f} s each); frames/s = "
                   "samples/s / plain samples per frame"),
    }


def tau_args(args):
    """stepper.train keyword arguments for --learned-tau (none: its defaults)."""
    if not args.learned_tau:
        return {}
    ts, tt = (float(x) for x in args.learned_tau.split(","))
    return {"tau_skip": ts, "tau_step": tt}


def learned_leg(ctx, v, t, dims, bufs, params, wimg, himg, plain_fps, plain_samples, args, batched=None):
    """Learned stepping on the same frames: train the stepping network on
    plain traces of seeded random TFs (the library's own kernels), then time
    the learned render of every view (learned march + 1/64 sparse plain
    re-march + judge + conditional fallback) and compare each frame with the
    plain one (PSNR over RGBA).  With `batched` (the plain batched leg's
    result): also every view in one batched launch with learned stepping
    (hipdvr_batch_launch_learned, the same configuration), compared with the
    plain batched launch; the speedup is then the batched one's."""
    from voreen_amd import hipdvr as H, stepper
    t0 = time.perf_counter()
    mlp, info = stepper.train(ctx, v, dims, n_tfs=args.learned_tfs, n_views=args.learned_views, **tau_args(args))
    train_s = time.perf_counter() - t0
    lr = H.Learned(ctx, v)
    nviews = len(bufs)
    dl = [ctx.alloc(wimg * himg * 16) for _ in range(nviews)]
    try:
        lr.prepare(t, params[0], mlp)
        per_view = [lr.render_device(t, de, dx, wimg, himg, p, o, stats=True)
                    for (de, dx, _), p, o in zip(bufs, params, dl)]
        for _ in range(args.warmup):
            for (de, dx, _), p, o in zip(bufs, params, dl):
                lr.render_device(t, de, dx, wimg, himg, p, o)
        ctx.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            for (de, dx, _), p, o in zip(bufs, params, dl):
                lr.render_device(t, de, dx, wimg, himg, p, o)
        ctx.synchronize()
        fps = nviews * args.steps / (time.perf_counter() - ta)
        psnrs = []
        for (_, _, do), o in zip(bufs, dl):
            a = do.download((himg, wimg, 4), np.float32).astype(np.float64)
            b = o.download((himg, wimg, 4), np.float32).astype(np.float64)
            mse = float(np.mean((a - b) ** 2))
            psnrs.append(999.0 if mse == 0 else 10 * np.log10(1 / mse))
        bres = None
        if batched:
            single = [o.download((himg, wimg, 4), np.float32) for o in dl]
            b = H.Batch(ctx, [(v, t, de, dx, o, wimg, himg, p) for (de, dx, _), p, o in zip(bufs, params, dl)])
            try:
                c = batched["config"]
                b.set_config(c["tile"], c["brick_log2"], c["batch"], c.get("resident", 0))
                b.set_learned([lr] * nviews)
                bst = b.launch_learned(stats=True)
                for _ in range(max(args.warmup, 1)):
                    b.launch_learned()
                ctx.synchronize()
                ctx.timing_begin(args.steps)
                ta = time.perf_counter()
                for _ in range(args.steps):
                    b.launch_learned()
                ctx.synchronize()
                bfps = nviews * args.steps / (time.perf_counter() - ta)
                bms = ctx.timing_read()
                bpsnr = []
                equal = True
                for (_, _, do), o, sref in zip(bufs, dl, single):
                    x = o.download((himg, wimg, 4), np.float32)
                    equal = equal and np.array_equal(x, sref)
                    mse = float(np.mean((do.download((himg, wimg, 4), np.float32).astype(np.float64) - x) ** 2))
                    bpsnr.append(999.0 if mse == 0 else 10 * np.log10(1 / mse))
            finally:
                b.close()
            bres = {"fps": round(bfps, 3), "speedup_vs_batched_plain": round(bfps / batched["fps"], 3),
                    "ms_per_launch": round(sum(bms) / len(bms), 4), "psnr_min_db": round(min(bpsnr), 2),
                    "views_fell_back": int(sum(x["fell_back"] for x in bst)),
                    "equals_single_view_learned": bool(equal)}
    finally:
        for o in dl:
            o.free()
        lr.close()
        mlp.close()
    res = {
        "fps": round(fps, 3),
        "speedup_vs_plain": round(fps / plain_fps, 3),
        "samples_equiv_gsamples_per_s": round(plain_samples * fps / 1e9, 3),
        "psnr_min_db": round(min(psnrs), 2),
        "psnr_mean_db": round(float(np.mean(psnrs)), 2),
        "views_fell_back": int(sum(s["fell_back"] for s in per_view)),
        "reconstructed_per_frame": int(sum(s["samples"] for s in per_view) / nviews),
        "train": {"tfs": info["tfs"], "views": info["views"], "rows": info["rows"], "steps": info["steps"],
                  "loss_first": round(info["loss_first"], 4), "loss_last": round(info["loss_last"], 4),
                  "seconds": round(train_s, 2), "tau_skip_step": args.learned_tau or "3e-3,8e-3"},
    }
    if bres:
        # the headline comparison: the same kind of launch on both sides
        res["per_frame"] = {"fps": res["fps"], "speedup_vs_per_frame_plain": res["speedup_vs_plain"],
                            "psnr_min_db": res["psnr_min_db"], "views_fell_back": res["views_fell_back"]}
        res["fps"] = bres["fps"]
        res["speedup_vs_plain"] = bres["speedup_vs_batched_plain"]
        res["samples_equiv_gsamples_per_s"] = round(plain_samples * bres["fps"] / 1e9, 3)
        res["psnr_min_db"] = min(res["psnr_min_db"], bres["psnr_min_db"])
        res["batched"] = bres
    return res


def host_path_leg(ctx, v, t, bufs, params, wimg, himg, kernel_ms, reps=20):
    """What HIPRaycaster::process pays per frame: hipdvr_render with the
    entry/exit images and the output in pageable host memory (numpy arrays),
    view 0, the staging buffers kept by the context after the first frame."""
    de, dx, _ = bufs[0]
    en = de.download((himg, wimg, 4), np.float32)
    ex = dx.download((himg, wimg, 4), np.float32)
    out = np.empty((himg, wimg, 4), np.float32)
    ctx.render(v, t, en, ex, params[0], stats=False, out=out)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.render(v, t, en, ex, params[0], stats=False, out=out)
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    return {"fps": round(1e3 / ms, 3), "ms_per_frame": round(ms, 3), "frames": reps,
            "march_kernel_ms": round(kernel_ms, 3), "transfer_and_host_ms": round(ms - kernel_ms, 3),
            "sample": f"view 0, median of {reps} hipdvr_render calls: {2 * wimg * himg * 16 >> 20} MiB up, "
                      f"{wimg * himg * 16 >> 20} MiB down per frame"}


def main_batch(args):
    """64 uneven volumes (log-uniform 64..384 per axis, u8/u16, blob volumes,
    own TF and orbit camera), 512^2 each, one persistent launch replayed from
    a captured hipGraph; one step = one batch = all jobs' frames."""
    from voreen_amd import hipdvr as H, workloads as W
    t_setup = time.perf_counter()
    specs = W.batch_specs(seed=0xC2, n=args.jobs)
    ctx = H.Context(0)
    res = []
    jobs = []
    for s in specs:
        vol = H.synth_blobs(s["vol_seed"], s["dims"], s["dtype"])
        v = ctx.volume(vol)
        t = ctx.transfunc(H.tf_from_keys(s["keys"], 256 if s["dtype"] == H.U8 else 4096))
        w, h = s["image"]
        de, dx, do = (ctx.alloc(w * h * 16) for _ in range(3))
        llf, urb = W.world_box(s["dims"])
        ctx.entry_exit_device(s["camera"], llf, urb, w, h, de, dx)
        p = H.default_params(W.step_for(s["dims"], 2.0), camera_pos=tuple(s["camera"].eye))
        res += [v, t, de, dx, do]
        jobs.append((v, t, de, dx, do, w, h, p))
    b = H.Batch(ctx, jobs)
    if args.c2_config:
        b.set_config(*(int(x) for x in args.c2_config.split(",")))
        tuned, tuned_ms = b.config(), None
    else:
        tuned, tuned_ms = (b.config(), None) if args.no_autotune else b.autotune(reps=2)
    st = b.launch(stats=True)
    if not args.no_graph:
        b.capture()
    setup_s = time.perf_counter() - t_setup
    for _ in range(args.warmup):
        b.launch()
    ctx.synchronize()
    ctx.timing_begin(args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    launch_ms = ctx.timing_read()
    frames = len(jobs) * args.steps
    vox_bytes = sum((1 if s["dtype"] == H.U8 else 2) for s in specs) / len(specs)
    bytes_batch = st["fetched"] * 8 * vox_bytes + 48 * sum(j[5] * j[6] for j in jobs)
    kernel_s = sum(launch_ms) / 1e3
    achieved = bytes_batch * args.steps / kernel_s / 1e9
    out = {
        "metric": "frames/s + Gsamples/s at 1024^2, 512^3 u16; learned vs plain at >=45 dB PSNR",
        "value": round(frames / elapsed, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c2: {len(jobs)} uneven volumes (64..384 per axis, u8/u16), 512x512 each, "
                               "one persistent launch from a captured hipGraph per step",
                   "autotuned": {"tile": ["8x8", "16x4", "32x2"][tuned[0]], "brick": 1 << tuned[1],
                                 "batch": tuned[2], "resident_per_cu": tuned[3], "ms": tuned_ms}},
        "gsamples_per_s": round(st["samples"] * args.steps / elapsed / 1e9, 4),
        "samples_per_batch": st["samples"], "fetched_per_batch": st["fetched"],
        "kernel_ms_avg": round(sum(launch_ms) / len(launch_ms), 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "setup_s": round(setup_s, 2),
    }
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        ca, cb = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_device(cb, ca, nb)
        ctx.synchronize()
        ca.free()
        cb.free()
    b.close()
    for r in res:
        (r.free if hasattr(r, "free") else r.close)()
    ctx.close()
    # counter passes of the tuned configuration: a child per pass, after this
    # process has released the GPU
    pmc = None if (args.no_pmc or args.c2_config) else pmc_leg(args, c2=tuned)
    if pmc:
        kms = sum(launch_ms) / len(launch_ms)
        out["roofline"].update({
            "traffic": pmc["traffic_bytes"], "l2_hit_rate": pmc["l2_hit_rate"],
            "counter_calibration": {"fetch": pmc["fetch_factor"], "write": pmc["write_factor"]},
            "traffic_source": pmc["source"], "algorithmic_bytes_per_launch": int(bytes_batch),
            "reuse": round(bytes_batch / pmc["traffic_bytes"], 3)})
        if pmc.get("valu_insts"):
            out["roofline"]["valu_insts_per_launch"] = int(pmc["valu_insts"])
            out["roofline"]["valu_issue_frac"] = round(pmc["valu_insts"] / (kms / 1e3 * VALU_ISSUE_PEAK), 4)
    print(json.dumps(out), flush=True)


def batched_mode(args, nviews):
    """Whether the bench times a step's views in one batched launch: steps of
    several views with the default kernel choice."""
    return nviews > 1 and not args.no_batched and args.kernel == "auto" and args.batch == 0


def batched_views_leg(ctx, v, t, bufs, params, wimg, himg, args, bytes_orbit):
    """All views of the step in one persistent launch (hipdvr_batch_*,
    autotuned configuration, captured graph), each view's image compared with
    its single-frame image (the per-view launches just timed left them in
    the output buffers).  A reported rate beside `value`, which stays the
    per-frame launches."""
    from voreen_amd import hipdvr as H
    nviews = len(bufs)
    ref = [do.download((himg, wimg, 4), np.float32) for _, _, do in bufs]
    outs = [ctx.alloc(wimg * himg * 16) for _ in range(nviews)]
    b = H.Batch(ctx, [(v, t, de, dx, o, wimg, himg, p) for (de, dx, _), p, o in zip(bufs, params, outs)])
    try:
        cfg, _ = b.autotune(reps=2)
        if not args.no_graph:
            b.capture()
        for _ in range(max(args.warmup, 1)):
            b.launch()
        ctx.synchronize()
        ctx.timing_begin(args.steps)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.launch()
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        ms = ctx.timing_read()
        equal = all(np.array_equal(o.download((himg, wimg, 4), np.float32), r) for o, r in zip(outs, ref))
    finally:
        b.close()
        for o in outs:
            o.free()
    launch_s = sum(ms) / len(ms) / 1e3
    achieved = bytes_orbit / launch_s / 1e9
    return {"fps": round(nviews * args.steps / elapsed, 3), "ms_per_launch": round(launch_s * 1e3, 4),
            "views_per_launch": nviews,
            "config": {"tile": cfg[0], "brick_log2": cfg[1], "batch": cfg[2], "resident": cfg[3]},
            "equals_single_frames": bool(equal),
            "roofline": {"achieved": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 4)}}


def progress(msg):
    print(f"[c4] {msg}", file=sys.stderr, flush=True)


def psnr_db(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 999.0 if mse == 0 else 10 * np.log10(1 / mse)


def main_c4(args):
    """Large-volume stress: 2048x2048x1024 uint16 turbulence (8 GiB, 40% of
    the voxels empty) made on the device, 8 views of 3840x2160 with Phong.
    One step = the 8 views in one batched persistent launch (plain path).
    Beside it: the same views as single frames, learned stepping per view and
    all 8 learned views in one batched launch (hipdvr_batch_launch_learned;
    network trained on plain traces of a 512^3 volume of the same generator:
    the trace kernel addresses the tiled copy with 32-bit offsets, below the
    large volume's 8 GiB) with each view's PSNR against the plain one, and
    every 32nd pixel of an unshaded view 0 against the CPU oracle."""
    # counter passes first: child processes under rocprofv3, before this
    # process touches the GPU
    pmc = None if args.no_pmc else pmc_leg(args, c4=True)
    from voreen_amd import hipdvr as H, workloads as W, stepper
    cfg = W.CONFIGS["c4"]
    dims = cfg["dims"]
    wimg, himg = cfg["image"]
    nviews = args.views or cfg["views"]
    t_setup = time.perf_counter()
    ctx = H.Context(0)
    buf, empty = W.make_volume_device(ctx, "c4")
    llf, urb = W.world_box(dims)
    sp = tuple(float((urb[i] - llf[i]) / dims[i]) for i in range(3))
    off = tuple(float(llf[i] + 0.5 * sp[i]) for i in range(3))
    v = ctx.volume_device(buf, dims, H.U16, spacing=sp, offset=off, brick_log2=args.brick_log2)
    buf.free()
    keys, tf = W.make_tf("c4")
    t = ctx.transfunc(tf)
    step = W.step_for(dims, cfg["rate"])
    cams = [W.orbit_camera(i, nviews) for i in range(nviews)]
    img_bytes = wimg * himg * 16
    bufs = []
    for cam in cams:
        de, dx, do = ctx.alloc(img_bytes), ctx.alloc(img_bytes), ctx.alloc(img_bytes)
        ctx.entry_exit_device(cam, llf, urb, wimg, himg, de, dx)
        bufs.append((de, dx, do))
    params = [H.default_params(step, shading=cfg["shading"], camera_pos=tuple(cam.eye)) for cam in cams]
    progress(f"volume made on the device ({empty:.4f} empty), {nviews} views of {wimg}x{himg}: "
             f"{time.perf_counter() - t_setup:.1f} s")

    # single frames (per-view launches), counted once
    per_view = [ctx.render_device(v, t, de, dx, wimg, himg, p, do, stats=True)
                for (de, dx, do), p in zip(bufs, params)]
    plain_imgs = [do.download((himg, wimg, 4), np.float32) for _, _, do in bufs]
    samples = sum(s["samples"] for s in per_view)
    fetched = sum(s["fetched"] for s in per_view)
    shaded = sum(s["shaded"] for s in per_view)
    ctx.synchronize()
    t0 = time.perf_counter()
    for (de, dx, do), p in zip(bufs, params):
        ctx.render_device(v, t, de, dx, wimg, himg, p, do)
    ctx.synchronize()
    single_s = time.perf_counter() - t0
    progress(f"single frames: {nviews / single_s:.3f} views/s")

    # the batched launch: all views, one persistent grid, captured graph
    outs = [ctx.alloc(img_bytes) for _ in range(nviews)]
    b = H.Batch(ctx, [(v, t, de, dx, o, wimg, himg, p) for (de, dx, _), p, o in zip(bufs, params, outs)])
    b.set_config(H.TILE_8x8, args.brick_log2, 2)
    st = b.launch(stats=True)
    if not args.no_graph:
        b.capture()
    setup_s = time.perf_counter() - t_setup
    for _ in range(args.warmup):
        b.launch()
    ctx.synchronize()
    ctx.timing_begin(args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
        ctx.synchronize()
        progress("batched launch done")
    elapsed = time.perf_counter() - t0
    launch_ms = ctx.timing_read()
    batched_equal = all(np.array_equal(o.download((himg, wimg, 4), np.float32), ref)
                        for o, ref in zip(outs, plain_imgs))
    b.close()
    bytes_step = st["fetched"] * 16 + st["shaded"] * 48 + 48 * wimg * himg * nviews
    kernel_s = sum(launch_ms) / 1e3
    achieved = bytes_step * args.steps / kernel_s / 1e9
    out = {
        "metric": "frames/s + Gsamples/s at 1024^2, 512^3 u16; learned vs plain at >=45 dB PSNR",
        "value": round(nviews * args.steps / elapsed, 4), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"c4: {dims[0]}x{dims[1]}x{dims[2]} u16 turbulence ({empty:.3f} empty, made on "
                                f"the device), {nviews} views of {wimg}x{himg}, phong, step 1/{int(round(1 / step))}, "
                                "all views in one batched launch per step"),
                   "brick": 1 << args.brick_log2, "cell_records_gib": round(dims[0] * dims[1] * dims[2] * 16 / 2**30, 1)},
        "gsamples_per_s": round(st["samples"] * args.steps / elapsed / 1e9, 4),
        "samples_per_step": st["samples"], "fetched_per_step": st["fetched"], "shaded_per_step": st["shaded"],
        "kernel_ms_avg": round(sum(launch_ms) / len(launch_ms), 3),
        "batched_equals_single_frames": bool(batched_equal),
        "single_frames_fps": round(nviews / single_s, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc["traffic_bytes"] if pmc else None},
        "setup_s": round(setup_s, 2),
    }
    if pmc:
        out["roofline"]["l2_hit_rate"] = pmc["l2_hit_rate"]
        out["roofline"]["counter_calibration"] = {"fetch": pmc["fetch_factor"], "write": pmc["write_factor"]}
        out["roofline"]["traffic_source"] = pmc["source"]
        out["roofline"]["algorithmic_bytes_per_launch"] = int(bytes_step)
        out["roofline"]["reuse"] = round(bytes_step / pmc["traffic_bytes"], 3)
        if pmc.get("valu_insts"):
            out["roofline"]["valu_insts_per_launch"] = int(pmc["valu_insts"])
            out["roofline"]["valu_issue_frac"] = round(
                pmc["valu_insts"] / (sum(launch_ms) / len(launch_ms) / 1e3 * VALU_ISSUE_PEAK), 4)
    assert st["samples"] == samples and st["fetched"] == fetched and st["shaded"] == shaded
    for o in outs:
        o.free()
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        ca, cb = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_device(cb, ca, nb)
        ctx.synchronize()
        ca.free()
        cb.free()
        out["copy_calibration_bytes"] = nb

    if not args.no_learned:
        tdims = (512, 512, 512)
        tb, _ = W.make_volume_device(ctx, "c4", dims=tdims)
        tl, tu = W.world_box(tdims)
        tsp = tuple(float((tu[i] - tl[i]) / tdims[i]) for i in range(3))
        toff = tuple(float(tl[i] + 0.5 * tsp[i]) for i in range(3))
        tv = ctx.volume_device(tb, tdims, H.U16, spacing=tsp, offset=toff)
        tb.free()
        t0 = time.perf_counter()
        mlp, info = stepper.train(ctx, tv, tdims, n_tfs=args.learned_tfs, n_views=args.learned_views,
                                  **tau_args(args))
        train_s = time.perf_counter() - t0
        tv.close()
        progress(f"network trained on {tdims}: {train_s:.1f} s")
        lr = H.Learned(ctx, v)
        dl = [ctx.alloc(img_bytes) for _ in range(nviews)]
        bl = None
        try:
            lr.prepare(t, params[0], mlp)
            lst = [lr.render_device(t, de, dx, wimg, himg, p, o, stats=True)
                   for (de, dx, _), p, o in zip(bufs, params, dl)]
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                for (de, dx, _), p, o in zip(bufs, params, dl):
                    lr.render_device(t, de, dx, wimg, himg, p, o)
                ctx.synchronize()
                progress("learned views done")
            lfps = nviews * args.steps / (time.perf_counter() - t0)
            psnrs = [psnr_db(o.download((himg, wimg, 4), np.float32), ref) for o, ref in zip(dl, plain_imgs)]
            learned_imgs = [o.download((himg, wimg, 4), np.float32) for o in dl]
            # the same views in one batched launch with learned stepping
            bl = H.Batch(ctx, [(v, t, de, dx, o, wimg, himg, p) for (de, dx, _), p, o in zip(bufs, params, dl)])
            bl.set_config(H.TILE_8x8, args.brick_log2, 2)
            bl.set_learned([lr] * nviews)
            bst = bl.launch_learned(stats=True)
            for _ in range(args.warmup):
                bl.launch_learned()
            ctx.synchronize()
            ctx.timing_begin(args.steps)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                bl.launch_learned()
                ctx.synchronize()
                progress("batched learned launch done")
            bfps = nviews * args.steps / (time.perf_counter() - t0)
            bms = ctx.timing_read()
            bimgs = [o.download((himg, wimg, 4), np.float32) for o in dl]
            bpsnrs = [psnr_db(x, ref) for x, ref in zip(bimgs, plain_imgs)]
            b_equal = all(np.array_equal(x, y) for x, y in zip(bimgs, learned_imgs))
        finally:
            if bl is not None:
                bl.close()
            for o in dl:
                o.free()
            lr.close()
            mlp.close()
        out["learned"] = {"fps": round(lfps, 4), "speedup_vs_single_frames": round(lfps * single_s / nviews, 3),
                          "psnr_min_db": round(min(psnrs), 2), "psnr_mean_db": round(float(np.mean(psnrs)), 2),
                          "views_fell_back": int(sum(s["fell_back"] for s in lst)),
                          "reconstructed_per_view": int(sum(s["samples"] for s in lst) / nviews),
                          "batched": {"fps": round(bfps, 4), "speedup_vs_batched_plain": round(bfps / out["value"], 3),
                                      "ms_per_launch": round(sum(bms) / len(bms), 3),
                                      "psnr_min_db": round(min(bpsnrs), 2),
                                      "views_fell_back": int(sum(s["fell_back"] for s in bst)),
                                      "equals_single_view_learned": bool(b_equal),
                                      "reconstructed_per_launch": int(bst[0]["samples"])},
                          "train": {"volume": "512^3 turbulence, same generator", "tfs": info["tfs"],
                                    "views": info["views"], "rows": info["rows"], "seconds": round(train_s, 1),
                                    "tau_skip_step": args.learned_tau or "3e-3,8e-3"}}

    if not args.no_cpu:
        # every 32nd pixel of view 0, unshaded, against the oracle on the downloaded volume
        de, dx, do = bufs[0]
        p0 = H.default_params(step, camera_pos=tuple(cams[0].eye))
        ctx.render_device(v, t, de, dx, wimg, himg, p0, do)
        gpu = do.download((himg, wimg, 4), np.float32)[::32, ::32]
        en = de.download((himg, wimg, 4), np.float32)[::32, ::32]
        ex = dx.download((himg, wimg, 4), np.float32)[::32, ::32]
        vol = np.empty((dims[2], dims[1], dims[0]), np.uint16)
        ctx.check(ctx.lib.hipdvr_memcpy_d2h(ctx.h, vol.ctypes.data_as(C.c_void_p), v_linear_ptr(ctx, v),
                                            vol.nbytes), "hipdvr_memcpy_d2h")
        from oracle import oracle as O
        t0 = time.perf_counter()
        cpu, n = O.render(vol, tf, (0.0, 1.0), en, ex, step, spacing=sp, offset=off)
        out["oracle_subset"] = {"rays": int(np.sum(cpu[..., 3] > 0)), "max_abs": float(np.abs(gpu - cpu).max()),
                                "samples": int(n), "cpu_seconds": round(time.perf_counter() - t0, 2)}
        del vol
    for bset in bufs:
        for x in bset:
            x.free()
    v.close()
    t.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def v_linear_ptr(ctx, v):
    """Device pointer of the volume's linear copy (hipdvr_volume_device_ptr)."""
    p = C.c_void_p()
    ctx.check(ctx.lib.hipdvr_volume_device_ptr(v.h, C.byref(p)), "hipdvr_volume_device_ptr")
    return p


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.config == "c2":
        return main_batch(args)
    if args.config == "c4":
        return main_c4(args)
    # counter passes first: child processes under rocprofv3, before this
    # process touches the GPU
    pmc = None if args.no_pmc else pmc_leg(args)
    from voreen_amd import hipdvr as H, workloads as W

    cfg = dict(W.CONFIGS[args.config])
    if args.shading != "default":
        cfg["shading"] = H.SHADE_PHONG if args.shading == "phong" else H.SHADE_NONE
    dims = cfg["dims"]
    wimg, himg = cfg["image"]
    nviews = args.views or cfg["views"]
    kernel = {"auto": H.KERNEL_AUTO, "reference": H.KERNEL_REFERENCE,
              "bricked": H.KERNEL_BRICKED, "cell": H.KERNEL_CELL, "slab": H.KERNEL_SLAB}[args.kernel]
    t_setup = time.perf_counter()
    vol = W.make_volume(args.config)
    keys, tf = W.make_tf(args.config)
    step = W.step_for(dims, cfg["rate"])
    llf, urb = W.world_box(dims)
    spacing = tuple(float((urb[i] - llf[i]) / dims[i]) for i in range(3))
    offset = tuple(float(llf[i] + 0.5 * spacing[i]) for i in range(3))
    vsz = 1 if vol.dtype == np.uint8 else 2

    ctx = H.Context(0)
    v = ctx.volume(vol, spacing=spacing, offset=offset, brick_log2=args.brick_log2)
    t = ctx.transfunc(tf, (0.0, 1.0))
    img_bytes = wimg * himg * 16
    cams = [W.orbit_camera(i, nviews) for i in range(nviews)]
    bufs = []
    for cam in cams:
        de, dx, do = ctx.alloc(img_bytes), ctx.alloc(img_bytes), ctx.alloc(img_bytes)
        ctx.entry_exit_device(cam, llf, urb, wimg, himg, de, dx)
        bufs.append((de, dx, do))
    params = [H.default_params(step, shading=cfg["shading"], kernel=kernel, batch=args.batch,
                               camera_pos=tuple(cam.eye)) for cam in cams]
    ctx.synchronize()

    # counted pass (also the first warm-up): per-view sample statistics
    per_view = []
    for (de, dx, do), p in zip(bufs, params):
        per_view.append(ctx.render_device(v, t, de, dx, wimg, himg, p, do, stats=True))
    samples = sum(s["samples"] for s in per_view)
    fetched = sum(s["fetched"] for s in per_view)
    shaded = sum(s["shaded"] for s in per_view)
    rays = sum(s["rays"] for s in per_view)
    lane_slots = sum(s["lane_slots"] for s in per_view)
    shade_slots = sum(s["shade_slots"] for s in per_view)
    setup_s = time.perf_counter() - t_setup

    for _ in range(args.warmup):
        for (de, dx, do), p in zip(bufs, params):
            ctx.render_device(v, t, de, dx, wimg, himg, p, do)
    ctx.synchronize()

    ctx.timing_begin(nviews * args.steps)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for (de, dx, do), p in zip(bufs, params):
            ctx.render_device(v, t, de, dx, wimg, himg, p, do)
    ctx.synchronize()
    t1 = time.perf_counter()
    launch_ms = ctx.timing_read()
    elapsed = t1 - t0

    frames = nviews * args.steps
    fps = frames / elapsed
    # algorithmic bytes: 8*sizeof(T) per reconstructed sample, 24*sizeof(T)
    # more per shaded sample (32-voxel union of the 7 stencils), 48 B per pixel
    bytes_orbit = fetched * 8 * vsz + shaded * 24 * vsz + 48 * wimg * himg * nviews
    kernel_s = sum(launch_ms) / 1e3
    achieved = bytes_orbit * args.steps / kernel_s / 1e9
    per_frame = {"fps": round(fps, 4), "ms_per_step": round(elapsed / args.steps * 1e3, 3),
                 "kernel_ms_avg": round(sum(launch_ms) / len(launch_ms), 4),
                 "roofline": {"achieved": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 4)}}
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_device(b, a, nb)
        ctx.synchronize()
        a.free()
        b.free()
    # the step's views in one persistent launch: the timed figure when a step
    # has more than one view
    bv = None
    if batched_mode(args, nviews):
        bv = batched_views_leg(ctx, v, t, bufs, params, wimg, himg, args, bytes_orbit)
    if bv:
        value, step_s = bv["fps"], nviews / bv["fps"]
        kernel_ms, views_per_launch = bv["ms_per_launch"], nviews
        achieved = bv["roofline"]["achieved"]
    else:
        value, step_s = fps, elapsed / args.steps
        kernel_ms, views_per_launch = sum(launch_ms) / len(launch_ms), 1
    traffic = None
    if pmc:
        traffic = {"bytes": pmc["traffic_bytes"], "source": pmc["source"]}
    elif not bv and args.kernel == "auto" and args.batch == 0 and args.brick_log2 == 4:
        traffic = measured_traffic(args.config, "phong" if cfg["shading"] else "none")
    launch = (f"all {nviews} views of a step in one persistent batched launch (hipdvr_batch, autotuned, "
              "graph-replayed); per-frame launches under per_frame" if bv else "one launch per view")
    out = {
        "metric": "frames/s + Gsamples/s at 1024^2, 512^3 u16; learned vs plain at >=45 dB PSNR",
        "value": round(value, 4),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: {dims[0]}x{dims[1]}x{dims[2]} "
                         f"{'u8' if vsz == 1 else 'u16'}, {wimg}x{himg}, step 1/{int(round(1 / step))}, "
                         f"{'phong' if cfg['shading'] else 'no shading'}, {nviews} orbit views per step"),
            "launch": launch,
            "kernel": args.kernel,
            "batch": args.batch,
            "brick": 1 << args.brick_log2,
            "voxel": "u8" if vsz == 1 else "u16",
        },
        "gsamples_per_s": round(samples / nviews * value / 1e9, 4),
        "samples_per_frame": samples // nviews,
        "fetched_per_frame": fetched // nviews,
        "shaded_per_frame": shaded // nviews,
        "rays_per_frame": rays // nviews,
        "kernel_ms_avg": round(kernel_ms, 4),
        "views_per_launch": views_per_launch,
        "simd_efficiency": round(samples / max(lane_slots, 1), 4),
        "shade_efficiency": round(shaded / max(shade_slots, 1), 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic["bytes"] if traffic else None,
        },
        "per_frame": per_frame,
        "setup_s": round(setup_s, 2),
    }
    if bv:
        out["batched_views"] = {k: bv[k] for k in ("config", "equals_single_frames")}
    per_launch_bytes = bytes_orbit if bv else bytes_orbit / nviews
    if pmc:
        out["roofline"]["l2_hit_rate"] = pmc["l2_hit_rate"]
        out["roofline"]["counter_calibration"] = {"fetch": pmc["fetch_factor"], "write": pmc["write_factor"]}
        if pmc.get("valu_insts"):
            # the kernel's VALU wave instructions per launch over what the
            # chip can issue in the launch's time: the bound of this kernel
            out["roofline"]["valu_insts_per_launch"] = int(pmc["valu_insts"])
            out["roofline"]["valu_issue_frac"] = round(pmc["valu_insts"] / (kernel_ms / 1e3 * VALU_ISSUE_PEAK), 4)
    if traffic:
        out["roofline"]["traffic_source"] = traffic["source"]
        out["roofline"]["algorithmic_bytes_per_launch"] = round(per_launch_bytes)
        # reuse: algorithmic bytes per HBM byte (records served from L2 / Infinity Cache)
        out["roofline"]["reuse"] = round(per_launch_bytes / traffic["bytes"], 3)
    if args.copy_calibration:
        out["copy_calibration_bytes"] = args.copy_calibration << 20
    if not args.no_host:
        out["host_path"] = host_path_leg(ctx, v, t, bufs, params, wimg, himg, launch_ms[0])
    if not args.no_learned:
        out["learned"] = learned_leg(ctx, v, t, dims, bufs, params, wimg, himg, fps, samples / nviews,
                                     args, batched=bv)
    if not args.no_cpu:
        de, dx, _ = bufs[0]
        en = de.download((himg, wimg, 4), np.float32)
        ex = dx.download((himg, wimg, 4), np.float32)
        out["cpu_baseline"] = cpu_baseline(vol, keys, tf, en, ex, step, samples / nviews,
                                           args.cpu_seconds)
    for bset in bufs:
        for b in bset:
            b.free()
    v.close()
    t.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
