"""Training of the stepping network on plain-marcher traces.

For each of `n_tfs` seeded random transfer functions (4-8 keys) and each of
`n_views` orbit views, the plain marcher traces the volume at a reduced image
size (include/hipdvr_learned.h: hipdvr_learned_trace).  For every 16^3
brick and ray octant the trace records the worst transmittance-weighted error
of marching that brick segment at 2x, 4x and 8x the step, and of skipping it.
These errors become the training targets, with inputs from the features of the
brick statistics and the TF.  The network is then trained with the library's
Adam kernels.

    from voreen_amd import stepper
    mlp, info = stepper.train(ctx, volume, dims)
"""
import ctypes as C
import time

import numpy as np

from . import hipdvr as H
from . import workloads as W


class _View:
    """A device address inside a larger buffer (minibatch window)."""

    def __init__(self, addr):
        self.ptr = C.c_void_p(addr)


def _diagnose(mlp, dx, dt, count, n=200_000):
    n = min(n, count)
    x = dx.download((n, H.MLP_INPUTS), np.float32)
    t = dt.download((n, H.MLP_TARGETS), np.float32)
    y = mlp.forward(x)
    skip_l = t[:, 0] > 0.5
    skip_p = y[:, 0] > 0.0
    m_l = t[:, 1].astype(np.int64)
    m_p = np.clip(np.floor(y[:, 1] + 0.5), 0, 3).astype(np.int64)
    keep = ~skip_l
    return dict(label_skip=round(float(skip_l.mean()), 4),
                label_mult_hist=np.bincount(m_l[keep], minlength=4).tolist(),
                pred_skip=round(float(skip_p.mean()), 4),
                false_skip=round(float((skip_p & ~skip_l).mean()), 4),
                pred_mult_hist=np.bincount(m_p[keep & ~skip_p], minlength=4).tolist(),
                over_mult=round(float((m_p[keep] > m_l[keep]).mean()) if keep.any() else 0.0, 4))


def random_tf_keys(seed):
    r = W.SplitMix64(seed)
    n = 4 + r.randint(5)
    return W.random_keys(r.next(), n)


def train(ctx, vol, dims, n_tfs=200, n_views=16, trace_size=192, tf_width=None, rate=2.0,
          epochs=4, min_steps=2000, batch=8192, lr=2e-3, tau_skip=3e-3, tau_step=8e-3,
          seed=0x5EED, capacity=None, extra_tfs=(), diagnose=False, log=None):
    """Trace, collect and train; returns (mlp, info dict).

    extra_tfs: additional RGBA8 tables traced after the random ones.
    diagnose: add label statistics and the trained network's accuracy on up
    to 200k collected rows to the info dict."""
    t0 = time.perf_counter()
    tf_width = tf_width or (256 if vol.desc.dtype == H.U8 else 4096)
    step = W.step_for(dims, rate)
    learned = H.Learned(ctx, vol)
    rows_per_tf = ((dims[0] + 15) // 16) * ((dims[1] + 15) // 16) * ((dims[2] + 15) // 16) * 8
    capacity = capacity or min((n_tfs + len(extra_tfs)) * rows_per_tf, 24_000_000)
    dx = ctx.alloc(capacity * H.MLP_INPUTS * 4)
    dt = ctx.alloc(capacity * H.MLP_TARGETS * 4)
    llf, urb = W.world_box(dims)
    ee = [(ctx.alloc(trace_size * trace_size * 16), ctx.alloc(trace_size * trace_size * 16))
          for _ in range(n_views)]
    for v, (de, dxx) in enumerate(ee):
        ctx.entry_exit_device(W.orbit_camera(v, n_views, elevation_deg=10.0 + 50.0 * (v % 4) / 3.0),
                              llf, urb, trace_size, trace_size, de, dxx)
    params = H.default_params(step)
    count = 0
    try:
        tables = [H.tf_from_keys(random_tf_keys(seed + 7919 * k), tf_width) for k in range(n_tfs)]
        tables += [np.asarray(x, np.uint8) for x in extra_tfs]
        for table in tables:
            tf = ctx.transfunc(table)
            try:
                learned.prepare(tf, params)
                learned.trace_reset()
                for de, dxx in ee:
                    learned.trace(tf, params, de, dxx, trace_size, trace_size)
                count = learned.collect(tau_skip, tau_step, dx, dt, capacity, count)
            finally:
                tf.close()
            if count >= capacity:
                break
        t_data = time.perf_counter() - t0
        if count == 0:
            raise RuntimeError("the traces visited no bricks")
        mlp = H.Mlp(ctx, seed=seed)
        steps = max(min_steps, int(epochs * count / batch))
        rng = np.random.default_rng(seed)
        first = last = None
        offsets = rng.integers(0, max(1, count - batch), size=steps)
        rows = min(batch, count)
        for i, off in enumerate(offsets):
            sub_x = _View(dx.ptr.value + int(off) * H.MLP_INPUTS * 4)
            sub_t = _View(dt.ptr.value + int(off) * H.MLP_TARGETS * 4)
            want = i == 0 or i == steps - 1
            loss = mlp.train_step_device(sub_x, sub_t, rows, lr, want_loss=want)
            if i == 0:
                first = loss
            if i == steps - 1:
                last = loss
        ctx.synchronize()
        info = dict(rows=int(count), tfs=len(tables), views=n_views, trace_size=trace_size, steps=steps,
                    loss_first=first, loss_last=last, data_s=round(t_data, 2),
                    total_s=round(time.perf_counter() - t0, 2))
        if diagnose:
            info.update(_diagnose(mlp, dx, dt, count))
        if log:
            log(info)
        return mlp, info
    finally:
        for de, dxx in ee:
            de.free()
            dxx.free()
        dx.free()
        dt.free()
        learned.close()
