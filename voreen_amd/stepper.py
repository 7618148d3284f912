"""Training of the stepping network on plain-marcher traces.

A thin wrapper of hipdvr_learned_train (include/hipdvr_learned.h), which
runs the whole loop in the library: for each of `n_tfs` seeded random
transfer functions (4-8 keys) and each of `n_views` orbit views, the plain
marcher traces the volume at a reduced image size; for every 16^3 brick and
ray octant the trace records the worst transmittance-weighted error of
marching that brick segment at 2x, 4x and 8x the step, and of skipping it.
These errors become the training targets, with inputs from the features of
the brick statistics and the TF, and the network takes Adam steps on them
with the library's kernels.

    from voreen_amd import stepper
    mlp, info = stepper.train(ctx, volume, dims)
"""
from . import hipdvr as H


def train(ctx, vol, dims, n_tfs=200, n_views=16, trace_size=192, rate=2.0, epochs=4, min_steps=2000,
          batch=8192, lr=2e-3, tau_skip=3e-3, tau_step=8e-3, seed=0x5EED, capacity=None, extra_tfs=(),
          log=None):
    """Train a new network for `vol`; returns (mlp, info dict).  dims is the
    volume's (x, y, z) size (checked against the volume); extra_tfs are RGBA8
    tables traced after the random ones; capacity bounds the training rows."""
    if tuple(vol.desc.dims) != tuple(dims):
        raise ValueError(f"dims {tuple(dims)} differ from the volume's {tuple(vol.desc.dims)}")
    learned = H.Learned(ctx, vol)
    mlp = H.Mlp(ctx, seed=seed)
    try:
        info = learned.train(mlp, extra_tfs=extra_tfs, n_tfs=n_tfs, n_views=n_views, trace_size=trace_size,
                             sampling_rate=rate, tau_skip=tau_skip, tau_step=tau_step, min_steps=min_steps,
                             epochs=epochs, batch=batch, lr=lr, seed=seed, max_rows=capacity or 0)
    except Exception:
        mlp.close()
        raise
    finally:
        learned.close()
    if log:
        log(info)
    return mlp, info
