"""The stepping network: FP64 oracle pinned by finite differences (CPU), and
the GPU kernels (BF16 weights, FP32 accumulation on MFMA) checked against it:
forward outputs, the full parameter gradient, and training on a learnable
synthetic task."""
import numpy as np
import pytest

SEG = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248),
       "W3": (5248, 5376), "b3": (5376, 5378)}


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(5378, np.float32)
    p[0:1024] = rng.uniform(-1, 1, 1024) * np.sqrt(6 / 80) * scale
    p[1024:1088] = rng.normal(0, 0.1, 64)
    p[1088:5184] = rng.uniform(-1, 1, 4096) * np.sqrt(6 / 128) * scale
    p[5184:5248] = rng.normal(0, 0.1, 64)
    p[5248:5376] = rng.uniform(-1, 1, 128) * np.sqrt(6 / 66) * scale
    p[5376:5378] = rng.normal(0, 0.1, 2)
    return p


def _batch(seed, rows):
    rng = np.random.default_rng(seed)
    x = np.zeros((rows, 16), np.float32)
    x[:, :11] = rng.uniform(-1, 1, (rows, 11))
    t = np.zeros((rows, 3), np.float32)
    t[:, 0] = (x[:, 0] + 0.5 * x[:, 4] > 0).astype(np.float32)
    t[:, 1] = np.clip(1.5 + x[:, 1] - x[:, 5], 0, 3)
    t[:, 2] = rng.uniform(0.2, 1.0, rows)
    return x, t


def test_mlp_oracle_finite_differences(oracle):
    p = _rand_params(1).astype(np.float64)
    x, t = _batch(2, 37)
    loss, g = oracle.mlp_loss_grad(p, x, t, quantize=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(6, b - a), replace=False)
                          for a, b in SEG.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = oracle.mlp_loss_grad(pp, x, t, quantize=False)
        lm, _ = oracle.mlp_loss_grad(pm, x, t, quantize=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_mlp_oracle_closed_form(oracle):
    # all weights zero: y = b3, loss = softplus(b3_0) - t0*b3_0 + t2*(b3_1 - t1)^2
    p = np.zeros(5378, np.float32)
    p[5376:5378] = (0.3, -1.2)
    b0, b1 = float(p[5376]), float(p[5377])
    x, t = _batch(4, 5)
    y = oracle.mlp_forward(p, x)
    assert np.allclose(y, [[b0, b1]] * 5)
    loss, g = oracle.mlp_loss_grad(p.astype(np.float64), x, t)
    t64 = t.astype(np.float64)
    ref = np.mean(np.log1p(np.exp(b0)) - t64[:, 0] * b0 + t64[:, 2] * (b1 - t64[:, 1]) ** 2)
    assert abs(loss - ref) < 1e-12
    assert abs(g[5376] - np.mean(1 / (1 + np.exp(-b0)) - t64[:, 0])) < 1e-12
    assert abs(g[5377] - np.mean(2 * t64[:, 2] * (b1 - t64[:, 1]))) < 1e-12


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(ctx, oracle):
    from voreen_amd import hipdvr as H
    print("[case] mlp forward 1000 rows", flush=True)
    m = H.Mlp(ctx, seed=5)
    try:
        p = _rand_params(6)
        m.set_params(p)
        assert np.array_equal(m.params(), p)
        x, _ = _batch(7, 1000)
        y = m.forward(x)
        ref = oracle.mlp_forward(p, x, quantize=True)
        err = np.abs(y - ref)
        print(f"forward max abs err {err.max():.3g} (output scale {np.abs(ref).max():.3g})", flush=True)
        # activations are rounded to BF16 between layers on the GPU
        assert err.max() < 0.03 * max(1.0, np.abs(ref).max())
        assert np.corrcoef(y[:, 0], ref[:, 0])[0, 1] > 0.999
    finally:
        m.close()


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle(ctx, oracle):
    from voreen_amd import hipdvr as H
    print("[case] mlp gradient 2048 rows", flush=True)
    m = H.Mlp(ctx, seed=8)
    try:
        p = _rand_params(9)
        m.set_params(p)
        x, t = _batch(10, 2048)
        g, loss = m.gradient(x, t)
        lref, gref = oracle.mlp_loss_grad(p.astype(np.float64), x, t, quantize=True)
        assert abs(loss - lref) < 0.01 * abs(lref)
        for name, (a, b) in SEG.items():
            gg, rr = g[a:b].astype(np.float64), gref[a:b]
            rel = np.linalg.norm(gg - rr) / max(np.linalg.norm(rr), 1e-12)
            cos = gg @ rr / max(np.linalg.norm(gg) * np.linalg.norm(rr), 1e-30)
            print(f"grad {name}: rel L2 err {rel:.3g} cos {cos:.6f}", flush=True)
            assert rel < 0.05 and cos > 0.998, name
    finally:
        m.close()


@pytest.mark.gpu
def test_mlp_training_reduces_loss(ctx):
    from voreen_amd import hipdvr as H
    print("[case] mlp training 300 steps", flush=True)
    m = H.Mlp(ctx, seed=11)
    x, t = _batch(12, 8192)
    dx, dt = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
    try:
        dx.upload(x)
        dt.upload(t)
        first = m.train_step_device(dx, dt, 8192, 3e-3, want_loss=True)
        for _ in range(300):
            m.train_step_device(dx, dt, 8192, 3e-3)
        last = m.train_step_device(dx, dt, 8192, 3e-3, want_loss=True)
        print(f"loss {first:.4f} -> {last:.4f}", flush=True)
        assert last < 0.5 * first
        y = m.forward(x)
        acc = np.mean((y[:, 0] > 0) == (t[:, 0] > 0.5))
        assert acc > 0.9
    finally:
        dx.free()
        dt.free()
        m.close()
