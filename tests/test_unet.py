"""Learned preconditioner: the FP64 CPU network pinned by finite differences,
and the GPU forward/backward (BF16 weights on MFMA, FP32 accumulation)
checked against it."""
import numpy as np
import pytest

import unet_oracle as U

C = (16, 16, 16)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, n, n, 2))
    x[..., 0] = rng.normal(size=(n, n, n))
    m = np.zeros((n, n, n))
    m[0, :, :] = m[-1, :, :] = m[:, 0, :] = m[:, -1, :] = m[:, :, 0] = m[:, :, -1] = 1
    x[..., 1] = m
    t = rng.normal(size=(n, n, n))
    return x, t


def test_unet_oracle_gradient_matches_finite_differences():
    n = 8
    x, t = _inputs(n, 1)
    p = U.init_params(*C, seed=3)
    loss, g = U.loss_grad(p, x, t, *C)
    rng = np.random.default_rng(4)
    idx = rng.choice(p.size, 40, replace=False)
    # always include parameters of the first and last layers
    idx = np.concatenate([idx, [0, 5, p.size - 1, p.size - 20]])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (U.loss_grad(pp, x, t, *C)[0] - U.loss_grad(pm, x, t, *C)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (q, fd, g[q])


def test_unet_oracle_shapes_and_param_count():
    assert U.n_params(16, 32, 64) == sum(co * t * ci + co for ci, co, t in U.layer_specs(16, 32, 64))
    x, _ = _inputs(8, 2)
    y = U.forward(U.init_params(*C, seed=1), x, *C)
    assert y.shape == (8, 8, 8)


def _ctx(pg, n):
    return pg.Context(n, n, n, [0, 0, 0], [1.0, 1.0, 1.0])


@pytest.mark.gpu
def test_gpu_unet_forward_and_gradient_match_oracle():
    import pic_gpu as pg

    n = 16
    ctx = _ctx(pg, n)
    ctx.unet_create(*C, seed=5)
    p = U.init_params(*C, seed=7).astype(np.float32)
    ctx.unet_set_params(p)
    # the device uses BF16 weights: compare against the oracle on BF16-rounded weights
    pb = _bf16(p).astype(np.float64)
    x, t = _inputs(n, 9)
    y = ctx.unet_forward(x.astype(np.float32))
    y_ref = U.forward(pb, x, *C)
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y - y_ref.ravel())) <= 3e-2 * scale
    loss, g = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    loss_ref, g_ref = U.loss_grad(pb, x, t, *C)
    assert loss == pytest.approx(loss_ref, rel=3e-2)
    off = 0
    for ci, co, taps in U.layer_specs(*C):
        nw = co * taps * ci + co
        a, b = g[off:off + nw], g_ref[off:off + nw]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
        assert cos > 0.99, (ci, co, taps, cos)
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b), (ci, co, taps)
        off += nw


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.gpu
def test_gpu_unet_adam_reduces_loss():
    import pic_gpu as pg

    n = 16
    ctx = _ctx(pg, n)
    ctx.unet_create(*C, seed=5)
    x, t = _inputs(n, 11)
    t = np.tanh(x[..., 0])  # a learnable target
    l0, _ = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    for _ in range(60):
        _, g = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
        ctx.unet_adam_step(g, 3e-3)
    l1, _ = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    assert l1 < 0.5 * l0
