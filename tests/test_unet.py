"""Learned preconditioner: the FP64 CPU network pinned by finite differences,
and the GPU forward/backward (BF16 weights on MFMA, FP32 accumulation) and
the whole preconditioner checked against it."""
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


def _fd_check(channels, n, seed):
    x, t = _inputs(n, seed)
    p = U.init_params(channels, seed=3)
    loss, g = U.loss_grad(p, x, t, channels)
    rng = np.random.default_rng(seed + 1)
    idx = rng.choice(p.size, 30, replace=False)
    # always include parameters of the first and last layers
    idx = np.concatenate([idx, [0, 5, p.size - 1, p.size - 20]])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (U.loss_grad(pp, x, t, channels)[0] - U.loss_grad(pm, x, t, channels)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (q, fd, g[q])


def test_unet_oracle_gradient_matches_finite_differences():
    _fd_check(C, 8, 1)


@pytest.mark.parametrize("channels", [(16, 16), (16, 16, 16, 16)])
def test_unet_oracle_gradient_other_depths(channels):
    _fd_check(channels, 8, 2)


def test_unet_oracle_shapes_and_param_count():
    specs = U.layer_specs((16, 32, 64))
    # in, B0, A1, B1, A2, B2, U1, Dec1, U0, Dec0, out
    assert [s[:2] for s in specs] == [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64),
                                      (64, 32), (32, 32), (32, 16), (16, 16), (16, 1)]
    assert U.n_params((16, 32, 64)) == sum(co * t * ci + co for ci, co, t in specs)
    assert len(U.layer_specs((32, 64, 64, 64))) == 2 + 2 * 3 + 2 * 3 + 1
    x, _ = _inputs(8, 2)
    y = U.forward(U.init_params(C, seed=1), x, C)
    assert y.shape == (8, 8, 8)


def test_grid_transfers():
    rng = np.random.default_rng(0)
    n = 8
    regular = np.zeros((n, n, n), bool)
    regular[1:-1, 1:-1, 1:-1] = True
    # the mean over regular nodes reproduces a constant; the mask is the non-regular fraction
    m, mask = U.restrict_mean(np.full((n, n, n), 3.0), regular)
    assert np.allclose(m, 3.0)
    assert mask[0, 0, 0] == pytest.approx(7 / 8) and mask[1, 1, 1] == 0 and mask[0, 1, 1] == 0.5
    f = rng.normal(size=(n, n, n))
    m, _ = U.restrict_mean(f, regular)
    assert m[1, 1, 2] == pytest.approx(f[2:4, 2:4, 4:6].mean())
    assert m[1, 1, 3] == pytest.approx(f[2:4, 2:4, 6].mean())
    assert m[0, 1, 1] == pytest.approx(f[1, 2:4, 2:4].mean())
    # cell-centred trilinear interpolation reproduces linear functions away from the ends
    c = np.arange(n // 2) * 2 + 0.5  # aggregate centres in node units
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    y = 1.0 + 2.0 * X - 3.0 * Y + 0.5 * Z
    fine = U.prolong_trilinear(y)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    exact = 1.0 + 2.0 * i - 3.0 * j + 0.5 * k
    assert np.allclose(fine[1:-1, 1:-1, 1:-1], exact[1:-1, 1:-1, 1:-1])
    assert np.allclose(U.prolong_trilinear(np.full((4, 4, 4), 2.5)), 2.5)


def test_precondition_oracle_is_two_level_cycle():
    """With a zero network the preconditioner is Gauss-Seidel sweeps (one
    red-black, then the closing black-red ones); its output then satisfies
    the red update equations."""
    n = 8
    regular = np.zeros((n, n, n), bool)
    regular[1:-1, 1:-1, 1:-1] = True
    r = np.where(regular, np.random.default_rng(1).normal(size=(n, n, n)), 0.0)
    ch = (16, 16)
    z, corr = U.precondition(r, regular, 1e-4, np.zeros(U.n_params(ch)), ch, 2)
    assert np.all(corr == 0)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    red = regular & ((i + j + k) % 2 == 0)
    nb = np.zeros_like(z)
    nb[1:-1, 1:-1, 1:-1] = (z[:-2, 1:-1, 1:-1] + z[2:, 1:-1, 1:-1] + z[1:-1, :-2, 1:-1] +
                            z[1:-1, 2:, 1:-1] + z[1:-1, 1:-1, :-2] + z[1:-1, 1:-1, 2:])
    # the last half-sweep was red: red nodes are exact for the final black values
    assert np.allclose(z[red], (nb[red] - 1e-4 * r[red]) / 6.0)
    assert np.all(z[~regular] == 0)


def test_precondition_oracle_closing_sweeps_match_node_loops():
    """The oracle's sweeps (numpy slices) against Gauss-Seidel written as
    node-by-node loops, with a zero network (no coarse correction):
    pre_sweeps red-black sweeps from 0, then post_sweeps black-red sweeps;
    and with many closing sweeps M(r) converges to the solution of A z = r."""
    n = 8
    h2 = 1e-4
    regular = np.zeros((n, n, n), bool)
    regular[1:-1, 1:-1, 1:-1] = True
    r = np.where(regular, np.random.default_rng(5).normal(size=(n, n, n)), 0.0)
    ch = (16, 16)
    p0 = np.zeros(U.n_params(ch))

    def gs_loops(pre, post):
        z = np.zeros((n, n, n))

        def half(colour):
            for k in range(1, n - 1):
                for j in range(1, n - 1):
                    for i in range(1, n - 1):
                        if (i + j + k) % 2 == colour:
                            s = (z[k - 1, j, i] + z[k + 1, j, i] + z[k, j - 1, i] + z[k, j + 1, i] +
                                 z[k, j, i - 1] + z[k, j, i + 1])
                            z[k, j, i] = (s - h2 * r[k, j, i]) / 6.0

        for _ in range(pre):
            half(0)
            half(1)
        for _ in range(post):
            half(1)
            half(0)
        return z

    for pre, post in ((1, 1), (3, 3), (2, 1)):
        z, corr = U.precondition(r, regular, h2, p0, ch, 2, pre_sweeps=pre, post_sweeps=post)
        assert np.all(corr == 0)
        assert np.allclose(z, gs_loops(pre, post), rtol=1e-12, atol=1e-14 * np.abs(z).max())
    z, _ = U.precondition(r, regular, h2, p0, ch, 2, post_sweeps=300)
    Az = np.zeros_like(z)
    Az[1:-1, 1:-1, 1:-1] = (z[:-2, 1:-1, 1:-1] + z[2:, 1:-1, 1:-1] + z[1:-1, :-2, 1:-1] + z[1:-1, 2:, 1:-1] +
                            z[1:-1, 1:-1, :-2] + z[1:-1, 1:-1, 2:] - 6.0 * z[1:-1, 1:-1, 1:-1]) / h2
    assert np.linalg.norm((Az - r)[regular]) <= 1e-10 * np.linalg.norm(r[regular])


def _ctx(pg, n):
    return pg.Context(n, n, n, [0, 0, 0], [1.0, 1.0, 1.0])


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("channels,n", [(C, 16), ((16, 32, 32, 32), 16), ((32, 32), 8),
                                        ((16, 64, 64), 16),
                                        ((16, 32, 32), 32)])  # level 1 at 16^3 pools in its conv
def test_gpu_unet_forward_and_gradient_match_oracle(channels, n):
    import pic_gpu as pg

    print("U-Net parity", channels, n, flush=True)
    ctx = _ctx(pg, n)
    ctx.unet_create(channels, seed=5)
    p = U.init_params(channels, seed=7).astype(np.float32)
    ctx.unet_set_params(p)
    # the device uses BF16 weights: compare against the oracle on BF16-rounded weights
    pb = _bf16(p).astype(np.float64)
    x, t = _inputs(n, 9)
    y = ctx.unet_forward(x.astype(np.float32))
    y_ref = U.forward(pb, x, channels)
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y - y_ref.ravel())) <= 3e-2 * scale
    loss, g = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    loss_ref, g_ref = U.loss_grad(pb, x, t, channels)
    assert loss == pytest.approx(loss_ref, rel=3e-2)
    off = 0
    for ci, co, taps in U.layer_specs(channels):
        nw = co * taps * ci + co
        a, b = g[off:off + nw], g_ref[off:off + nw]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
        assert cos > 0.99, (ci, co, taps, cos)
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b), (ci, co, taps)
        off += nw


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [(32, 32), (16, 32), (32, 64)])
def test_gpu_unet_forward_large_grid_matches_oracle(channels):
    """64^3: the large-grid convolution variants (register-blocked weights for
    32 and 64 output channels, the LDS-staged kernel for 16, the input and
    output layers, the fused up-convolution) against the FP64 oracle."""
    import pic_gpu as pg

    n = 64
    print("U-Net forward parity", channels, n, flush=True)
    ctx = _ctx(pg, n)
    ctx.unet_create(channels, seed=5)
    p = U.init_params(channels, seed=13).astype(np.float32)
    ctx.unet_set_params(p)
    x, _ = _inputs(n, 17)
    y = ctx.unet_forward(x.astype(np.float32))
    y_ref = U.forward(_bf16(p).astype(np.float64), x, channels)
    err = np.abs(y - y_ref.ravel())
    scale = np.max(np.abs(y_ref))
    assert np.max(err) <= 3e-2 * scale
    assert np.sqrt(np.mean(err ** 2)) <= 1e-2 * np.sqrt(np.mean(y_ref ** 2)) + 1e-3 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("coarsen", [1, 2])
def test_gpu_learned_preconditioner_matches_oracle(coarsen):
    """M(r) on the device (FP64 smoothing arithmetic on FP32 intermediate
    vectors, BF16 network) against the FP64 oracle with the same BF16-rounded
    weights; the smoothing carries FP32 rounding (well below 1e-6 relative),
    the network's correction the BF16 error."""
    import pic_gpu as pg

    n = 32
    ch = (16, 16, 16)
    h = 0.01
    L = h * (n - 1)
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    ctx.unet_create(ch, seed=1, coarsen=coarsen)
    p = U.init_params(ch, seed=11).astype(np.float32)
    ctx.unet_set_params(p)
    regular = np.zeros((n, n, n), bool)
    regular[1:-1, 1:-1, 1:-1] = True
    r = np.where(regular, np.random.default_rng(2).normal(size=(n, n, n)), 0.0)
    z = ctx.learned_precondition(r.ravel()).reshape(n, n, n)
    z_ref, corr = U.precondition(r, regular, h * h, _bf16(p).astype(np.float64), ch, coarsen)
    err = float(np.linalg.norm(z - z_ref))
    nc, nz = float(np.linalg.norm(corr)), float(np.linalg.norm(z_ref))
    assert nc > 1e-3 * nz  # the line-searched correction is present
    assert err <= 3e-2 * nc + 1e-6 * nz, (err, nc, nz)
    assert np.all(z[~regular] == 0)


@pytest.mark.gpu
def test_gpu_learned_smoothing_pair_and_four_node_kernels_match_oracle():
    """The cycle's smoothing sweeps against the FP64 oracle, with a zero
    network (no coarse correction), on the per-pair stencil kernels (30^3:
    i-lines not a multiple of 4) and the four-node ones (32^3).  Only the
    FP32 storage of the sweeps separates them from the oracle (3.8e-8
    relative measured)."""
    import pic_gpu as pg

    ch = (16, 16)
    h = 0.01
    for n in (30, 32):
        L = h * (n - 1)
        ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
        ctx.dirichlet_walls()
        ctx.unet_create(ch, seed=1, coarsen=1)
        p = np.zeros(U.n_params(ch), np.float32)
        ctx.unet_set_params(p)
        regular = np.zeros((n, n, n), bool)
        regular[1:-1, 1:-1, 1:-1] = True
        r = np.where(regular, np.random.default_rng(2).normal(size=(n, n, n)), 0.0)
        z = ctx.learned_precondition(r.ravel()).reshape(n, n, n)
        z_ref, corr = U.precondition(r, regular, h * h, p.astype(np.float64), ch, 1)
        assert np.all(corr == 0)
        err = float(np.linalg.norm(z - z_ref))
        nz = float(np.linalg.norm(z_ref))
        assert err <= 3e-7 * nz, (n, err, nz)
        assert np.all(z[~regular] == 0)
        ctx.close()


@pytest.mark.gpu
def test_gpu_unet_adam_reduces_loss():
    import pic_gpu as pg

    n = 16
    ctx = _ctx(pg, n)
    ctx.unet_create(C, seed=5)
    x, t = _inputs(n, 11)
    t = np.tanh(x[..., 0])  # a learnable target
    l0, _ = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    for _ in range(60):
        _, g = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
        ctx.unet_adam_step(g, 3e-3)
    l1, _ = ctx.unet_loss_grad(x.astype(np.float32), t.astype(np.float32))
    assert l1 < 0.5 * l0
