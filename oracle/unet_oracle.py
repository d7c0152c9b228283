"""FP64 numpy restatement of the learned preconditioner's U-Net (forward and
backward), the checker for the GPU training and inference kernels.

TEST INFRASTRUCTURE ONLY.  Layout and parameter order match the device
library: activations channels-last with voxel order (z, y, x) (= node order
i-fastest); each layer's weights are [cout][taps * cin] with k = tap * cin +
ci and tap = (dz+1)*9 + (dy+1)*3 + (dx+1), followed by [cout] biases; layers
in the order in, c1b, c2a, c2b, c3a, c3b, u2, c2c, u1, c1c, out.  The
backward pass is pinned by central finite differences in the tests.
"""
import numpy as np


def layer_specs(c1, c2, c3):
    return [(2, c1, 27), (c1, c1, 27), (c1, c2, 27), (c2, c2, 27), (c2, c3, 27), (c3, c3, 27),
            (c3, c2, 1), (c2, c2, 27), (c2, c1, 1), (c1, c1, 27), (c1, 1, 27)]


def unpack(params, c1, c2, c3):
    out, off = [], 0
    for cin, cout, taps in layer_specs(c1, c2, c3):
        nw = cout * taps * cin
        W = params[off:off + nw].reshape(cout, taps * cin)
        off += nw
        b = params[off:off + cout]
        off += cout
        out.append((W, b, taps))
    assert off == params.size
    return out


def n_params(c1, c2, c3):
    return sum(co * t * ci + co for ci, co, t in layer_specs(c1, c2, c3))


def conv(x, W, b, taps):
    nz, ny, nx, cin = x.shape
    if taps == 1:
        return x @ W.T + b
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.broadcast_to(b, (nz, ny, nx, W.shape[0])).astype(np.float64).copy()
    for t in range(27):
        ox, oy, oz = t % 3, (t // 3) % 3, t // 9
        y += xp[oz:oz + nz, oy:oy + ny, ox:ox + nx, :] @ W[:, t * cin:(t + 1) * cin].T
    return y


def conv_bwd(x, W, taps, dy):
    nz, ny, nx, cin = x.shape
    db = dy.sum(axis=(0, 1, 2))
    if taps == 1:
        dW = np.einsum("zyxo,zyxi->oi", dy, x)
        return dy @ W, dW, db
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    for t in range(27):
        ox, oy, oz = t % 3, (t // 3) % 3, t // 9
        sl = (slice(oz, oz + nz), slice(oy, oy + ny), slice(ox, ox + nx))
        dW[:, t * cin:(t + 1) * cin] = np.einsum("zyxo,zyxi->oi", dy, xp[sl])
        dxp[sl] += dy @ W[:, t * cin:(t + 1) * cin]
    return dxp[1:-1, 1:-1, 1:-1, :], dW, db


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_bwd(dc):
    return np.repeat(np.repeat(np.repeat(dc, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(df):
    nz, ny, nx, c = df.shape
    return df.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


relu = lambda a: np.maximum(a, 0.0)


def forward(params, x, c1, c2, c3, keep=False):
    """x: (nz, ny, nx, 2) -> y (nz, ny, nx)"""
    L = unpack(params, c1, c2, c3)
    a = {}
    a["x"] = x
    a["a1a"] = relu(conv(x, *L[0]))
    a["s1"] = relu(conv(a["a1a"], *L[1]))
    a["p1"] = pool(a["s1"])
    a["a2a"] = relu(conv(a["p1"], *L[2]))
    a["s2"] = relu(conv(a["a2a"], *L[3]))
    a["p2"] = pool(a["s2"])
    a["a3a"] = relu(conv(a["p2"], *L[4]))
    a["a3b"] = relu(conv(a["a3a"], *L[5]))
    a["u2"] = conv(a["a3b"], *L[6])
    a["a2"] = up(a["u2"]) + a["s2"]
    a["c2c"] = relu(conv(a["a2"], *L[7]))
    a["u1"] = conv(a["c2c"], *L[8])
    a["a1"] = up(a["u1"]) + a["s1"]
    a["c1c"] = relu(conv(a["a1"], *L[9]))
    y = conv(a["c1c"], *L[10])[..., 0]
    return (y, a) if keep else y


def loss_grad(params, x, target, c1, c2, c3):
    """0.5 mean over non-Dirichlet nodes (mask channel <= 0.5) of (y - t)^2"""
    L = unpack(params, c1, c2, c3)
    y, a = forward(params, x, c1, c2, c3, keep=True)
    valid = (x[..., 1] <= 0.5).astype(np.float64)
    nv = max(valid.sum(), 1.0)
    r = (y - target) * valid
    loss = 0.5 * np.sum(r * r) / nv
    grads = [None] * len(L)
    dy = (r / nv)[..., None]
    d, gw, gb = conv_bwd(a["c1c"], L[10][0], 27, dy)
    grads[10] = (gw, gb)
    d = d * (a["c1c"] > 0)
    d, gw, gb = conv_bwd(a["a1"], L[9][0], 27, d)
    grads[9] = (gw, gb)
    g_s1 = d.copy()
    d_u1 = up_bwd(d)
    d, gw, gb = conv_bwd(a["c2c"], L[8][0], 1, d_u1)
    grads[8] = (gw, gb)
    d = d * (a["c2c"] > 0)
    d, gw, gb = conv_bwd(a["a2"], L[7][0], 27, d)
    grads[7] = (gw, gb)
    g_s2 = d.copy()
    d_u2 = up_bwd(d)
    d, gw, gb = conv_bwd(a["a3b"], L[6][0], 1, d_u2)
    grads[6] = (gw, gb)
    d = d * (a["a3b"] > 0)
    d, gw, gb = conv_bwd(a["a3a"], L[5][0], 27, d)
    grads[5] = (gw, gb)
    d = d * (a["a3a"] > 0)
    d, gw, gb = conv_bwd(a["p2"], L[4][0], 27, d)
    grads[4] = (gw, gb)
    g_s2 = (g_s2 + pool_bwd(d)) * (a["s2"] > 0)
    d, gw, gb = conv_bwd(a["a2a"], L[3][0], 27, g_s2)
    grads[3] = (gw, gb)
    d = d * (a["a2a"] > 0)
    d, gw, gb = conv_bwd(a["p1"], L[2][0], 27, d)
    grads[2] = (gw, gb)
    g_s1 = (g_s1 + pool_bwd(d)) * (a["s1"] > 0)
    d, gw, gb = conv_bwd(a["a1a"], L[1][0], 27, g_s1)
    grads[1] = (gw, gb)
    d = d * (a["a1a"] > 0)
    _, gw, gb = conv_bwd(a["x"], L[0][0], 27, d)
    grads[0] = (gw, gb)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, flat


def init_params(c1, c2, c3, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout, taps in layer_specs(c1, c2, c3):
        s = np.sqrt(6.0 / (taps * cin))
        parts.append(rng.uniform(-s, s, cout * taps * cin))
        parts.append(rng.uniform(-0.1, 0.1, cout))
    return np.concatenate(parts)
