"""FP64 numpy restatement of the learned preconditioner's U-Net (forward and
backward), the checker for the GPU training and inference kernels.

TEST INFRASTRUCTURE ONLY.  Layout and parameter order match the device
library: activations channels-last with voxel order (z, y, x) (x fastest);
each layer's weights are [cout][taps * cin] with k = tap * cin + ci and
tap = (dz+1)*9 + (dy+1)*3 + (dx+1), followed by [cout] biases.  With channels
C_0..C_{D-1} the layers are
    in (2->C0), B0 (C0->C0),
    A_l (C_{l-1}->C_l), B_l (C_l->C_l)          for l = 1..D-1,
    U_l (1x1, C_{l+1}->C_l), Dec_l (C_l->C_l)   for l = D-2..0,
    out (C0->1).
The backward pass is pinned by central finite differences in the tests.
"""
import numpy as np


def layer_specs(channels):
    C = list(channels)
    D = len(C)
    s = [(2, C[0], 27), (C[0], C[0], 27)]
    for l in range(1, D):
        s += [(C[l - 1], C[l], 27), (C[l], C[l], 27)]
    for l in range(D - 2, -1, -1):
        s += [(C[l + 1], C[l], 1), (C[l], C[l], 27)]
    s.append((C[0], 1, 27))
    return s


def unpack(params, channels):
    out, off = [], 0
    for cin, cout, taps in layer_specs(channels):
        nw = cout * taps * cin
        W = params[off:off + nw].reshape(cout, taps * cin)
        off += nw
        b = params[off:off + cout]
        off += cout
        out.append((W, b, taps))
    assert off == params.size
    return out


def n_params(channels):
    return sum(co * t * ci + co for ci, co, t in layer_specs(channels))


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


def _index(channels):
    """layer indices: in, B[l], A[l], U[l], Dec[l], out"""
    D = len(channels)
    idx = {"in": 0, "B": {0: 1}, "A": {}, "U": {}, "Dec": {}}
    k = 2
    for l in range(1, D):
        idx["A"][l], idx["B"][l] = k, k + 1
        k += 2
    for l in range(D - 2, -1, -1):
        idx["U"][l], idx["Dec"][l] = k, k + 1
        k += 2
    idx["out"] = k
    return idx


def forward(params, x, channels, keep=False):
    """x: (nz, ny, nx, 2) -> y (nz, ny, nx)"""
    L = unpack(params, channels)
    I = _index(channels)
    D = len(channels)
    a = {"x": x}
    a["a0"] = relu(conv(x, *L[I["in"]]))
    a["s0"] = relu(conv(a["a0"], *L[I["B"][0]]))
    for l in range(1, D):
        a["p%d" % l] = pool(a["s%d" % (l - 1)])
        a["a%d" % l] = relu(conv(a["p%d" % l], *L[I["A"][l]]))
        a["s%d" % l] = relu(conv(a["a%d" % l], *L[I["B"][l]]))
    for l in range(D - 2, -1, -1):
        src = a["s%d" % (D - 1)] if l == D - 2 else a["h%d" % (l + 1)]
        a["u%d" % l] = conv(src, *L[I["U"][l]])
        a["m%d" % l] = up(a["u%d" % l]) + a["s%d" % l]
        a["h%d" % l] = relu(conv(a["m%d" % l], *L[I["Dec"][l]]))
    y = conv(a["h0"], *L[I["out"]])[..., 0]
    return (y, a) if keep else y


def loss_grad(params, x, target, channels):
    """0.5 mean over voxels holding a regular node (mask channel < 1) of (y - t)^2"""
    L = unpack(params, channels)
    I = _index(channels)
    D = len(channels)
    y, a = forward(params, x, channels, keep=True)
    valid = (x[..., 1] < 0.999).astype(np.float64)
    nv = max(valid.sum(), 1.0)
    r = (y - target) * valid
    loss = 0.5 * np.sum(r * r) / nv
    grads = [None] * len(L)
    dy = (r / nv)[..., None]
    d, gw, gb = conv_bwd(a["h0"], L[I["out"]][0], 27, dy)
    grads[I["out"]] = (gw, gb)
    gh = d * (a["h0"] > 0)  # at h0's pre-activation
    gm = {}
    for l in range(0, D - 1):
        d, gw, gb = conv_bwd(a["m%d" % l], L[I["Dec"][l]][0], 27, gh)
        grads[I["Dec"][l]] = (gw, gb)
        gm[l] = d
        gu = up_bwd(d)
        src = a["s%d" % (D - 1)] if l == D - 2 else a["h%d" % (l + 1)]
        d, gw, gb = conv_bwd(src, L[I["U"][l]][0], 1, gu)
        grads[I["U"][l]] = (gw, gb)
        gh = d * (src > 0)  # at src's pre-activation
    gs = gh  # at s_{D-1}'s pre-activation
    gp = None
    for l in range(D - 1, 0, -1):
        if l < D - 1:
            gs = (gm[l] + pool_bwd(gp)) * (a["s%d" % l] > 0)
        d, gw, gb = conv_bwd(a["a%d" % l], L[I["B"][l]][0], 27, gs)
        grads[I["B"][l]] = (gw, gb)
        ga = d * (a["a%d" % l] > 0)
        gp, gw, gb = conv_bwd(a["p%d" % l], L[I["A"][l]][0], 27, ga)
        grads[I["A"][l]] = (gw, gb)
    gs = (gm[0] + pool_bwd(gp)) * (a["s0"] > 0)
    d, gw, gb = conv_bwd(a["a0"], L[I["B"][0]][0], 27, gs)
    grads[I["B"][0]] = (gw, gb)
    ga = d * (a["a0"] > 0)
    _, gw, gb = conv_bwd(a["x"], L[I["in"]][0], 27, ga)
    grads[I["in"]] = (gw, gb)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, flat


def init_params(channels, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout, taps in layer_specs(channels):
        s = np.sqrt(6.0 / (taps * cin))
        parts.append(rng.uniform(-s, s, cout * taps * cin))
        parts.append(rng.uniform(-0.1, 0.1, cout))
    return np.concatenate(parts)


# ---- the two-level preconditioner's grid transfers (aggregates of 2x2x2 nodes)

def restrict_mean(f, regular):
    """mean over each aggregate's regular nodes; arrays (nz, ny, nx)"""
    nz, ny, nx = f.shape
    r = regular.astype(np.float64)
    s = (f * r).reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(1, 3, 5))
    n = r.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(1, 3, 5))
    return np.where(n > 0, s / np.maximum(n, 1), 0.0), 1.0 - n / 8.0


def prolong_trilinear(y):
    """cell-centred trilinear interpolation from aggregate centres: per axis,
    weight 3/4 on the node's aggregate and 1/4 on the neighbour on its side
    (the aggregate itself at the grid's ends)"""
    out = y
    for ax in range(3):
        n = out.shape[ax]
        idx = np.arange(2 * n)
        c = idx // 2
        nb = np.clip(c + np.where(idx % 2 == 1, 1, -1), 0, n - 1)
        out = 0.75 * np.take(out, c, axis=ax) + 0.25 * np.take(out, nb, axis=ax)
    return out


def precondition(r, regular, h2, params, channels, coarsen, pre_sweeps=3, post_sweeps=3):
    """z = M(r) of the learned solver on a Dirichlet-walled uniform grid;
    arrays (nz, ny, nx).  pre_sweeps red-black Gauss-Seidel sweeps on
    A z = r from 0 (A = 7-point Laplacian / h2 on regular nodes; red =
    (i+j+k) even), the network's coarse correction c of the remaining
    residual res = r - A z, added as z += omega c with the A-norm optimal
    omega = (c.res)/(c.A c), post_sweeps sweeps in the reverse colour order
    (black, red).  Returns (z, omega c)."""
    nz, ny, nx = r.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    par = (i + j + k) % 2
    z = np.zeros_like(r, dtype=np.float64)

    def nbr_sum(v):
        s = np.zeros_like(v)
        s[1:-1, 1:-1, 1:-1] = (v[:-2, 1:-1, 1:-1] + v[2:, 1:-1, 1:-1] + v[1:-1, :-2, 1:-1] +
                               v[1:-1, 2:, 1:-1] + v[1:-1, 1:-1, :-2] + v[1:-1, 1:-1, 2:])
        return s

    def sweep(color):
        m = regular & (par == color)
        z[m] = (nbr_sum(z)[m] - h2 * r[m]) / 6.0

    for _ in range(pre_sweeps):
        sweep(0)
        sweep(1)
    Az = (nbr_sum(z) - 6.0 * z) / h2
    res = np.where(regular, r - Az, 0.0)
    if coarsen == 2:
        xin, mask = restrict_mean(-h2 * res, regular)
    else:
        xin, mask = np.where(regular, -h2 * res, 0.0), 1.0 - regular
    valid = mask < 0.999
    s = np.sqrt(np.mean(xin[valid] ** 2))
    x = np.stack([xin / s, mask], axis=-1)
    y = forward(params, x, channels)
    c = prolong_trilinear(y) if coarsen == 2 else y
    c = np.where(regular, c, 0.0)
    Ac = np.where(regular, (nbr_sum(c) - 6.0 * c) / h2, 0.0)
    den = np.sum(c * Ac)
    omega = np.sum(c * res) / den if den != 0 else 0.0
    corr = omega * c
    z += corr
    for _ in range(post_sweeps):
        sweep(1)
        sweep(0)
    return z, corr
