"""Train the learned Poisson preconditioner on the device and save its weights.

    python engineering-degree-in-plasma-simulations_amd/train_unet.py \
        [--nodes 128] [--channels 32 64 64 64] [--coarsen 2] [--rounds 10] \
        [--samples 500] [--epochs 2] [--out FILE]

Each round generates fresh training pairs on the device: a residual t
(random Gaussian blobs plus white noise, the particle-noise component of a
PIC charge density), what the cycle's pre-smoothing (three red-black
Gauss-Seidel sweeps) leaves of it (restricted to the network grid: the
input) and the exact error of that
remainder from the plain red-black SOR solver (restricted: the target); then
Adam runs over them.  The weights (FP32, the
parameter order of pic_unet_get_params) go to data/unet_<nodes>.f32 with a
JSON sidecar holding the architecture; bench.py and the learned solver tests
load them.  Finally the trained network is compared on a held-out density
with plain PCG, plain SOR and the same two-level cycle without a network.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import pic_gpu as pg  # noqa: E402


def weights_path(nodes):
    return os.path.join(HERE, "data", "unet_%d.f32" % nodes)


def held_out_density(n, seed=12345):
    """blobs plus particle-like noise, i-fastest, in C/m^3 for dx = 1 cm"""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = np.zeros((n, n, n))
    for _ in range(6):
        c = rng.uniform(4, n - 5, 3)
        s = rng.uniform(3, 0.15 * n)
        rho += rng.choice([-1, 1]) * np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)
                                            / (2 * s * s))
    rho += 0.1 * rng.normal(size=rho.shape)
    rho[0, :, :] = rho[-1, :, :] = rho[:, 0, :] = rho[:, -1, :] = rho[:, :, 0] = rho[:, :, -1] = 0
    return 1e-8 * rho.reshape(-1)


def compare(ctx, n, rho, params):
    """cold-start solves to 1e-6 of the held-out density"""
    ctx.set_field(pg.FIELD_RHO, rho)
    res = {}
    omega = 2.0 / (1.0 + np.sin(np.pi / (n - 1)))
    runs = (("pcg", pg.SOLVER_PCG, {}), ("sor", pg.SOLVER_SOR_RB, dict(omega=omega, check_every=25)),
            ("learned", pg.SOLVER_LEARNED, {}), ("zero_network", pg.SOLVER_LEARNED, {}))
    for name, m, kw in runs:
        if name == "zero_network":
            ctx.unet_set_params(np.zeros_like(params))
        ctx.set_field(pg.FIELD_PHI, np.zeros(ctx.nn))
        s = ctx.poisson(method=m, tol=1e-6, tol_kind=pg.TOL_REL, max_it=20000, **kw)
        res[name] = {"seconds": s["seconds"], "iterations": s["iterations"],
                     "rel_residual": s["rel_residual"], "converged": s["converged"]}
        print("%s: %.4f s, %d it, rel %.2e" % (name, s["seconds"], s["iterations"],
                                             s["rel_residual"]), flush=True)
    ctx.unet_set_params(params)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--channels", type=int, nargs="+", default=[32, 64, 64, 64])
    ap.add_argument("--coarsen", type=int, default=2)
    ap.add_argument("--rounds", type=int, default=10)
    ap.add_argument("--samples", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--noise", type=float, default=0.3)
    ap.add_argument("--sigma", type=float, nargs=2, default=(2.0, 20.0))
    ap.add_argument("--blobs", type=int, nargs=2, default=(0, 10))
    ap.add_argument("--rel-weight", type=float, default=0.9)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    n = args.nodes
    L = 1e-2 * (n - 1)
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    npar = ctx.unet_create(args.channels, seed=args.seed, coarsen=args.coarsen)
    if args.init:
        ctx.unet_set_params(np.fromfile(args.init, dtype="<f4"))
    print("U-Net channels %s, coarsen %d, %d parameters" % (args.channels, args.coarsen, npar),
          flush=True)
    hist = []
    t0 = time.time()
    for r in range(args.rounds):
        lr = args.lr * (0.5 ** (r // max(1, args.rounds // 3)))
        st = ctx.unet_train(samples=args.samples, epochs=args.epochs, batch=args.batch, lr=lr,
                            seed=args.seed * 1000 + r, blobs=tuple(args.blobs),
                            sigma=tuple(args.sigma), noise=args.noise, rel_weight=args.rel_weight)
        hist.append(st)
        print("round %d: loss %.4e -> %.4e (data %.1fs, train %.1fs, lr %.1e)"
              % (r, st["loss_first"], st["loss_last"], st["seconds_data"], st["seconds_train"], lr),
              flush=True)
    p = ctx.unet_get_params()
    out = args.out or weights_path(n)
    p.astype("<f4").tofile(out)
    meta = {"nodes": n, "channels": list(args.channels), "coarsen": args.coarsen,
            "params": int(p.size), "rounds": args.rounds, "samples_per_round": args.samples,
            "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "noise": args.noise,
            "sigma": list(args.sigma), "blobs": list(args.blobs), "seed": args.seed,
            "rel_weight": args.rel_weight,
            "loss_per_round": [h["loss_last"] for h in hist],
            "train_seconds": time.time() - t0}
    json.dump(meta, open(out + ".json", "w"), indent=1)
    meta["held_out"] = compare(ctx, n, held_out_density(n), p)
    json.dump(meta, open(out + ".json", "w"), indent=1)
    print(json.dumps(meta["held_out"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
