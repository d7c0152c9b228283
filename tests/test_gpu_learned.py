"""The learned preconditioner end to end on the device: train on pairs made by
the plain solver, then solve a new problem with flexible PCG and compare
with the plain method.  The final residual is checked by the FP64 oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _blob_density(n, seed):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(3, n - 4, 3)
        s = rng.uniform(2, 6)
        rho += rng.choice([-1, 1]) * np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)
                                            / (2 * s * s))
    rho = np.transpose(rho, (2, 1, 0)).reshape(-1)  # i-fastest
    rho += 0.05 * rng.normal(size=rho.size)
    return 1e-8 * rho


def test_learned_preconditioner_trains_and_beats_plain_pcg():
    import pic_gpu as pg

    n = 32
    L = 0.01 * (n - 1)
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    ctx.unet_create(16, 16, 32, seed=3)
    st = ctx.unet_train(samples=128, epochs=20, batch=4, lr=2e-3, seed=5, blobs=(1, 6),
                        sigma=(2.0, 8.0), noise=0.05)
    assert st["loss_last"] < 0.2 * st["loss_first"], st
    rho = _blob_density(n, 42)
    ctx.set_field(pg.FIELD_RHO, rho)
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    results = {}
    for name, method in (("plain", pg.SOLVER_PCG), ("learned", pg.SOLVER_LEARNED)):
        ctx.set_field(pg.FIELD_PHI, np.zeros(ctx.nn))
        s = ctx.poisson(method=method, tol=1e-6, tol_kind=pg.TOL_REL, max_it=5000)
        assert s["converged"] and s["rel_residual"] <= 1e-6, (name, s)
        phi = ctx.get_field(pg.FIELD_PHI)
        r2, b2 = O.residual_dirichlet(g, phi, rho)
        assert r2 / b2 <= 1.05e-6, name
        results[name] = (s, phi)
    it_plain = results["plain"][0]["iterations"]
    it_learned = results["learned"][0]["iterations"]
    assert it_learned < 0.5 * it_plain, (it_learned, it_plain)
    a, b = results["plain"][1], results["learned"][1]
    assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(a))


def test_learned_solver_refuses_without_network():
    import pic_gpu as pg

    ctx = pg.Context(16, 16, 16, [0, 0, 0], [0.15, 0.15, 0.15])
    ctx.dirichlet_walls()
    with pytest.raises(pg.PicError):
        ctx.poisson(method=pg.SOLVER_LEARNED, tol=1e-6)
