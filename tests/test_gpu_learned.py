"""The learned preconditioner end to end on the device: train on pairs made by
the plain solver, then solve a new problem with flexible CG and compare with
the same two-level cycle without a network (a zero network: a symmetric
Gauss-Seidel preconditioner) and with plain PCG.  The final residual is
checked by the FP64 oracle."""
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


def _solve(ctx, pg, method, rho, g):
    ctx.set_field(pg.FIELD_RHO, rho)
    ctx.set_field(pg.FIELD_PHI, np.zeros(ctx.nn))
    s = ctx.poisson(method=method, tol=1e-6, tol_kind=pg.TOL_REL, max_it=5000)
    assert s["converged"] and s["rel_residual"] <= 1e-6, s
    phi = ctx.get_field(pg.FIELD_PHI)
    r2, b2 = O.residual_dirichlet(g, phi, rho)
    assert r2 / b2 <= 1.05e-6
    return s, phi


@pytest.mark.parametrize("coarsen", [2])
def test_learned_preconditioner_trains_and_beats_untrained_cycle(coarsen):
    import pic_gpu as pg

    n = 32
    L = 0.01 * (n - 1)
    print("learned solver, coarsen", coarsen, flush=True)
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    rho = _blob_density(n, 42)
    ctx.unet_create((16, 16, 16), seed=3, coarsen=coarsen)
    st = ctx.unet_train(samples=128, epochs=20, batch=4, lr=2e-3, seed=5, blobs=(1, 6),
                        sigma=(2.0, 8.0), noise=0.05)
    assert st["loss_last"] < 0.4 * st["loss_first"], st
    s_learned, phi_l = _solve(ctx, pg, pg.SOLVER_LEARNED, rho, g)
    trained = ctx.unet_get_params()
    ctx.unet_set_params(np.zeros_like(trained))
    s_zero, _ = _solve(ctx, pg, pg.SOLVER_LEARNED, rho, g)
    ctx.unet_set_params(trained)
    s_plain, phi_p = _solve(ctx, pg, pg.SOLVER_PCG, rho, g)
    print("iterations: plain PCG %d, zero network %d, trained %d" % (
        s_plain["iterations"], s_zero["iterations"], s_learned["iterations"]), flush=True)
    assert s_learned["iterations"] < 0.7 * s_zero["iterations"]
    assert s_learned["iterations"] < 0.4 * s_plain["iterations"]
    assert np.max(np.abs(phi_l - phi_p)) <= 1e-4 * np.max(np.abs(phi_p))


def test_learned_solver_refuses_without_network():
    import pic_gpu as pg

    ctx = pg.Context(16, 16, 16, [0, 0, 0], [0.15, 0.15, 0.15])
    ctx.dirichlet_walls()
    with pytest.raises(pg.PicError):
        ctx.poisson(method=pg.SOLVER_LEARNED, tol=1e-6)


def test_learned_solver_time_loop_checks():
    """Learned solves in sequence on one context, as in a time loop: after the
    first, a solve reads its initial residual with its first convergence check
    and queues the final FP64 residual with the check where it expects to
    converge.  Every solve reports the oracle's residual of the phi it leaves;
    a solve that starts converged (the same density again) stays converged."""
    import pic_gpu as pg

    n = 32
    L = 0.01 * (n - 1)
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    ctx.unet_create((16, 16, 16), seed=3, coarsen=2)
    ctx.unet_set_params(ctx.unet_get_params())  # the random initial network: more iterations, same checks
    rho0 = _blob_density(n, 7)
    drho = 0.05 * _blob_density(n, 8)
    ctx.set_field(pg.FIELD_PHI, np.zeros(ctx.nn))
    kw = dict(method=pg.SOLVER_LEARNED, tol=1e-6, tol_kind=pg.TOL_REL, max_it=5000)
    its = []
    for t in range(6):
        rho = rho0 + t * drho
        ctx.set_field(pg.FIELD_RHO, rho)
        s = ctx.poisson(flags=pg.SOLVE_EXTRAPOLATE, **kw)
        phi = ctx.get_field(pg.FIELD_PHI)
        r2, b2 = O.residual_dirichlet(g, phi, rho)
        print("step", t, s["iterations"], s["rel_residual"], r2 / b2, flush=True)
        assert s["converged"] and s["rel_residual"] <= 1e-6, (t, s)
        assert s["rel_residual"] == pytest.approx(r2 / b2, rel=1e-6), t
        its.append(s["iterations"])
    # a density changing linearly in time: the extrapolated start is almost
    # the solution, and the checks must not carry the solve far past it
    assert its[-1] < 0.6 * its[0], its
    s = ctx.poisson(**kw)  # already converged: stays so
    phi2 = ctx.get_field(pg.FIELD_PHI)
    r2, b2 = O.residual_dirichlet(g, phi2, rho)
    assert s["converged"] and r2 / b2 <= 1e-6 and s["rel_residual"] == pytest.approx(r2 / b2, rel=1e-6)
    assert np.max(np.abs(phi2 - phi)) <= 1e-5 * np.max(np.abs(phi))
