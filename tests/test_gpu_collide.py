"""GPU parity of DSMC_MEX (equal-weight DSMC with Bird's VHS cross section,
Interactions.cpp:143-284) and of the beam sources (Source.cpp:3-185)
against the oracle under the same Philox draws, on the reference fixture's
inputs."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TAG_DSMC = 0x44534D01
TAG_SRC = 0x53524301


@pytest.fixture(scope="module")
def pg():
    import pic_gpu

    return pic_gpu


def _grid_ctx(pg, gr):
    g = O.make_grid(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
    return g, pg.Context(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])


def _f32(p):
    q = p.reshape(-1, 7).copy()
    q[:, 3:7] = q[:, 3:7].astype(np.float32).astype(np.float64)
    return q.reshape(-1)


@pytest.mark.parametrize("two", [False, True])
def test_dsmc_mex_matches_oracle(pg, golden, two):
    gr = golden["dm_grid"]
    g, c = _grid_ctx(pg, gr)
    dt = gr[9]
    a0, b0 = _f32(golden["dm_a_in"]), _f32(golden["dm_b_in"])
    sa = c.species_add(16 * O.AMU, 0.0, 1e9, a0.size // 7 + 64)
    sb = c.species_add(40 * O.AMU, 0.0, 1e9, b0.size // 7 + 64)
    c.upload(sa, a0)
    c.upload(sb, b0)
    did = c.dsmc_create(sa, sb if two else None)
    sv = 1e-14
    for step in range(3):
        c.sort(sa)
        c.sort(sb)
        a, b = c.download(sa), c.download(sb)  # the device's cell order
        st = c.dsmc_apply(did, dt, 55, step)
        ra, rb, sv, nc, ng = O.dsmc_apply(g, a, 16 * O.AMU, 1e9, dt, sv,
                                          O.rng_philox(55, step, TAG_DSMC),
                                          p2=b if two else None, mass2=40 * O.AMU)
        print("step", step, st, nc, ng, flush=True)
        assert st["groups"] == ng and st["collisions"] == nc > 100
        assert st["sigma_v_rel_max"] == pytest.approx(sv, rel=1e-6)
        ga = c.download(sa).reshape(-1, 7)
        vs = np.max(np.abs(ra.reshape(-1, 7)[:, 3:6]))
        assert np.max(np.abs(ga[:, 3:6] - ra.reshape(-1, 7)[:, 3:6])) <= 1e-5 * vs
        assert np.array_equal(ga[:, :3], a.reshape(-1, 7)[:, :3])  # positions untouched
        if two:
            gb = c.download(sb).reshape(-1, 7)
            assert np.max(np.abs(gb[:, 3:6] - rb.reshape(-1, 7)[:, 3:6])) <= 1e-5 * vs


@pytest.mark.parametrize("face,T", [(0, 0.0), (1, 0.0), (3, 4000.0), (5, 5000.0)])
def test_source_matches_oracle(pg, golden, face, T):
    gr = golden["grid"]
    g, c = _grid_ctx(pg, gr)
    c.set_dt(2e-9)
    ef = golden["ef_analytic"]
    c.set_field(pg.FIELD_EF, ef)
    L = np.array(g.dx[:]) * (np.array([g.ni, g.nj, g.nk]) - 1)
    xc = 0.5 * (np.array(g.x0[:]) + np.array(g.xm[:]))
    rect = ([xc[0], xc[1], g.xm[2]], [L[0] * 0.2, L[1] * 0.15, L[2] * 0.1], 0.0)
    c.add_rectangle(rect[0], 2 * np.array(rect[1]), 0.0)
    sid = c.species_add(16 * O.AMU, O.QE, 5e3, 200000)
    v, den = (3e3, 2e14) if T else (2e5, 2e12)
    src = c.source_create(sid, face, v, den, T)
    oid, nt = c.node_flags()
    for call in range(2):
        n = c.source_sample(src, 77)
        ref = O.source_sample(g, O.rng_philox(77, call, TAG_SRC), O.QE, 16 * O.AMU, 5e3, face,
                              v, den, T, 2e-9, ef=ef, rects=[rect])
        assert n >= ref.size // 7 > 100
    got = c.download(sid).reshape(-1, 7)
    # both calls' particles, in injection order (the species is not sorted)
    r0 = O.source_sample(g, O.rng_philox(77, 0, TAG_SRC), O.QE, 16 * O.AMU, 5e3, face, v, den, T,
                         2e-9, ef=ef, rects=[rect])
    r1 = O.source_sample(g, O.rng_philox(77, 1, TAG_SRC), O.QE, 16 * O.AMU, 5e3, face, v, den, T,
                         2e-9, ef=ef, rects=[rect])
    ref = np.concatenate([r0, r1]).reshape(-1, 7)
    assert got.shape == ref.shape
    assert np.max(np.abs(got[:, :3] - ref[:, :3]) / L) < 1e-6
    assert np.max(np.abs(got[:, 3:6] - ref[:, 3:6])) <= 1e-5 * np.max(np.abs(ref[:, 3:6]))
    # the inlet face is Dirichlet (World::addInlet)
    ax, plus = face // 2, face & 1
    nn = (g.ni, g.nj, g.nk)
    idx = np.arange(g.ni * g.nj * g.nk)
    ijk = [idx % g.ni, (idx // g.ni) % g.nj, idx // (g.ni * g.nj)]
    on_face = ijk[ax] == (nn[ax] - 1 if plus else 0)
    assert np.all(nt[on_face] == 2)
