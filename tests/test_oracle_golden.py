"""Pin the CPU oracle against vectors written by the compiled reference.

tests/golden/ref_v3.npz was produced by tests/golden/gen_golden.py, which
links the reference's own ch4/v3 sources with our harness main.  Where the
oracle follows the reference's operation order the comparison is bit-exact.
"""
import numpy as np
import pytest

import oracle as O


def _grid(gr):
    return O.make_grid(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])


def _L(g):
    return np.array(g.dx[:]) * (np.array([g.ni, g.nj, g.nk]) - 1)


def _xc(g):
    return 0.5 * (np.array(g.xm[:]) + np.array(g.x0[:]))


@pytest.fixture(scope="module")
def G(golden):
    return _grid(golden["grid"])


def test_node_volumes_exact(G, golden):
    assert np.array_equal(O.node_volumes(G), golden["node_vol"])


def test_efield_exact(G, golden):
    assert np.array_equal(O.efield(G, golden["phi_analytic"]), golden["ef_analytic"])


def test_loader_mt19937_exact(G, golden):
    ef = golden["ef_analytic"]
    dt0 = golden["grid"][9]
    e = O.load_box_thermal(G, O.rng_mt(12345), -O.QE, O.ME, dt0, _xc(G), _L(G) * 0.98, 1e12,
                           2.0 * 11604.52, 1.5e5, ef=ef)
    assert np.array_equal(e, golden["ele_loaded"])
    i = O.load_box_thermal(G, O.rng_mt(777), O.QE, 16 * O.AMU, dt0, _xc(G), _L(G) * 0.9, 1e12,
                           0.1 * 11604.52, 2e5, ef=ef)
    assert np.array_equal(i, golden["ion_loaded"])


def test_push_exact(G, golden):
    ef = golden["ef_analytic"]
    e = O.push(G, ef, golden["ele_loaded"], -O.QE, O.ME, 2e-9)
    assert np.array_equal(e, golden["ele_pushed"])
    i = O.push(G, ef, golden["ion_loaded"], O.QE, 16 * O.AMU, 2e-9)
    assert np.array_equal(i, golden["ion_pushed"])


def test_push_removal_order_exact(G, golden):
    """Particles leaving the box are swap-removed exactly as the reference does."""
    ef = golden["ef_analytic"]
    pre = golden["ele_pre_long"]
    e = O.push(G, ef, pre, -O.QE, O.ME, 4e-8)
    assert e.size < pre.size
    assert np.array_equal(e, golden["ele_post_long"])
    # the stable variant keeps the same set
    s = O.push(G, ef, pre, -O.QE, O.ME, 4e-8, stable=True)
    a = np.sort(e.reshape(-1, 7), axis=0)
    b = np.sort(s.reshape(-1, 7), axis=0)
    assert np.array_equal(a, b)


def test_density_moments_counts_rho_exact(G, golden):
    ele = golden["ele_post_long"]
    ion = golden["ion_pushed"]
    den_e = O.number_density(G, ele)
    den_i = O.number_density(G, ion)
    assert np.array_equal(den_e, golden["ele_den"])
    assert np.array_equal(den_i, golden["ion_den"])
    acc = O.sample_moments(G, ele)
    vel, T = O.gas_properties(G, O.ME, acc)
    assert np.array_equal(vel, golden["ele_vel"])
    assert np.array_equal(T, golden["ele_T"])
    assert np.array_equal(O.macro_count(G, ele), golden["ele_macro_count"])
    rho = np.zeros(O.nn(G))
    O.lib().orc_charge_accumulate(rho.size, rho, -O.QE, den_e)
    O.lib().orc_charge_accumulate(rho.size, rho, O.QE, den_i)
    assert np.array_equal(rho, golden["rho"])


def test_diagnostic_sums(G, golden):
    phi = golden["phi_analytic"]
    d = golden["diag"]
    se = O.species_sums(G, golden["ele_post_long"], O.ME, -O.QE, phi)
    si = O.species_sums(G, golden["ion_pushed"], 16 * O.AMU, O.QE, phi)
    assert np.array_equal(se, d[0:6])
    assert np.array_equal(si, d[6:12])
    assert O.field_energy(G, golden["ef_analytic"]) == d[12]


def test_gs_v3_neumann_object_exact(G, golden):
    oid = golden["gs_object_id"].astype(np.int32)
    assert oid.sum() > 0
    phi, ok, it, l2 = O.gs_v3(G, oid, golden["gs_phi_in"], golden["rho"], 0.0, 0.0, 1e20,
                              4000, 1e-4)
    assert ok == bool(golden["gs_ok"][0])
    assert np.array_equal(phi, golden["gs_phi_out"])


def test_nrpcg_dirichlet_linear_exact(G, golden):
    nt = np.zeros(O.nn(G), np.int32)
    O.lib().orc_dirichlet_walls(G, nt)
    oid = np.zeros(O.nn(G), np.int32)
    phi, ok, nit, pit, l2 = O.nrpcg(G, oid, nt, np.zeros(O.nn(G)), golden["rho"], 0.0, 0.0,
                                    1e20, 500, 1e-6)
    assert ok == 1 and bool(golden["pcg_ok"][0])
    assert np.array_equal(phi, golden["pcg_phi_out"])
    assert np.array_equal(O.efield(G, phi), golden["pcg_ef"])


def test_nrpcg_boltzmann_exact(G, golden):
    nt = np.zeros(O.nn(G), np.int32)
    O.lib().orc_dirichlet_walls(G, nt)
    oid = np.zeros(O.nn(G), np.int32)
    rho = np.full(O.nn(G), O.QE * 1.5e12)
    phi, ok, nit, pit, l2 = O.nrpcg(G, oid, nt, np.zeros(O.nn(G)), rho, 0.0, 1.0e12, 3.0, 500,
                                    1e-8)
    assert ok == 1 and bool(golden["nr_ok"][0])
    assert nit >= 2  # genuinely nonlinear
    assert np.array_equal(phi, golden["nr_phi_out"])


def test_mcc_apply_mt19937_exact(golden, sigma_table_path):
    """Three MC_MEX_Ionization::apply calls, draw for draw against the reference."""
    gr = golden["mcc_grid"]
    g = _grid(gr)
    dt = gr[9]
    sE, sV = O.load_sigma_table(sigma_table_path)
    E_ion = 1313.9 * 1000 / 6.02214076e23
    prm = O.mcc_params(16 * O.AMU, O.ME, E_ion, sE, sV, dt, g.dx[0] * g.dx[1] * g.dx[2],
                       pool_append=True)
    neu, ele = golden["mcc_neu_in"], golden["mcc_ele_in"]
    ion = np.zeros(0)
    wsv = golden["mcc_wsv0"][0]
    rng = O.rng_mt(99)
    total_iz = 0
    for call in range(3):
        neu, ion, ele, wsv, st = O.mcc_apply(g, prm, wsv, rng, neu, ion, ele)
        total_iz += st.ionizations
        s = str(call)
        assert wsv == golden["mcc_wsv_" + s][0]
        assert np.array_equal(neu, golden["mcc_neu_" + s])
        assert np.array_equal(ion, golden["mcc_ion_" + s])
        assert np.array_equal(ele, golden["mcc_ele_" + s])
    assert total_iz > 100


def test_push_threaded_exact(G, golden):
    """The reference's threaded electron push (3 threads): every particle is
    pushed, and none is removed -- the per-thread deletion buffers reach the
    workers as std::bind copies (ThreadPool.h:63, :77-78), so the buffers
    Species.cpp:287-308 reads stay empty."""
    ef = golden["ef_analytic"]
    p = golden["ele_pre_long"].copy()
    n = O.push_threaded(G, ef, p, -O.QE, O.ME, 4e-8, 3)
    assert n == p.size // 7
    assert np.array_equal(p, golden["thr_post_long"])
    # the pushed positions equal the serial push's (which removes leavers)
    lo, hi = np.array(G.x0[:]), np.array(G.xm[:])
    rows = p.reshape(-1, 7)
    inside = rows[np.all((rows[:, :3] >= lo) & (rows[:, :3] < hi), axis=1)]
    a = np.sort(inside, axis=0)
    b = np.sort(golden["ele_post_long"].reshape(-1, 7), axis=0)
    assert np.array_equal(a, b)


def test_push_threaded_with_removal_keeps_serial_set(G, golden):
    """remove=True applies the removal the source intends (descending
    swap-with-last over the merged buffers): the survivors are exactly the
    serial push's survivors."""
    ef = golden["ef_analytic"]
    for threads in (1, 3, 8):
        p = golden["ele_pre_long"].copy()
        n = O.push_threaded(G, ef, p, -O.QE, O.ME, 4e-8, threads, remove=True)
        a = np.sort(p[: 7 * n].reshape(-1, 7), axis=0)
        b = np.sort(golden["ele_post_long"].reshape(-1, 7), axis=0)
        assert np.array_equal(a, b)


def _surface_setup(golden):
    G = _grid(golden["grid"])
    r = golden["sf_rects"].reshape(-1, 6)
    rects = [(row[:3], row[3:], 0.0) for row in r]
    return G, rects, golden["sf_ef"], golden["sf_dt"]


def test_push_electrons_absorbed_in_objects_exact(golden):
    G, rects, ef, dts = _surface_setup(golden)
    e = O.push(G, ef, golden["sf_ele_in"], -O.QE, O.ME, dts[0], rects=rects)
    assert e.size < golden["sf_ele_in"].size
    assert np.array_equal(e, golden["sf_ele_out"])


def test_push_ions_recycled_into_neutrals_exact(golden):
    """Ions hitting an object are removed and emit int(w/mpw0_n + U)
    neutrals with diffusely reflected velocities; addParticle rejects an
    emission point that still tests inside the object (Species.cpp:220-236)."""
    G, rects, ef, dts = _surface_setup(golden)
    rng = O.rng_mt(556)
    ion, neu, st = O.push_nonelectron(G, ef, golden["sf_ion_in"], O.QE, 16 * O.AMU, dts[1],
                                      rects, 1e10, rng, neu=golden["sf_neu_in"])
    assert st[0] > 100 and st[1] > 50 and st[2] > 0  # hits, emissions, rejected emissions
    assert np.array_equal(ion, golden["sf_ion_out"])
    assert np.array_equal(neu, golden["sf_neu_mid"])


def test_push_neutrals_reflected_exact(golden):
    """Neutrals reflect diffusely at 300 K and drift on for the rest of the
    step (Species.cpp:216-219, :835-861)."""
    G, rects, ef, dts = _surface_setup(golden)
    rng = O.rng_mt(557)
    neu, _, st = O.push_nonelectron(G, ef, golden["sf_neu_mid"], 0.0, 16 * O.AMU, dts[2], rects,
                                    1e10, rng)
    assert st[0] > 100 and st[1] == 0
    assert np.array_equal(neu, golden["sf_neu_out"])


def test_line_intersect_closed_forms():
    """Entry time, point and outward normal of a segment into a box; a
    segment through an edge gets the normalised sum of both normals."""
    box = ([0.0, 0.0, 0.0], [1.0, 2.0, 3.0], 0.0)
    t, pos, n = O.rect_line_intersect(box, [-3.0, 0.5, 0.1], [0.0, 0.5, 0.1])
    assert t == pytest.approx(2.0 / 3.0) and np.allclose(pos, [-1.0, 0.5, 0.1])
    assert np.array_equal(n, [-1.0, 0.0, 0.0])
    t, pos, n = O.rect_line_intersect(box, [0.2, 5.0, -1.0], [0.2, 1.0, -1.0])
    assert t == pytest.approx(0.75) and np.allclose(pos, [0.2, 2.0, -1.0])
    assert np.array_equal(n, [0.0, 1.0, 0.0])
    t, pos, n = O.rect_line_intersect(box, [-2.0, -3.0, 0.0], [0.0, -1.0, 0.0])
    assert t == pytest.approx(0.5) and np.allclose(pos, [-1.0, -2.0, 0.0])
    assert np.allclose(n, [-np.sqrt(0.5), -np.sqrt(0.5), 0.0])


def test_pcg_gs_fallback_exact(G, golden):
    """Newton-PCG whose PCG stops after 2 iterations: every Newton step then
    runs the Gauss-Seidel fallback (PotentialSolver.cpp:222-224, :315-347)."""
    nt = np.zeros(O.nn(G), np.int32)
    O.lib().orc_dirichlet_walls(G, nt)
    oid = np.zeros(O.nn(G), np.int32)
    phi, ok, nit, pit, l2 = O.nrpcg(G, oid, nt, np.zeros(O.nn(G)), golden["rho"], 0.0, 0.0,
                                    1e20, 2, 1e-6)
    assert ok == bool(golden["fb_ok"][0])
    assert nit > 1
    assert np.array_equal(phi, golden["fb_phi_out"])


@pytest.mark.parametrize("case,charge,mass,seed", [("neu", 0.0, 16 * 1.660538921e-27, 31415),
                                                   ("ele", -1.602176565e-19, 9.10938215e-31,
                                                    27182)])
def test_merge_exact(G, golden, case, charge, mass, seed):
    """Species::merge draw for draw: velocity bins over the cell and
    particles[0], pairs conserving weight, momentum and per-axis energy,
    appended through addParticle (half-step rewind), zero-weight
    particles swap-removed (Species.cpp:987-1145)."""
    pin = golden["mg_%s_in" % case]
    out, st = O.merge(G, pin, charge, mass, 2e-9, O.rng_mt(seed), ef=golden["ef_analytic"])
    assert st[0] >= 10
    assert out.size < pin.size
    assert np.array_equal(out, golden["mg_%s_out" % case])


def test_dsmc_mex_one_and_two_species_exact(golden):
    """DSMC_MEX::apply draw for draw: equal-weight pairs per cell, Bird's VHS
    cross section, isotropic scattering (Interactions.cpp:143-284)"""
    gr = golden["dm_grid"]
    g = _grid(gr)
    dt = gr[9]
    a, b = golden["dm_a_in"], golden["dm_b_in"]
    rng = O.rng_mt(8081)
    sv = 1e-14
    total = 0
    for call in range(2):
        a, _, sv, nc, ng = O.dsmc_apply(g, a, 16 * O.AMU, 1e9, dt, sv, rng)
        total += nc
        assert np.array_equal(a, golden["dm_one_%d" % call])
    assert total > 500
    rng = O.rng_mt(8082)
    sv = 1e-14
    for call in range(2):
        a, b, sv, nc, ng = O.dsmc_apply(g, a, 16 * O.AMU, 1e9, dt, sv, rng, p2=b, mass2=40 * O.AMU)
        assert nc > 100
        assert np.array_equal(a, golden["dm_two_a_%d" % call])
        assert np.array_equal(b, golden["dm_two_b_%d" % call])


def test_dsmc_mex_conserves_momentum_and_energy(golden):
    gr = golden["dm_grid"]
    g = _grid(gr)
    a, b = golden["dm_a_in"].reshape(-1, 7), golden["dm_b_in"].reshape(-1, 7)
    a2, b2, sv, nc, ng = O.dsmc_apply(g, a.reshape(-1), 16 * O.AMU, 1e9, gr[9], 1e-14,
                                      O.rng_philox(5, 0, 1), p2=b.reshape(-1), mass2=40 * O.AMU)
    a2, b2 = a2.reshape(-1, 7), b2.reshape(-1, 7)
    ma, mb = 16 * O.AMU, 40 * O.AMU
    p0 = ma * a[:, 3:6].sum(0) + mb * b[:, 3:6].sum(0)
    p1 = ma * a2[:, 3:6].sum(0) + mb * b2[:, 3:6].sum(0)
    e0 = ma * (a[:, 3:6] ** 2).sum() + mb * (b[:, 3:6] ** 2).sum()
    e1 = ma * (a2[:, 3:6] ** 2).sum() + mb * (b2[:, 3:6] ** 2).sum()
    assert nc > 100
    assert np.allclose(p0, p1, rtol=0, atol=1e-12 * np.sqrt(e0 * ma))
    assert e1 == pytest.approx(e0, rel=1e-12)


def test_sources_exact(G, golden):
    """ColdBeamSource (two samples, x- face) and WarmBeamSource (z+ face, a
    block in front of part of it) draw for draw (Source.cpp:3-185)"""
    L = _L(G)
    c = _xc(G)
    rect = ([c[0], c[1], G.xm[2]], [L[0] * 0.2, L[1] * 0.15, L[2] * 0.1], 0.0)
    ef = golden["ef_analytic"]
    rng = O.rng_mt(1234)
    cold = np.concatenate([O.source_sample(G, rng, -O.QE, O.ME, 1e4, 0, 2e5, 1e12, 0.0, 2e-9,
                                           ef=ef, rects=[rect]) for _ in range(2)])
    assert cold.size > 0
    assert np.array_equal(cold, golden["src_cold"])
    rng = O.rng_mt(4321)
    warm = O.source_sample(G, rng, O.QE, 16 * O.AMU, 5e3, 5, 3e3, 2e12, 5000.0, 2e-9, ef=ef,
                           rects=[rect])
    assert warm.size > 0
    assert np.array_equal(warm, golden["src_warm"])
