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
