"""DSMC (MC_MEX_Ionization::apply, Interactions.cpp:567-762) on the GPU
against the oracle: the electron-heavier branch with ionization and
neutral removal under the same Philox draws, and seeded ensembles of the
device (Philox) against the oracle's restatement of the reference with its
own mt19937 streams, compared in their ionization, collision and electron
energy moments within statistical error."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TAG_MCC = 0x4D434301


@pytest.fixture(scope="module")
def pg():
    import pic_gpu

    return pic_gpu


def _rows_sorted(p):
    q = p.reshape(-1, 7)
    return q[np.lexsort(np.round(q[:, [6, 5, 4, 3, 2, 1, 0]], 7).T)]


def _same_multiset(a, b, L, rel=1e-5):
    a, b = _rows_sorted(a), _rows_sorted(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if len(a) == 0:
        return
    vs = max(np.max(np.abs(a[:, 3:6])), 1.0)
    assert np.max(np.abs(a[:, :3] - b[:, :3]) / L) < rel
    assert np.max(np.abs(a[:, 3:6] - b[:, 3:6])) < rel * vs
    assert np.allclose(a[:, 6], b[:, 6], rtol=1e-6)


def test_electron_heavier_branch_matches_oracle(pg, sigma_table_path):
    """Electrons ten times heavier than the neutrals and energetic enough to
    ionize: every collision splits the electron (Interactions.cpp:705-726),
    ionizations remove the neutral from the cell's pool (the pool's last
    neutral takes its slot), and the next step sorts the neutrals whole.
    Two steps, each compared with the oracle under the same draws."""
    n, dx = 9, 1e-4
    L = dx * (n - 1)
    c = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    rng = np.random.default_rng(11)

    def parts(count, vth, w):
        p = np.zeros((count, 7))
        p[:, :3] = rng.uniform(0.01 * L, 0.99 * L, (count, 3))
        p[:, 3:6] = rng.normal(0, vth, (count, 3)).astype(np.float32)
        p[:, 6] = w
        return p.reshape(-1)

    sn = c.species_add(16 * O.AMU, 0.0, 1e4, 40000)
    si = c.species_add(16 * O.AMU, O.QE, 1.0, 40000)
    se = c.species_add(O.ME, -O.QE, 1.0, 60000)
    c.upload(sn, parts(30 * 512, 300.0, 1.0))
    c.upload(se, parts(10 * 512, 3e6, 10.0))  # ~75 eV mean: well above E_ion
    sE, sV = O.load_sigma_table(sigma_table_path)
    E_ion = 1313.9 * 1000 / 6.02214076e23
    # W sigma v_max starts at its working value (1.5e-12 for these weights
    # and energies), so most groups collide
    dt, wsv = 5e-3, 1.5e-12
    mid = c.mcc_create(sn, si, se, E_ion, sE, sV, wsv_max_init=wsv)
    prm = O.mcc_params(16 * O.AMU, O.ME, E_ion, sE, sV, dt, g.dx[0] * g.dx[1] * g.dx[2],
                       pool_append=True)
    for step in range(2):
        c.sort(sn)
        c.sort(se)
        neu, ele, ion = c.download(sn), c.download(se), c.download(si)
        st = c.mcc_apply(mid, dt, 13, step)
        on, oi, oe, wsv_o, ost = O.mcc_apply(g, prm, wsv, O.rng_philox(13, step, TAG_MCC), neu,
                                             ion, ele)
        print("step", step, st, ost.collisions, ost.ionizations, flush=True)
        assert st["groups"] == ost.groups
        assert st["collisions"] == ost.collisions
        assert st["ionizations"] == ost.ionizations > 100
        assert st["wsv_max"] == pytest.approx(wsv_o, rel=1e-6)
        for sid, ref in ((sn, on), (si, oi), (se, oe)):
            _same_multiset(ref, c.download(sid), L)
        assert c.count(sn) < neu.size // 7  # ionised neutrals left
        wsv = wsv_o


def test_appended_neutrals_indexed_then_joined_match_oracle(pg, sigma_table_path):
    """Neutrals ten times heavier than the electrons: collisions split them
    (Interactions.cpp:684-703) and the split neutrals are appended, as are
    the ionizations' electrons.  Neutrals and electrons are sorted once;
    every later step indexes their appended parts by cell and joins the
    neutrals' with the sorted part once it outgrows a fraction of it.  Each
    step is compared with the oracle under the same draws on the downloaded
    arrays, which list the appended parts cell by cell in pick-pool order."""
    n, dx = 9, 1e-4
    L = dx * (n - 1)
    c = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    rng = np.random.default_rng(12)

    def parts(count, vth, w):
        p = np.zeros((count, 7))
        p[:, :3] = rng.uniform(0.01 * L, 0.99 * L, (count, 3))
        p[:, 3:6] = rng.normal(0, vth, (count, 3)).astype(np.float32)
        p[:, 6] = w
        return p.reshape(-1)

    n0 = 30 * 512
    # (the species' mpw0 as MC_MEX_Ionization requires, Interactions.cpp:479-489; the
    # particles' own weights are uploaded)
    sn = c.species_add(16 * O.AMU, 0.0, 100.0, 200000)
    si = c.species_add(16 * O.AMU, O.QE, 1.0, 200000)
    se = c.species_add(O.ME, -O.QE, 1.0, 200000)
    c.upload(sn, parts(n0, 300.0, 10.0))
    c.upload(se, parts(10 * 512, 3e6, 1.0))
    sE, sV = O.load_sigma_table(sigma_table_path)
    E_ion = 1313.9 * 1000 / 6.02214076e23
    dt, wsv = 2e-2, 1.5e-12
    mid = c.mcc_create(sn, si, se, E_ion, sE, sV, wsv_max_init=wsv)
    prm = O.mcc_params(16 * O.AMU, O.ME, E_ion, sE, sV, dt, g.dx[0] * g.dx[1] * g.dx[2],
                       pool_append=True)

    def sorted_prefix(p):
        q = p.reshape(-1, 7)
        ijk = np.clip(np.floor(q[:, :3] / dx + 1e-8).astype(np.int64), 0, n - 2)
        key = (ijk[:, 2] * (n - 1) + ijk[:, 1]) * (n - 1) + ijk[:, 0]
        drop = np.nonzero(np.diff(key) < 0)[0]
        return len(key) if len(drop) == 0 else int(drop[0]) + 1

    c.sort(sn)
    c.sort(se)
    prefixes, counts = [], []
    for step in range(6):
        neu, ele, ion = c.download(sn), c.download(se), c.download(si)
        prefixes.append(sorted_prefix(neu))
        counts.append(neu.size // 7)
        st = c.mcc_apply(mid, dt, 17, step)
        on, oi, oe, wsv_o, ost = O.mcc_apply(g, prm, wsv, O.rng_philox(17, step, TAG_MCC), neu,
                                             ion, ele)
        print("step", step, "neutrals", counts[-1], "sorted prefix", prefixes[-1], st,
              ost.collisions, flush=True)
        assert st["groups"] == ost.groups
        assert st["collisions"] == ost.collisions > 1000
        assert st["ionizations"] == ost.ionizations
        assert st["pool_overflow"] == 0
        assert st["wsv_max"] == pytest.approx(wsv_o, rel=1e-6)
        for sid, ref in ((sn, on), (si, oi), (se, oe)):
            _same_multiset(ref, c.download(sid), L)
        wsv = wsv_o
    # an appended part was indexed (the arrays unsorted past the sorted part)
    # and later joined (the sorted part grew past the first sort's)
    assert any(p < cnt for p, cnt in zip(prefixes[1:], counts[1:]))
    assert max(prefixes) > 1.2 * n0, prefixes


def test_species_operations_after_indexed_births(pg, sigma_table_path):
    """After DSMC steps leave births indexed behind the sorted parts, every
    other operation takes the species as they are: the density of the
    neutrals equals the oracle's of their downloaded particles, a merge of
    the electrons keeps their weight, a push of the electrons sorts them,
    and the next DSMC step still equals the oracle's under the same draws."""
    n, dx = 9, 1e-4
    L = dx * (n - 1)
    c = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    g = O.make_grid(n, n, n, [0, 0, 0], [L, L, L])
    rng = np.random.default_rng(31)

    def parts(count, vth, w):
        p = np.zeros((count, 7))
        p[:, :3] = rng.uniform(0.01 * L, 0.99 * L, (count, 3))
        p[:, 3:6] = rng.normal(0, vth, (count, 3)).astype(np.float32)
        p[:, 6] = w
        return p.reshape(-1)

    sn = c.species_add(16 * O.AMU, 0.0, 100.0, 200000)
    si = c.species_add(16 * O.AMU, O.QE, 1.0, 200000)
    se = c.species_add(O.ME, -O.QE, 1.0, 200000)
    c.upload(sn, parts(30 * 512, 300.0, 10.0))
    c.upload(se, parts(10 * 512, 3e6, 1.0))
    sE, sV = O.load_sigma_table(sigma_table_path)
    E_ion = 1313.9 * 1000 / 6.02214076e23
    dt, wsv = 2e-2, 1.5e-12
    mid = c.mcc_create(sn, si, se, E_ion, sE, sV, wsv_max_init=wsv)
    prm = O.mcc_params(16 * O.AMU, O.ME, E_ion, sE, sV, dt, g.dx[0] * g.dx[1] * g.dx[2],
                       pool_append=True)
    c.sort(sn)
    c.sort(se)
    for step in range(2):
        wsv = c.mcc_apply(mid, dt, 5, step)["wsv_max"]
    # the neutrals' density from the arrays as they are
    neu = c.download(sn)
    c.deposit(sn, pg.DEPOSIT_DENSITY)
    ref = O.number_density(g, neu)
    assert np.max(np.abs(c.species_field(sn, pg.SFIELD_DEN) - ref)) <= 1e-5 * np.max(ref)
    # a merge of the electrons keeps their weight
    w0 = np.sum(_rows_sorted(c.download(se))[:, 6])
    c.merge(se, seed=3)
    assert np.sum(_rows_sorted(c.download(se))[:, 6]) == pytest.approx(w0, rel=1e-6)
    # a push of the electrons (no field) leaves them sorted by cell
    c.push_sort(se, 1e-13)
    # the next DSMC step from these states equals the oracle's
    neu, ele, ion = c.download(sn), c.download(se), c.download(si)
    st = c.mcc_apply(mid, dt, 5, 2)
    on, oi, oe, wsv_o, ost = O.mcc_apply(g, prm, wsv, O.rng_philox(5, 2, TAG_MCC), neu, ion, ele)
    assert st["groups"] == ost.groups
    assert st["collisions"] == ost.collisions > 100
    assert st["ionizations"] == ost.ionizations
    for sid, want in ((sn, on), (si, oi), (se, oe)):
        _same_multiset(want, c.download(sid), L)


def _ensemble_stats(ele, ion_iz, coll):
    e = ele.reshape(-1, 7)
    ke = np.sum(e[:, 6] * np.sum(e[:, 3:6] ** 2, axis=1)) / np.sum(e[:, 6])
    return [ion_iz, coll, 0.5 * O.ME * ke / O.QE, len(e)]


def test_seeded_ensemble_matches_reference_moments(pg, golden, sigma_table_path):
    """K seeds of three apply calls on the reference fixture's dense atomic-O
    box (~30 neutrals and ~10 electrons of 60 eV per cell): the device's
    Philox streams against the oracle's restatement of the reference (split
    neutrals join the pool) with mt19937 streams.  The ensemble means of
    ionizations, collisions, mean electron energy and electron count agree
    within 4 standard errors."""
    gr = golden["mcc_grid"]
    g = O.make_grid(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
    dt = gr[9]
    sE, sV = O.load_sigma_table(sigma_table_path)
    E_ion = 1313.9 * 1000 / 6.02214076e23
    neu = golden["mcc_neu_in"].reshape(-1, 7).copy()
    ele = golden["mcc_ele_in"].reshape(-1, 7).copy()
    for p in (neu, ele):
        p[:, 3:7] = p[:, 3:7].astype(np.float32).astype(np.float64)
    prm = O.mcc_params(16 * O.AMU, O.ME, E_ion, sE, sV, dt, g.dx[0] * g.dx[1] * g.dx[2],
                       pool_append=True)
    K = 24
    dev, ref = [], []
    for seed in range(1, K + 1):
        c = pg.Context(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
        sn = c.species_add(16 * O.AMU, 0.0, 5e11, 8 * len(neu))
        si = c.species_add(16 * O.AMU, O.QE, 100, 40000)
        se = c.species_add(O.ME, -O.QE, 100, 8 * len(ele))
        c.upload(sn, neu)
        c.upload(se, ele)
        mid = c.mcc_create(sn, si, se, E_ion, sE, sV)
        iz = coll = 0
        for call in range(3):
            st = c.mcc_apply(mid, dt, seed, call)
            assert st["pool_overflow"] == 0
            iz += st["ionizations"]
            coll += st["collisions"]
        dev.append(_ensemble_stats(c.download(se), iz, coll))
        c.close()
        on, oi, oe = neu.reshape(-1), np.zeros(0), ele.reshape(-1)
        wsv = golden["mcc_wsv0"][0]
        rng = O.rng_mt(1000 + seed)
        iz = coll = 0
        for call in range(3):
            on, oi, oe, wsv, st = O.mcc_apply(g, prm, wsv, rng, on, oi, oe)
            iz += st.ionizations
            coll += st.collisions
        ref.append(_ensemble_stats(oe, iz, coll))
    dev, ref = np.array(dev, float), np.array(ref, float)
    md, mr = dev.mean(0), ref.mean(0)
    se = np.sqrt(dev.var(0, ddof=1) / K + ref.var(0, ddof=1) / K)
    for name, a, b, s in zip(("ionizations", "collisions", "mean e- energy [eV]", "electrons"),
                             md, mr, se):
        print("%-20s device %.6g  reference %.6g  (%.2f standard errors)" % (name, a, b,
                                                                             abs(a - b) / s),
              flush=True)
        assert abs(a - b) <= 4 * s, name
