"""GPU parity of the object surfaces against the oracle (advanceElectrons'
removal, advanceNonElectron's ion recycling and neutral reflection,
Species.cpp:170-256, :333-338, :835-861; Rectangle::lineIntersect,
Object.cpp:271-317), on the reference fixture's three rectangles (two
plates and a block with faces, edges and corners in reach).

Particles carry distinct weights 2^33 + i 2^10 (exact in FP32), so a
device particle is matched to its oracle counterpart by weight.  Draws are
the same Philox stream on both sides; the oracle keeps emissions at the
entry point (the device's defined behaviour) and the survivors' order.
Tolerances: positions 1e-5 of the box, velocities 1e-5 of the largest;
a handful of particles within FP32 rounding of a face may decide
differently and are allowed."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TAG_SURF = 0x53524601
REL = 1e-5


@pytest.fixture(scope="module")
def pg():
    import pic_gpu

    return pic_gpu


def _setup(pg, golden):
    gr = golden["grid"]
    g = O.make_grid(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
    c = pg.Context(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
    r = golden["sf_rects"].reshape(-1, 6)
    rects = [(row[:3], row[3:], 0.0) for row in r]
    for row in r:
        c.add_rectangle(row[:3], 2 * row[3:], 0.0)
    ef = golden["sf_ef"]
    c.set_field(pg.FIELD_EF, ef)
    return g, c, rects, ef, golden["sf_dt"]


def _tagged(parts):
    """FP32-exact velocities, unique FP32-exact weights"""
    q = parts.reshape(-1, 7).copy()
    q[:, 3:6] = q[:, 3:6].astype(np.float32).astype(np.float64)
    q[:, 6] = 2.0 ** 33 + np.arange(len(q)) * 2.0 ** 10
    return q


def _off_faces(parts, golden):
    """drop the neutrals that the ion recycling left exactly on an object
    face: upload places particles by addParticle's containment test
    (Species.cpp:420-426, faces count as inside) and drops these, while the
    oracle would push them; particles emitted during a push stay on the
    device"""
    gr = golden["grid"]
    x0, xm, n = gr[3:6], gr[6:9], gr[:3]
    dx = (xm - x0) / (n - 1)
    q = parts.reshape(-1, 7)
    lc = (q[:, :3] - x0) / dx
    inside = np.zeros(len(q), bool)
    for row in golden["sf_rects"].reshape(-1, 6):
        lo, hi = (row[:3] - row[3:] - x0) / dx, (row[:3] + row[3:] - x0) / dx
        inside |= np.all((lc >= lo) & (lc <= hi), axis=1)
    return q[~inside].copy()


def _L(g):
    return np.array(g.dx[:]) * (np.array([g.ni, g.nj, g.nk]) - 1)


def _match(ref, got, g, allow):
    """match rows by weight; returns the number of unmatched rows"""
    ref, got = ref.reshape(-1, 7), got.reshape(-1, 7)
    rk, gk = {w: i for i, w in enumerate(ref[:, 6])}, {w: i for i, w in enumerate(got[:, 6])}
    common = sorted(set(rk) & set(gk))
    a = ref[[rk[w] for w in common]]
    b = got[[gk[w] for w in common]]
    vs = max(np.max(np.abs(a[:, 3:6])), 1.0)
    dpos = np.max(np.abs(a[:, :3] - b[:, :3]) / _L(g), axis=1)
    dvel = np.max(np.abs(a[:, 3:6] - b[:, 3:6]), axis=1) / vs
    bad = int(np.count_nonzero((dpos > REL) | (dvel > REL)))
    unmatched = len(rk) + len(gk) - 2 * len(common)
    assert bad + unmatched <= allow, (bad, unmatched, len(common))
    return len(common)


def test_electrons_absorbed_in_objects(pg, golden):
    g, c, rects, ef, dts = _setup(pg, golden)
    p = _tagged(golden["sf_ele_in"])
    se = c.species_add(O.ME, -O.QE, 2e4, len(p) + 64)
    c.upload(se, p)
    c.push(se, dts[0])
    ref = O.push(g, ef, p.reshape(-1), -O.QE, O.ME, dts[0], rects=rects, stable=True)
    assert ref.size < p.size  # some entered objects or left
    assert _match(ref, c.download(se), g, allow=2) > 0.9 * (ref.size // 7)


@pytest.mark.parametrize("fused", [False, True])
def test_ions_recycled_into_neutrals(pg, golden, fused):
    """ions ending in an object leave and emit int(w / mpw0_n + U)
    neutrals at the entry point with reflected velocities"""
    g, c, rects, ef, dts = _setup(pg, golden)
    ions = _tagged(golden["sf_ion_in"])
    mpw0_n = 3e9  # 2^33 / 3e9 = 2.86: two or three neutrals per ion hit
    sn = c.species_add(16 * O.AMU, 0.0, mpw0_n, 8 * len(ions))
    si = c.species_add(16 * O.AMU, O.QE, 2.5e10, len(ions) + 64)
    c.upload(si, ions)
    c.set_surface(si, pg.SURF_RECYCLE, sn, seed=91)
    if fused:
        c.push_sort(si, dts[1])
    else:
        c.push(si, dts[1])
    rng = O.rng_philox(91, 0, TAG_SURF)
    ref_i, ref_n, st = O.push_nonelectron(g, ef, ions.reshape(-1), O.QE, 16 * O.AMU, dts[1],
                                          rects, mpw0_n, rng, accept_on_surface=True,
                                          stable=True)
    assert st[0] > 100 and st[1] > 200 and st[2] == 0
    _match(ref_i, c.download(si), g, allow=3)
    got_n = c.download(sn).reshape(-1, 7)
    ref_n = ref_n.reshape(-1, 7)
    assert abs(len(got_n) - len(ref_n)) <= 9
    assert np.all(got_n[:, 6] == np.float32(mpw0_n))
    # emitted neutrals as a multiset: every column sorted
    m = min(len(got_n), len(ref_n))
    L = _L(g)
    vs = np.max(np.abs(ref_n[:, 3:6]))
    for col in range(6):
        a, b = np.sort(ref_n[:, col])[:m], np.sort(got_n[:, col])[:m]
        scale = L[col] if col < 3 else vs
        assert np.mean(np.abs(a - b) <= REL * scale) > 0.98, col


@pytest.mark.parametrize("fused", [False, True])
def test_neutrals_reflected(pg, golden, fused):
    g, c, rects, ef, dts = _setup(pg, golden)
    neu = _tagged(_off_faces(golden["sf_neu_mid"], golden))
    sn = c.species_add(16 * O.AMU, 0.0, 1e10, len(neu) + 64)
    c.upload(sn, neu)
    c.set_surface(sn, pg.SURF_REFLECT, seed=17)
    if fused:
        c.push_sort(sn, dts[2])
    else:
        c.push(sn, dts[2])
    rng = O.rng_philox(17, 0, TAG_SURF)
    ref, _, st = O.push_nonelectron(g, ef, neu.reshape(-1), 0.0, 16 * O.AMU, dts[2], rects, 1e10,
                                    rng, stable=True)
    assert st[0] > 100
    _match(ref, c.download(sn), g, allow=3)


def test_reflection_steps_are_keyed_by_push_count(pg, golden):
    """two pushes of a reflecting species draw from two streams (the
    species' push count is a key word): the second push of a copy started
    from the first push's output differs from repeating the first"""
    g, c, rects, ef, dts = _setup(pg, golden)
    neu = _tagged(_off_faces(golden["sf_neu_mid"], golden))
    a = c.species_add(16 * O.AMU, 0.0, 1e10, len(neu) + 64)
    c.upload(a, neu)
    c.set_surface(a, pg.SURF_REFLECT, seed=5)
    c.push(a, dts[2])
    rng = O.rng_philox(5, 0, TAG_SURF)
    r1, _, _ = O.push_nonelectron(g, ef, neu.reshape(-1), 0.0, 16 * O.AMU, dts[2], rects, 1e10,
                                  rng, stable=True)
    # compacted and cell-sorted: the second push's particle indices are the
    # device's order, which the oracle then takes as its input order
    c.sort(a)
    mid = c.download(a)
    _match(r1, mid, g, allow=3)
    c.push(a, dts[2])
    rng = O.rng_philox(5, 1, TAG_SURF)
    r2, _, _ = O.push_nonelectron(g, ef, mid, 0.0, 16 * O.AMU, dts[2], rects, 1e10, rng,
                                  stable=True)
    _match(r2, c.download(a), g, allow=3)
    # the same input pushed with the first push's stream differs
    rng = O.rng_philox(5, 0, TAG_SURF)
    r2x, _, _ = O.push_nonelectron(g, ef, mid, 0.0, 16 * O.AMU, dts[2], rects, 1e10, rng,
                                   stable=True)
    with pytest.raises(AssertionError):
        _match(r2x, c.download(a), g, allow=3)
