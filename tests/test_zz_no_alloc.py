"""pic_deposit_density_noalloc (the density with its corner sums in a fixed
device array, no allocation) against pic_deposit(PIC_DEPOSIT_DENSITY)
through the C ABI: the same kernels on the same sorted particles give the
same density bit for bit, on two calls in a row, the second on a smaller
input (a smaller grid and fewer particles, so the first call's contents of
the shared array must not leak into it), then again on the first context
with fewer particles; and the entry point's refusals (unsorted species, a
grid beyond its 2^21 cells)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME, QE = 9.10938215e-31, 1.602176565e-19
PIC_ERR_CAPACITY, PIC_ERR_STATE = -3, -4


@pytest.fixture(scope="module")
def pg():
    import pic_gpu

    return pic_gpu


def _parts(rng, n, L):
    p = np.zeros((n, 7))
    p[:, :3] = rng.uniform(0.01, 0.99, (n, 3)) * np.asarray(L)
    p[:, 3:6] = rng.normal(0.0, 1e5, (n, 3))
    p[:, 6] = rng.uniform(0.5, 2.0, n) * 1e6
    return p.reshape(-1)


def _both(pg, c, sid):
    """the no-alloc density, then pic_deposit's, of the sorted species"""
    c.sort(sid)
    rc = c.deposit_density_noalloc(sid)
    assert rc == pg.PIC_OK, rc
    a = c.species_field(sid, pg.SFIELD_DEN)
    c.deposit(sid, pg.DEPOSIT_DENSITY)
    b = c.species_field(sid, pg.SFIELD_DEN)
    return a, b


def test_noalloc_density_matches_pic_deposit(pg):
    rng = np.random.default_rng(3)
    print("no-alloc density: 65^3 nodes, 400k particles", flush=True)
    L1 = [0.64, 0.64, 0.64]
    c1 = pg.Context(65, 65, 65, [0, 0, 0], L1)
    s1 = c1.species_add(ME, -QE, 1e6, 500000)
    c1.upload(s1, _parts(rng, 400000, L1))
    a, b = _both(pg, c1, s1)
    assert np.array_equal(a, b) and np.sum(a) > 0
    print("no-alloc density: 33x29x17 nodes, 100k particles", flush=True)
    L2 = [0.32, 0.28, 0.16]
    c2 = pg.Context(33, 29, 17, [0, 0, 0], L2)
    s2 = c2.species_add(ME, -QE, 1e6, 200000)
    c2.upload(s2, _parts(rng, 100000, L2))
    a2, b2 = _both(pg, c2, s2)
    assert np.array_equal(a2, b2) and np.sum(a2) > 0
    print("no-alloc density: 65^3 nodes again, 50k particles", flush=True)
    c1.upload(s1, _parts(rng, 50000, L1))
    a3, b3 = _both(pg, c1, s1)
    assert np.array_equal(a3, b3) and np.sum(a3) > 0
    c2.close()
    c1.close()


def test_noalloc_density_refusals(pg):
    rng = np.random.default_rng(4)
    L = [0.16, 0.16, 0.16]
    c = pg.Context(17, 17, 17, [0, 0, 0], L)
    s = c.species_add(ME, -QE, 1e6, 20000)
    c.upload(s, _parts(rng, 10000, L))  # uploaded particles are not sorted yet
    assert c.deposit_density_noalloc(s) == PIC_ERR_STATE
    c.close()
    # 130 x 129 x 129 nodes: 129 * 128 * 128 cells > 2^21
    big = pg.Context(130, 129, 129, [0, 0, 0], [1.29, 1.28, 1.28])
    sb = big.species_add(ME, -QE, 1e6, 1000)
    big.sort(sb)
    assert big.deposit_density_noalloc(sb) == PIC_ERR_CAPACITY
    big.close()
