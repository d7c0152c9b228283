"""ctypes binding of the FP64 CPU oracle (oracle/pic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by bench.py's cpu_baseline leg.  The product library never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("x0", C.c_double * 3), ("xm", C.c_double * 3),
                ("dx", C.c_double * 3), ("inv_dx", C.c_double * 3)]


class Rect(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("half", C.c_double * 3), ("phi", C.c_double)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int), ("mt", MT), ("key", C.c_uint32 * 2), ("tag", C.c_uint32),
                ("a", C.c_uint32), ("b", C.c_uint32)]


class MccParams(C.Structure):
    _fields_ = [("m_neutral", C.c_double), ("m_electron", C.c_double), ("E_ion_J", C.c_double),
                ("c0", C.c_double), ("c1", C.c_double), ("c2", C.c_double), ("B_inc", C.c_double),
                ("sig_E", C.POINTER(C.c_double)), ("sig_v", C.POINTER(C.c_double)),
                ("nsig", C.c_int), ("dt", C.c_double), ("dV", C.c_double),
                ("pool_append", C.c_int)]


class MccStats(C.Structure):
    _fields_ = [("collisions", C.c_longlong), ("ionizations", C.c_longlong),
                ("groups", C.c_longlong), ("wsv_max_step", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        G = C.POINTER(Grid)
        R = C.POINTER(Rect)
        L.orc_grid_init.argtypes = [G, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_node_volumes.argtypes = [G, _dp]
        L.orc_object_id.argtypes = [G, R, C.c_int, _ip, _ip, _dp]
        L.orc_dirichlet_walls.argtypes = [G, _ip]
        L.orc_gather3.argtypes = [G, _dp, _dp, _dp]
        L.orc_gather1.argtypes = [G, _dp, _dp]
        L.orc_gather1.restype = C.c_double
        L.orc_push.argtypes = [G, _dp, _dp, C.c_size_t, C.c_double, C.c_double, C.c_double,
                               R, C.c_int, C.c_int]
        L.orc_push.restype = C.c_size_t
        L.orc_push_threaded.argtypes = [G, _dp, _dp, C.c_size_t, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int]
        L.orc_push_threaded.restype = C.c_size_t
        L.orc_number_density.argtypes = [G, _dp, C.c_size_t, _dp, _dp]
        L.orc_charge_accumulate.argtypes = [C.c_size_t, _dp, C.c_double, _dp]
        L.orc_sample_moments.argtypes = [G, _dp, C.c_size_t, _dp, _dp, _dp, _dp, _dp]
        L.orc_gas_properties.argtypes = [G, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_macro_count.argtypes = [G, _dp, C.c_size_t, _dp]
        L.orc_sort_indexes.argtypes = [G, _dp, C.c_size_t, _ip, _ip]
        L.orc_species_sums.argtypes = [G, _dp, C.c_size_t, C.c_double, C.c_double, C.c_void_p,
                                       _dp]
        L.orc_field_energy.argtypes = [G, _dp, _dp]
        L.orc_field_energy.restype = C.c_double
        L.orc_efield.argtypes = [G, _dp, _dp]
        L.orc_sor_dirichlet.argtypes = [G, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                        C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_gs_v3.argtypes = [G, _ip, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.c_double, C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
        L.orc_nrpcg.argtypes = [G, _ip, _ip, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_residual_dirichlet.argtypes = [G, _dp, _dp, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
        L.orc_dense_poisson.argtypes = [G, _dp, _dp]
        L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
        L.orc_mt_next.argtypes = [C.POINTER(MT)]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_mt_uniform.argtypes = [C.POINTER(MT)]
        L.orc_mt_uniform.restype = C.c_double
        L.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]
        L.orc_rng_mt.argtypes = [C.POINTER(Rng), C.c_uint32]
        L.orc_rng_philox.argtypes = [C.POINTER(Rng), C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_rng_draw.argtypes = [C.POINTER(Rng), C.c_int]
        L.orc_rng_draw.restype = C.c_double
        L.orc_load_box_thermal.argtypes = [G, R, C.c_int, C.c_void_p, C.c_double, C.c_double,
                                           C.c_double, _dp, _dp, C.c_double, C.c_double,
                                           C.c_double, C.POINTER(Rng), C.c_int, C.c_size_t, _dp]
        L.orc_load_box_thermal.restype = C.c_size_t
        L.orc_sigma_el.argtypes = [C.POINTER(MccParams), C.c_double]
        L.orc_sigma_el.restype = C.c_double
        L.orc_sigma_iz.argtypes = [C.POINTER(MccParams), C.c_double]
        L.orc_sigma_iz.restype = C.c_double
        L.orc_mcc_collide.argtypes = [C.POINTER(MccParams), C.POINTER(Rng), _dp, _dp, _dp,
                                      C.c_double]
        L.orc_mcc_apply.argtypes = [G, C.POINTER(MccParams), C.POINTER(C.c_double),
                                    C.POINTER(Rng), _dp, C.POINTER(C.c_size_t), C.c_size_t,
                                    _dp, C.POINTER(C.c_size_t), C.c_size_t, _dp,
                                    C.POINTER(C.c_size_t), C.c_size_t, C.POINTER(MccStats)]
        _i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_rect_line_intersect.argtypes = [R, _dp, _dp, C.POINTER(C.c_double), _dp, _dp]
        L.orc_push_nonelectron.argtypes = [G, _dp, _dp, C.c_size_t, C.c_double, C.c_double,
                                           C.c_double, R, C.c_int, _dp, C.POINTER(C.c_size_t),
                                           C.c_size_t, C.c_double, C.POINTER(Rng), C.c_int,
                                           C.c_int, _i64]
        L.orc_push_nonelectron.restype = C.c_size_t
        L.orc_merge.argtypes = [G, C.c_void_p, _dp, C.POINTER(C.c_size_t), C.c_size_t,
                                C.c_double, C.c_double, C.c_double, C.POINTER(Rng), C.c_int, _i64]
        L.orc_dsmc_apply.argtypes = [G, _dp, C.c_size_t, C.c_double, C.c_void_p, C.c_size_t,
                                     C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(Rng), C.POINTER(C.c_longlong)]
        L.orc_dsmc_apply.restype = C.c_longlong
        L.orc_source_sample.argtypes = [G, R, C.c_int, C.c_void_p, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.POINTER(Rng), _dp, C.c_size_t]
        L.orc_source_sample.restype = C.c_size_t
        _lib = L
    return _lib


# physical constants (reference all.h:13-23)
EPS0 = 8.85418782e-12
QE = 1.602176565e-19
AMU = 1.660538921e-27
ME = 9.10938215e-31
KB = 1.380648e-23
PI_REF = 3.141592653


def make_grid(ni, nj, nk, x0, xm):
    g = Grid()
    lib().orc_grid_init(C.byref(g), ni, nj, nk, np.asarray(x0, np.float64),
                        np.asarray(xm, np.float64))
    return g


def rects_arg(rects):
    if not rects:
        return None, 0
    arr = (Rect * len(rects))()
    for n, (c, half, phi) in enumerate(rects):
        arr[n].c[:] = list(c)
        arr[n].half[:] = list(half)
        arr[n].phi = phi
    return arr, len(rects)


def nn(g):
    return g.ni * g.nj * g.nk


def node_volumes(g):
    v = np.zeros(nn(g))
    lib().orc_node_volumes(C.byref(g), v)
    return v


def efield(g, phi):
    ef = np.zeros(3 * nn(g))
    lib().orc_efield(C.byref(g), np.ascontiguousarray(phi, np.float64), ef)
    return ef


def push(g, ef, parts, charge, mass, dt, rects=None, stable=False):
    p = np.ascontiguousarray(parts, np.float64).copy()
    ra, nr = rects_arg(rects)
    n = lib().orc_push(C.byref(g), np.ascontiguousarray(ef, np.float64), p, p.size // 7,
                       charge, mass, dt, ra, nr, int(stable))
    return p[: 7 * n]


def push_threaded(g, ef, parts, charge, mass, dt, nthreads, remove=False):
    """the reference's threaded electron push in place; returns the count
    (remove=False: as the reference, which removes nothing)"""
    n = lib().orc_push_threaded(C.byref(g), ef, parts, parts.size // 7, charge, mass, dt,
                                nthreads, int(remove))
    return n


def number_density(g, parts, vol=None):
    if vol is None:
        vol = node_volumes(g)
    den = np.zeros(nn(g))
    p = np.ascontiguousarray(parts, np.float64)
    lib().orc_number_density(C.byref(g), p, p.size // 7, vol, den)
    return den


def sample_moments(g, parts, acc=None):
    n = nn(g)
    if acc is None:
        acc = [np.zeros(n), np.zeros(3 * n), np.zeros(n), np.zeros(n), np.zeros(n)]
    p = np.ascontiguousarray(parts, np.float64)
    lib().orc_sample_moments(C.byref(g), p, p.size // 7, *acc)
    return acc


def gas_properties(g, mass, acc):
    n = nn(g)
    vel = np.zeros(3 * n)
    T = np.zeros(n)
    lib().orc_gas_properties(C.byref(g), mass, acc[0], acc[1], acc[2], acc[3], acc[4], vel, T)
    return vel, T


def macro_count(g, parts):
    out = np.zeros((g.ni - 1) * (g.nj - 1) * (g.nk - 1))
    p = np.ascontiguousarray(parts, np.float64)
    lib().orc_macro_count(C.byref(g), p, p.size // 7, out)
    return out


def sort_indexes(g, parts):
    nc = (g.ni - 1) * (g.nj - 1) * (g.nk - 1)
    off = np.zeros(nc + 1, np.int32)
    p = np.ascontiguousarray(parts, np.float64)
    idx = np.zeros(max(1, p.size // 7), np.int32)
    lib().orc_sort_indexes(C.byref(g), p, p.size // 7, off, idx)
    return off, idx[: p.size // 7]


def species_sums(g, parts, mass, charge, phi=None):
    out = np.zeros(6)
    p = np.ascontiguousarray(parts, np.float64)
    ph = None
    if phi is not None:
        phi = np.ascontiguousarray(phi, np.float64)
        ph = phi.ctypes.data_as(C.c_void_p)
    lib().orc_species_sums(C.byref(g), p, p.size // 7, mass, charge, ph, out)
    return out


def field_energy(g, ef, vol=None):
    if vol is None:
        vol = node_volumes(g)
    return lib().orc_field_energy(C.byref(g), np.ascontiguousarray(ef, np.float64), vol)


def sor_dirichlet(g, phi, rho, max_it, tol, omega=1.4):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    it = C.c_int()
    l2 = C.c_double()
    ok = lib().orc_sor_dirichlet(C.byref(g), phi, np.ascontiguousarray(rho, np.float64), max_it,
                                 tol, omega, C.byref(it), C.byref(l2))
    return phi, bool(ok), it.value, l2.value


def gs_v3(g, object_id, phi, rho, phi0, n0, Te0, max_it, tol, omega=1.4):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    it = C.c_int()
    l2 = C.c_double()
    ok = lib().orc_gs_v3(C.byref(g), np.ascontiguousarray(object_id, np.int32), phi,
                         np.ascontiguousarray(rho, np.float64), phi0, n0, Te0, max_it, tol,
                         omega, C.byref(it), C.byref(l2))
    return phi, bool(ok), it.value, l2.value


def nrpcg(g, object_id, node_type, phi, rho, phi0, n0, Te0, pcg_max_it, tol, axis_swap=True):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    nit = C.c_int()
    pit = C.c_int()
    l2 = C.c_double()
    ok = lib().orc_nrpcg(C.byref(g), np.ascontiguousarray(object_id, np.int32),
                         np.ascontiguousarray(node_type, np.int32), phi,
                         np.ascontiguousarray(rho, np.float64), phi0, n0, Te0, pcg_max_it, tol,
                         int(axis_swap), C.byref(nit), C.byref(pit), C.byref(l2))
    return phi, ok, nit.value, pit.value, l2.value


def residual_dirichlet(g, phi, rho):
    r = C.c_double()
    b = C.c_double()
    lib().orc_residual_dirichlet(C.byref(g), np.ascontiguousarray(phi, np.float64),
                                 np.ascontiguousarray(rho, np.float64), C.byref(r), C.byref(b))
    return r.value, b.value


def dense_poisson(g, phi, rho):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    rc = lib().orc_dense_poisson(C.byref(g), phi, np.ascontiguousarray(rho, np.float64))
    assert rc == 0
    return phi


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return list(o)


def rng_mt(seed):
    r = Rng()
    lib().orc_rng_mt(C.byref(r), seed)
    return r


def rng_philox(seed, seq, tag):
    r = Rng()
    lib().orc_rng_philox(C.byref(r), seed, seq, tag)
    return r


def load_box_thermal(g, rng, charge, mass, rewind_dt, xc, sides, num_den, T, mpw0, mode=0,
                     ef=None, rects=None, max_out=None):
    box_vol = sides[0] * sides[1] * sides[2]
    nmax = int(num_den * box_vol / mpw0) + 1 if max_out is None else max_out
    out = np.zeros(7 * max(nmax, 1))
    ra, nr = rects_arg(rects)
    efp = None
    if ef is not None:
        ef = np.ascontiguousarray(ef, np.float64)
        efp = ef.ctypes.data_as(C.c_void_p)
    n = lib().orc_load_box_thermal(C.byref(g), ra, nr, efp, charge, mass, rewind_dt,
                                   np.asarray(xc, np.float64), np.asarray(sides, np.float64),
                                   num_den, T, mpw0, C.byref(rng), mode, nmax, out)
    return out[: 7 * n]


def load_sigma_table(path):
    t = np.loadtxt(path)
    return np.ascontiguousarray(t[:, 0]), np.ascontiguousarray(t[:, 1])


def mcc_params(m_neutral, m_electron, E_ion_J, sig_E, sig_v, dt, dV, pool_append=True,
               c=(1.015e-18, 9.793, 61.81), B_inc=10.0):
    p = MccParams()
    p.m_neutral = m_neutral
    p.m_electron = m_electron
    p.E_ion_J = E_ion_J
    p.c0, p.c1, p.c2 = c
    p.B_inc = B_inc
    p._keep = (sig_E, sig_v)
    p.sig_E = sig_E.ctypes.data_as(C.POINTER(C.c_double))
    p.sig_v = sig_v.ctypes.data_as(C.POINTER(C.c_double))
    p.nsig = sig_E.size
    p.dt = dt
    p.dV = dV
    p.pool_append = int(pool_append)
    return p


def mcc_apply(g, params, wsv_max, rng, neu, ion, ele, extra=None):
    """One MC_MEX_Ionization::apply; returns (neu, ion, ele, wsv_max, stats)."""
    nn_, ni_, ne_ = neu.size // 7, ion.size // 7, ele.size // 7
    if extra is None:
        extra = 4 * (nn_ + ne_) + 64
    bn = np.zeros(7 * (nn_ + extra))
    bi = np.zeros(7 * (ni_ + extra))
    be = np.zeros(7 * (ne_ + extra))
    bn[: neu.size] = neu
    bi[: ion.size] = ion
    be[: ele.size] = ele
    cn, ci, ce = C.c_size_t(nn_), C.c_size_t(ni_), C.c_size_t(ne_)
    w = C.c_double(wsv_max)
    st = MccStats()
    rc = lib().orc_mcc_apply(C.byref(g), C.byref(params), C.byref(w), C.byref(rng), bn,
                             C.byref(cn), nn_ + extra, bi, C.byref(ci), ni_ + extra, be,
                             C.byref(ce), ne_ + extra, C.byref(st))
    assert rc == 0, "oracle capacity exceeded"
    return bn[: 7 * cn.value], bi[: 7 * ci.value], be[: 7 * ce.value], w.value, st


def rect_line_intersect(rect, x1, x2):
    """Rectangle::lineIntersect: (t_entry, entry point, normal)"""
    ra, _ = rects_arg([rect])
    t = C.c_double()
    pos = np.zeros(3)
    n = np.zeros(3)
    lib().orc_rect_line_intersect(ra, np.asarray(x1, np.float64), np.asarray(x2, np.float64),
                                  C.byref(t), pos, n)
    return t.value, pos, n


def push_nonelectron(g, ef, parts, charge, mass, dt, rects, mpw0_n, rng, neu=None,
                     accept_on_surface=False, stable=False, extra=None):
    """advanceNonElectron with object surfaces; returns (survivors, neutrals
    with the emitted ones appended, stats[hits, emitted, rejected, removed])"""
    p = np.ascontiguousarray(parts, np.float64).copy()
    neu = np.zeros(0) if neu is None else np.ascontiguousarray(neu, np.float64)
    n0 = neu.size // 7
    if extra is None:
        extra = 8 * (p.size // 7) + 64
    buf = np.zeros(7 * (n0 + extra))
    buf[: neu.size] = neu
    cn = C.c_size_t(n0)
    ra, nr = rects_arg(rects)
    st = np.zeros(4, np.int64)
    n = lib().orc_push_nonelectron(C.byref(g), np.ascontiguousarray(ef, np.float64), p,
                                   p.size // 7, charge, mass, dt, ra, nr, buf, C.byref(cn),
                                   n0 + extra, mpw0_n, C.byref(rng), int(accept_on_surface),
                                   int(stable), st)
    assert n != C.c_size_t(-1).value, "oracle neutral capacity exceeded"
    return p[: 7 * n], buf[: 7 * cn.value], st


def merge(g, parts, charge, mass, dt, rng, ef=None, gpu_mode=False, extra=None):
    """Species::merge; returns (particles, stats[bins merged, particles in them])"""
    p = np.ascontiguousarray(parts, np.float64)
    n0 = p.size // 7
    if extra is None:
        extra = n0 + 64
    buf = np.zeros(7 * (n0 + extra))
    buf[: p.size] = p
    cn = C.c_size_t(n0)
    efp = None
    if ef is not None:
        ef = np.ascontiguousarray(ef, np.float64)
        efp = ef.ctypes.data_as(C.c_void_p)
    st = np.zeros(2, np.int64)
    rc = lib().orc_merge(C.byref(g), efp, buf, C.byref(cn), n0 + extra, charge, mass, dt,
                         C.byref(rng), int(gpu_mode), st)
    assert rc == 0, "oracle merge capacity exceeded"
    return buf[: 7 * cn.value], st


def dsmc_apply(g, p1, mass1, mpw0, dt, svmax, rng, p2=None, mass2=0.0):
    """DSMC_MEX::apply on copies; returns (p1, p2, svmax, collisions, groups)"""
    a = np.ascontiguousarray(p1, np.float64).copy()
    b = None if p2 is None else np.ascontiguousarray(p2, np.float64).copy()
    sv = C.c_double(svmax)
    ng = C.c_longlong()
    nc = lib().orc_dsmc_apply(C.byref(g), a, a.size // 7, mass1,
                              None if b is None else b.ctypes.data_as(C.c_void_p),
                              0 if b is None else b.size // 7, mass2, mpw0, dt, C.byref(sv),
                              C.byref(rng), C.byref(ng))
    return a, b, sv.value, nc, ng.value


def source_sample(g, rng, charge, mass, mpw0, face, v_drift, den, T, dt, ef=None, rects=None,
                  max_out=100000):
    """one ColdBeamSource (T <= 0) / WarmBeamSource sample; returns particles"""
    out = np.zeros(7 * max_out)
    ra, nr = rects_arg(rects)
    efp = None
    if ef is not None:
        ef = np.ascontiguousarray(ef, np.float64)
        efp = ef.ctypes.data_as(C.c_void_p)
    n = lib().orc_source_sample(C.byref(g), ra, nr, efp, charge, mass, mpw0, face, v_drift, den, T,
                                dt, C.byref(rng), out, max_out)
    assert n != C.c_size_t(-1).value
    return out[: 7 * n]
