"""ctypes wrapper around the CPU FP64 oracle (oracle/liborpic.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product path never uses it.
All arrays are numpy float64/int32/uint32, C-contiguous, in the AoS layouts
documented in orpic.h.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborpic.so")

NODE, CELL, FACE_I, FACE_J, FACE_K, EDGE_I, EDGE_J, EDGE_K = range(8)

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def _declare(L):
    c_int, c_i64, c_double = ctypes.c_int, ctypes.c_int64, ctypes.c_double
    L.orc_spline.argtypes = [c_int, c_int, _D, c_double, _D, _I32, _D, _D, _D, _D]
    L.orc_apic_p2g.argtypes = [c_int, c_int, _D, c_double, _I32, _I32, c_i64,
                               _D, _D, _D, _D, _D]
    L.orc_apic_g2p.argtypes = [c_int, c_int, _D, c_double, _I32, _I32, c_i64,
                               _D, _D, _D]
    L.orc_polypic_p2g.argtypes = [c_int, c_int, _D, c_double, _I32, _I32, c_i64,
                                  _D, _D, _D, _D, c_int, c_double, _D, _D]
    L.orc_polypic_g2p.argtypes = [c_int, c_int, _D, c_double, _I32, _I32, c_i64,
                                  _D, _D, c_int, _D]
    L.orc_g2p_gradu.argtypes = [c_int, c_int, _D, c_double, _I32, _I32, c_i64,
                                _D, _D, _D]
    L.orc_g2p_interp.argtypes = [c_int, c_int, c_int, _D, c_double, _I32, _I32, c_i64, _D,
                                 _D, c_int, _D]
    L.orc_p2g_interp.argtypes = [c_int, c_int, c_int, _D, c_double, _I32, _I32, c_i64, _D,
                                 _D, c_int, _D]
    L.orc_grid_normalize.argtypes = [c_i64, c_int, _D, _D, _D]
    L.orc_cell_keys.argtypes = [c_i64, _D, _D, c_double, _I32, _I32, _I32, c_int,
                                _U32, _I32]
    L.orc_bin_sort.argtypes = [c_i64, _U32, ctypes.c_uint32, _U32, _U32]
    L.orc_grid_halo.argtypes = [c_int, _I32, _I32, c_int, _D, _I32, _I32, _I32]
    L.orc_uniform01.argtypes = [ctypes.c_uint64, ctypes.c_uint64, c_int]
    L.orc_uniform01.restype = c_double
    L.orc_init_particles.argtypes = [c_int, _D, c_double, _I32, _I32, c_int, ctypes.c_uint64,
                                     _D, _D]
    L.orc_init_particles.restype = c_i64
    L.orc_det3.argtypes = [_D]
    L.orc_det3.restype = c_double
    L.orc_inverse3.argtypes = [_D, _D]
    L.orc_update_F.argtypes = [c_i64, _D, c_double, _D, _D, _D]
    L.orc_apic_step_omp.argtypes = [c_int, _D, c_double, _I32, _I32, c_i64,
                                    _D, _D, _D, _D, _D, _D, c_int, c_int]
    L.orc_polypic_nmode.argtypes = [c_int]
    L.orc_entity_offset.argtypes = [c_int, c_int, c_double]
    L.orc_entity_offset.restype = c_double


def _p(a, t=_D):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise OracleError(f"{what} failed with code {rc}")
    return rc


def spline(order, entity, low, h, x):
    s = np.zeros((3, 4), np.int32)
    w, g, g2, d = (np.zeros((3, 4)) for _ in range(4))
    low, x = _f64(low), _f64(x)
    nk = _check(lib().orc_spline(order, entity, _p(low), h, _p(x), _p(s, _I32),
                                 _p(w), _p(g), _p(g2), _p(d)), "spline")
    return dict(nk=nk, s=s[:, :nk], w=w[:, :nk], g=g[:, :nk], g2=g2[:, :nk],
                d=d[:, :nk])


def apic_p2g(order, entity, low, h, lo, n, x, m, c, gm, gmu):
    """Accumulate into gm [n0,n1,n2] and gmu [n0,n1,n2,ncomp] (in place)."""
    x, m, c = _f64(x), _f64(m), _f64(c)
    assert gm.flags.c_contiguous and gmu.flags.c_contiguous
    _check(lib().orc_apic_p2g(order, entity, _p(_f64(low)), h, _p(_i32(lo), _I32),
                              _p(_i32(n), _I32), x.shape[0], _p(x), _p(m), _p(c),
                              _p(gm), _p(gmu)), "apic_p2g")


def apic_g2p(order, entity, low, h, lo, n, x, gu):
    x = _f64(x)
    c = np.zeros((x.shape[0], 4, 3))
    gu = _f64(gu)
    _check(lib().orc_apic_g2p(order, entity, _p(_f64(low)), h, _p(_i32(lo), _I32),
                              _p(_i32(n), _I32), x.shape[0], _p(x), _p(gu), _p(c)),
           "apic_g2p")
    return c


def polypic_nmode(order):
    return lib().orc_polypic_nmode(order)


def polypic_p2g(order, entity, low, h, lo, n, x, m, u, c, dt, gmu, gm):
    x, m, u, c = _f64(x), _f64(m), _f64(u), _f64(c)
    ncomp = c.shape[-1]
    _check(lib().orc_polypic_p2g(order, entity, _p(_f64(low)), h,
                                 _p(_i32(lo), _I32), _p(_i32(n), _I32),
                                 x.shape[0], _p(x), _p(m), _p(u), _p(c), ncomp,
                                 dt, _p(gmu), _p(gm)), "polypic_p2g")


def polypic_g2p(order, entity, low, h, lo, n, x, gu, ncomp, c=None):
    x, gu = _f64(x), _f64(gu)
    nmode = polypic_nmode(order)
    if c is None:
        c = np.zeros((x.shape[0], nmode, ncomp))
    _check(lib().orc_polypic_g2p(order, entity, _p(_f64(low)), h,
                                 _p(_i32(lo), _I32), _p(_i32(n), _I32),
                                 x.shape[0], _p(x), _p(gu), ncomp, _p(c)),
           "polypic_g2p")
    return c


def g2p_gradu(order, entity, low, h, lo, n, x, gu):
    x, gu = _f64(x), _f64(gu)
    out = np.zeros((x.shape[0], 3, 3))
    _check(lib().orc_g2p_gradu(order, entity, _p(_f64(low)), h, _p(_i32(lo), _I32),
                               _p(_i32(n), _I32), x.shape[0], _p(x), _p(gu),
                               _p(out)), "g2p_gradu")
    return out


INTERP_VALUE, INTERP_GRADIENT, INTERP_DIVERGENCE = 0, 1, 2


def g2p_interp(op, order, entity, low, h, lo, n, x, f, ncomp):
    """G2P::value / gradient / divergence of a grid field f [nodes..., ncomp]
    at the particles: out [P, nout] (nout = ncomp, 3*ncomp or 1)."""
    x, f = _f64(x), _f64(f)
    nout = {INTERP_VALUE: ncomp, INTERP_GRADIENT: 3 * ncomp, INTERP_DIVERGENCE: 1}[op]
    out = np.zeros((x.shape[0], nout))
    _check(lib().orc_g2p_interp(op, order, entity, _p(_f64(low)), h, _p(_i32(lo), _I32),
                                _p(_i32(n), _I32), x.shape[0], _p(x), _p(f), ncomp,
                                _p(out)), "g2p_interp")
    return out


def p2g_interp(op, order, entity, low, h, lo, n, x, v, ncomp, f=None):
    """P2G::value / gradient / divergence of particle values v [P, ncomp],
    accumulated into a grid field [nodes..., ngrid] (zeros when f is None)."""
    x, v = _f64(x), _f64(v)
    ngrid = {INTERP_VALUE: ncomp, INTERP_GRADIENT: 3, INTERP_DIVERGENCE: ncomp // 3}[op]
    if f is None:
        f = np.zeros(tuple(int(k) for k in n) + (ngrid,))
    else:
        f = _f64(f).copy()
    _check(lib().orc_p2g_interp(op, order, entity, _p(_f64(low)), h, _p(_i32(lo), _I32),
                                _p(_i32(n), _I32), x.shape[0], _p(x), _p(v), ncomp,
                                _p(f)), "p2g_interp")
    return f


def grid_normalize(m, mu):
    m, mu = _f64(m), _f64(mu)
    ncomp = mu.shape[-1]
    u = np.zeros_like(mu)
    _check(lib().orc_grid_normalize(m.size, ncomp, _p(m), _p(mu), _p(u)),
           "grid_normalize")
    return u


def cell_keys(x, low, h, cell_lo, ncell, brick=(1, 1, 1), dual=0):
    x = _f64(x)
    keys = np.zeros(x.shape[0], np.uint32)
    cells = np.zeros((x.shape[0], 3), np.int32)
    _check(lib().orc_cell_keys(x.shape[0], _p(x), _p(_f64(low)), h,
                               _p(_i32(cell_lo), _I32), _p(_i32(ncell), _I32),
                               _p(_i32(brick), _I32), int(dual), _p(keys, _U32),
                               _p(cells, _I32)), "cell_keys")
    return keys, cells


def grid_halo(op, lo, g, own_lo, own_n, periodic):
    """op 1 scatter-sum, 0 gather, in place on g [n0, n1, n2(, ncomp)]."""
    assert g.flags.c_contiguous and g.dtype == np.float64
    ncomp = 1 if g.ndim == 3 else g.shape[3]
    _check(lib().orc_grid_halo(op, _p(_i32(lo), _I32), _p(_i32(g.shape[:3]), _I32), ncomp,
                               _p(g), _p(_i32(own_lo), _I32), _p(_i32(own_n), _I32),
                               _p(_i32([int(bool(v)) for v in periodic]), _I32)), "grid_halo")
    return g


def uniform01(seed, pid, d):
    return lib().orc_uniform01(seed, pid, d)


def init_particles(mode, low, h, cell_lo, ncell, ppc, seed=0):
    """Positions [np, 3] and volumes [np] of initializeParticles without the
    creation functor (mode 0 uniform lattice, 1 random)."""
    npc = ppc ** 3 if mode == 0 else ppc
    np_ = int(np.prod(ncell)) * npc
    x = np.zeros((np_, 3))
    vol = np.zeros(np_)
    n = lib().orc_init_particles(mode, _p(_f64(low)), h, _p(_i32(cell_lo), _I32),
                                 _p(_i32(ncell), _I32), ppc, seed, _p(x), _p(vol))
    _check(n if n < 0 else 0, "init_particles")
    return x, vol


def bin_sort(keys, nkeys):
    keys = np.ascontiguousarray(keys, np.uint32)
    perm = np.zeros(keys.size, np.uint32)
    start = np.zeros(nkeys + 1, np.uint32)
    _check(lib().orc_bin_sort(keys.size, _p(keys, _U32), nkeys, _p(perm, _U32),
                              _p(start, _U32)), "bin_sort")
    return perm, start


def det3(a):
    return lib().orc_det3(_p(_f64(a)))


def inverse3(a):
    r = np.zeros(9)
    lib().orc_inverse3(_p(_f64(a)), _p(r))
    return r.reshape(3, 3)


def update_F(gradu, dt, F):
    gradu = _f64(gradu)
    F = _f64(F).copy()
    n = F.shape[0]
    det = np.zeros(n)
    Finv = np.zeros((n, 3, 3))
    _check(lib().orc_update_F(n, _p(gradu), dt, _p(F), _p(det), _p(Finv)),
           "update_F")
    return F, det, Finv


def apic_step_omp(order, low, h, lo, n, x, m, c, mode=1, nthreads=0):
    """p2g -> normalise -> g2p on a zeroed node window; c is updated in place."""
    x, m = _f64(x), _f64(m)
    assert c.dtype == np.float64 and c.flags.c_contiguous
    nn = int(np.prod(n))
    gm = np.empty(nn)
    gmu = np.empty((nn, 3))
    gu = np.empty((nn, 3))
    _check(lib().orc_apic_step_omp(order, _p(_f64(low)), h, _p(_i32(lo), _I32),
                                   _p(_i32(n), _I32), x.shape[0], _p(x), _p(m),
                                   _p(c), _p(gm), _p(gmu), _p(gu), mode, nthreads),
           "apic_step_omp")
    return gm, gmu, gu
