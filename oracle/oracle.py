"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker (and the timed CPU path).  The product never imports this module.
Each function names the reference code it restates (see oracle.cpp).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is missing: run python oracle/build.py")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def _declare(L):
    f32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
    i32p, i64p, u32p = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint32)
    L.or_quantize.argtypes = [C.c_int, f32p, C.c_int, f32p, u32p]
    L.or_assign_rc.argtypes = [C.c_int, f32p, C.c_int, f32p, C.c_int, C.c_int, C.c_int,
                               f32p, f32p, i32p]
    for fn in (L.or_build_pairs_brute, L.or_build_pairs_cells):
        fn.argtypes = [C.c_int, f32p, C.c_int, f32p, f32p, C.c_float, i64p, i32p]
        fn.restype = C.c_int64
    L.or_tail_point.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, f32p, f32p,
                                f64p, C.c_int, C.c_double, f64p, f64p]
    L.or_forces.argtypes = [C.c_int, f32p, C.c_int, f32p, i32p, C.c_int, f32p, f32p,
                            f32p, f32p, i64p, i32p, f64p, f64p, f64p, f64p]
    L.or_forces.restype = C.c_int64
    L.or_tail_atoms.argtypes = [C.c_int, f32p, C.c_int, f32p, i32p, C.c_int, f32p, f32p,
                                C.c_int, f32p, f64p, C.c_int, C.c_int, f64p]
    L.or_grad_pairs.argtypes = [C.c_int, f32p, C.c_int, f32p, i32p, C.c_int, f32p, f32p,
                                f32p, i64p, i32p, C.c_double, f64p, f64p, f64p]
    L.or_grad_tail.argtypes = [C.c_int, f32p, C.c_int, f32p, i32p, C.c_int, f32p, f32p,
                               C.c_int, f32p, f64p, C.c_int, C.c_double, f64p, C.c_double,
                               f64p, f64p, f64p]
    L.or_num_threads.restype = C.c_int
    L.or_set_threads.argtypes = [C.c_int]


def _xyz(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    assert x.ndim == 2 and x.shape[1] in (3, 4)
    return x, x.shape[1]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def quantize(xyz, box):
    x, st = _xyz(xyz)
    u = np.empty((x.shape[0], 3), np.uint32)
    lib().or_quantize(x.shape[0], _p(x, C.c_float), st, _p(_f32(box), C.c_float),
                      _p(u, C.c_uint32))
    return u


def assign_rc(xyz, box, table):
    """Per-atom cutoff from the rc lattice table[nx, ny, nz] (AssignRCut.cu:77-104)."""
    x, st = _xyz(xyz)
    t = _f32(table)
    nx, ny, nz = t.shape
    rc = np.empty(x.shape[0], np.float32)
    bz = np.empty(x.shape[0], np.int32)
    lib().or_assign_rc(x.shape[0], _p(x, C.c_float), st, _p(_f32(box), C.c_float),
                       nx, ny, nz, _p(t.ravel(), C.c_float), _p(rc, C.c_float),
                       _p(bz, C.c_int32))
    return rc, bz


def build_pairs(xyz, box, rc, skin, method="cells"):
    """CSR Verlet list (offsets[N+1], j) with rows ascending."""
    x, st = _xyz(xyz)
    n = x.shape[0]
    rc = _f32(rc)
    fn = lib().or_build_pairs_cells if method == "cells" else lib().or_build_pairs_brute
    off = np.zeros(n + 1, np.int64)
    b = _f32(box)
    total = fn(n, _p(x, C.c_float), st, _p(b, C.c_float), _p(rc, C.c_float), skin,
               _p(off, C.c_int64), None)
    j = np.empty(max(total, 1), np.int32)
    fn(n, _p(x, C.c_float), st, _p(b, C.c_float), _p(rc, C.c_float), skin,
       _p(off, C.c_int64), _p(j, C.c_int32))
    return off, j[:total]


def tail_point(z, R, ta, eps, sig, rho, Lz):
    eps, sig = _f32(eps), _f32(sig)
    nt = int(round(np.sqrt(eps.size)))
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(nt, -1)
    F, E = C.c_double(), C.c_double()
    lib().or_tail_point(z, R, ta, nt, _p(eps.ravel(), C.c_float), _p(sig.ravel(), C.c_float),
                        _p(rho.ravel(), C.c_double), rho.shape[1], Lz, C.byref(F), C.byref(E))
    return F.value, E.value


def tail_atoms(xyz, box, types, eps, sig, rc_z, rho, exact=False):
    x, st = _xyz(xyz)
    n = x.shape[0]
    eps, sig = _f32(eps), _f32(sig)
    nt = int(round(np.sqrt(eps.size)))
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(nt, -1)
    rcz = _f32(rc_z).ravel()
    t = np.ascontiguousarray(types, dtype=np.int32) if types is not None else None
    out = np.empty((n, 2), np.float64)
    lib().or_tail_atoms(n, _p(x, C.c_float), st, _p(_f32(box), C.c_float), _p(t, C.c_int32),
                        nt, _p(eps.ravel(), C.c_float), _p(sig.ravel(), C.c_float),
                        rcz.size, _p(rcz, C.c_float), _p(rho.ravel(), C.c_double),
                        rho.shape[1], 1 if exact else 0, _p(out, C.c_double))
    return out


def forces(xyz, box, types, eps, sig, rc, off, j, shift=None, tail=None):
    """FP64 forces (N,3), per-atom energy (N,) and virial diagonal (N,3), n_in_rc."""
    x, st = _xyz(xyz)
    n = x.shape[0]
    eps, sig = _f32(eps), _f32(sig)
    nt = int(round(np.sqrt(eps.size)))
    t = np.ascontiguousarray(types, dtype=np.int32) if types is not None else None
    sh = _f32(shift).ravel() if shift is not None else None
    tl = np.ascontiguousarray(tail, dtype=np.float64) if tail is not None else None
    F = np.empty((n, 3), np.float64)
    E = np.empty(n, np.float64)
    W = np.empty((n, 3), np.float64)
    off = np.ascontiguousarray(off, dtype=np.int64)
    j = np.ascontiguousarray(j, dtype=np.int32)
    nin = lib().or_forces(n, _p(x, C.c_float), st, _p(_f32(box), C.c_float), _p(t, C.c_int32),
                          nt, _p(eps.ravel(), C.c_float), _p(sig.ravel(), C.c_float),
                          _p(sh, C.c_float), _p(_f32(rc), C.c_float), _p(off, C.c_int64),
                          _p(j, C.c_int32), _p(tl, C.c_double), _p(F, C.c_double),
                          _p(E, C.c_double), _p(W, C.c_double))
    return F, E, W, nin


def grad_pairs(xyz, box, types, eps, sig, rc, off, j, dLdE, dLdF):
    x, st = _xyz(xyz)
    n = x.shape[0]
    eps, sig = _f32(eps), _f32(sig)
    nt = int(round(np.sqrt(eps.size)))
    t = np.ascontiguousarray(types, dtype=np.int32) if types is not None else None
    g = np.ascontiguousarray(dLdF, dtype=np.float64) if dLdF is not None else None
    ge = np.zeros(nt * nt)
    gs = np.zeros(nt * nt)
    lib().or_grad_pairs(n, _p(x, C.c_float), st, _p(_f32(box), C.c_float), _p(t, C.c_int32),
                        nt, _p(eps.ravel(), C.c_float), _p(sig.ravel(), C.c_float),
                        _p(_f32(rc), C.c_float), _p(np.ascontiguousarray(off, np.int64), C.c_int64),
                        _p(np.ascontiguousarray(j, np.int32), C.c_int32), float(dLdE),
                        _p(g, C.c_double), _p(ge, C.c_double), _p(gs, C.c_double))
    return ge.reshape(nt, nt), gs.reshape(nt, nt)


def grad_tail(xyz, box, types, eps, sig, rc_z, rho, dLdE, dLdF, h=1e-4):
    x, st = _xyz(xyz)
    n = x.shape[0]
    eps, sig = _f32(eps), _f32(sig)
    nt = int(round(np.sqrt(eps.size)))
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(nt, -1)
    rcz = _f32(rc_z).ravel()
    t = np.ascontiguousarray(types, dtype=np.int32) if types is not None else None
    g = np.ascontiguousarray(dLdF, dtype=np.float64) if dLdF is not None else None
    ge = np.zeros(nt * nt)
    gs = np.zeros(nt * nt)
    grc = np.zeros(rcz.size)
    lib().or_grad_tail(n, _p(x, C.c_float), st, _p(_f32(box), C.c_float), _p(t, C.c_int32), nt,
                       _p(eps.ravel(), C.c_float), _p(sig.ravel(), C.c_float), rcz.size,
                       _p(rcz, C.c_float), _p(rho.ravel(), C.c_double), rho.shape[1],
                       float(dLdE), _p(g, C.c_double), h, _p(ge, C.c_double),
                       _p(gs, C.c_double), _p(grc, C.c_double))
    return ge.reshape(nt, nt), gs.reshape(nt, nt), grc


def set_threads(t):
    lib().or_set_threads(int(t))


def num_threads():
    return lib().or_num_threads()
