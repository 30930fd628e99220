"""The adaptive-cutoff producer (sre_arc_* in include/sre.h) from Python.

    arc = Arc(box, refh=1.0)
    arc.set_candidates(3.0, 7.5, 0.25)
    arc.deposit(engine)                # every sampling period
    arc.cal_error(); arc.cal_rcut(0.0045); arc.refine(2)
    arc.apply(engine)                  # rc lattice of replica 0

Restates DensityProfile_PiecewiseConst and AdaptRCut
(H/src/DensityProfile.cu:27-99, H/src/AdaptRCut.cu:96-508) on the device.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from .engine import _ptr, check


class Arc:
    def __init__(self, box, refh=1.0, device=-1, stream=None):
        self._h = C.c_void_p()
        b = np.ascontiguousarray(np.asarray(box, np.float64).reshape(3))
        check(lib().sre_arc_create(C.byref(self._h), _ptr(b, C.c_double), float(refh),
                                   int(device), C.c_void_p(stream) if stream else None))
        d = np.zeros(3, np.int32)
        check(lib().sre_arc_dims(self._h, _ptr(d, C.c_int32), None))
        self.dims = tuple(int(v) for v in d)
        self.box = tuple(float(np.float32(v)) for v in box)
        self.rcs = None

    def close(self):
        if self._h:
            lib().sre_arc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nele(self):
        return int(np.prod(self.dims))

    def set_candidates(self, rcmin, rcmax, rcstep):
        check(lib().sre_arc_set_candidates(self._h, float(rcmin), float(rcmax),
                                           float(rcstep)))
        n = C.c_int()
        check(lib().sre_arc_dims(self._h, None, C.byref(n)))
        r = np.empty(n.value, np.float64)
        check(lib().sre_arc_rc_list(self._h, _ptr(r, C.c_double)))
        self.rcs = r
        return r

    def clear(self):
        check(lib().sre_arc_clear(self._h))

    def deposit(self, source, replica=0):
        """source: an Engine (its bound positions) or a CUDA float tensor (N, >=3)."""
        if hasattr(source, "_h") and hasattr(source, "n_replicas"):
            check(lib().sre_arc_deposit_engine(self._h, source._h, int(replica)))
            return
        t = source
        assert t.is_cuda and t.dtype.itemsize == 4 and t.is_contiguous() and t.dim() == 2
        check(lib().sre_arc_deposit(self._h, C.c_void_p(t.data_ptr()), int(t.shape[0]),
                                    int(t.shape[1])))

    def set_density(self, rho):
        r = np.ascontiguousarray(np.asarray(rho, np.float64).reshape(self.dims))
        check(lib().sre_arc_set_density(self._h, _ptr(r, C.c_double)))

    def density(self):
        out = np.empty(self.dims, np.float64)
        check(lib().sre_arc_density(self._h, _ptr(out, C.c_double)))
        return out

    def cal_error(self, keep_parts=False):
        check(lib().sre_arc_cal_error(self._h, 1 if keep_parts else 0))

    def cal_rcut(self, prec):
        check(lib().sre_arc_cal_rcut(self._h, float(prec)))

    def refine(self, times=1):
        for _ in range(int(times)):
            check(lib().sre_arc_refine(self._h))

    def result(self):
        rc = np.empty(self.dims, np.float32)
        err = np.empty(self.dims, np.float64)
        idx = np.empty(self.dims, np.int32)
        check(lib().sre_arc_result(self._h, _ptr(rc, C.c_float), _ptr(err, C.c_double),
                                   _ptr(idx, C.c_int32)))
        return rc, err, idx

    def set_index(self, idx):
        i = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(self.dims))
        check(lib().sre_arc_set_index(self._h, _ptr(i, C.c_int32)))

    def errors(self, parts=False):
        nrc = len(self.rcs)
        err = np.empty((nrc,) + self.dims, np.float64)
        p = np.empty((nrc,) + self.dims + (4,), np.float64) if parts else None
        check(lib().sre_arc_errors(self._h, _ptr(err, C.c_double),
                                   _ptr(p, C.c_double) if parts else None))
        return (err, p) if parts else err

    def kernels(self):
        nrc = len(self.rcs)
        S1 = np.empty((nrc,) + self.dims, np.float64)
        G = np.empty((nrc,) + self.dims, np.float64)
        check(lib().sre_arc_kernels(self._h, _ptr(S1, C.c_double), _ptr(G, C.c_double)))
        return S1, G

    def pressure(self):
        """(pxx, pyy, pzz) correction for the current density and candidates."""
        p = np.zeros(3, np.float64)
        check(lib().sre_arc_pressure(self._h, _ptr(p, C.c_double)))
        return p

    def pressure_integrals(self):
        nrc = len(self.rcs)
        I1 = np.empty((nrc,) + self.dims, np.float64)
        I5 = np.empty((nrc,) + self.dims, np.float64)
        check(lib().sre_arc_pressure_integrals(self._h, _ptr(I1, C.c_double),
                                               _ptr(I5, C.c_double)))
        return I1, I5

    def force_grid(self):
        """Mean error force (nx, ny, nz, 3) at each cell's chosen candidate
        (needs cal_error(keep_parts=True))."""
        f = np.empty(self.dims + (3,), np.float32)
        check(lib().sre_arc_force_grid(self._h, _ptr(f, C.c_float)))
        return f

    def apply_force_correction(self, engine, replica=0):
        check(lib().sre_arc_apply_force_correction(self._h, engine._h, int(replica)))

    def apply(self, engine, replica=0):
        check(lib().sre_arc_apply(self._h, engine._h, int(replica)))
