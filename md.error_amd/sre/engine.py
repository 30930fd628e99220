"""Python binding over the C ABI of libsre.so.

``Engine`` is a thin object wrapper; every method is one C call.  Errors
become exceptions named after the reference's (MDExcptCuda for device/runtime
failures, H/include/MDException.h:11-29; the list/cell/box flags of
H/include/error.h:19-26).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib


class SREError(RuntimeError):
    """Base of all engine errors; ``status`` is the sre_status code."""

    def __init__(self, status, msg):
        super().__init__(f"[sre status {status}] {msg}")
        self.status = status


class MDExcptCuda(SREError):
    """Runtime (HIP) failure, as checkCUDAError raises MDExcptCuda."""


class ShortNeighborList(SREError):
    """mdErrorShortNeighborList: the list could not be regrown."""


class ShortCellList(SREError):
    """mdErrorShortCellList: the cell grid does not fit the sort key."""


class BoxTooSmall(SREError):
    """rc_max + skin is not below half of a box edge."""


class ForceRange(SREError):
    """A pair force left the fixed-point range (overlapping atoms)."""


class SkinExceeded(SREError):
    """An atom moved more than skin/2 without a rebuild (graph mode)."""


class NotReady(SREError):
    """Call order error (e.g. compute before rebuild)."""


class InvalidArgument(SREError, ValueError):
    pass


class MDExcptCannotOpenFile(SREError, OSError):
    """A file could not be opened, read or written (the reference's
    MDExcptCannotOpenFile)."""


_EXC = {
    _lib.ERR_LIST_OVERFLOW: ShortNeighborList,
    _lib.ERR_CELL: ShortCellList,
    _lib.ERR_INVALID: InvalidArgument,
    _lib.ERR_HIP: MDExcptCuda,
    _lib.ERR_BOX_TOO_SMALL: BoxTooSmall,
    _lib.ERR_FORCE_RANGE: ForceRange,
    _lib.ERR_SKIN: SkinExceeded,
    _lib.ERR_NOT_READY: NotReady,
    _lib.ERR_IO: MDExcptCannotOpenFile,
}


def check(status):
    if status != _lib.OK:
        msg = lib().sre_last_error().decode(errors="replace")
        raise _EXC.get(status, SREError)(status, msg)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _flags(stats, tail, fcorr=False):
    return ((_lib.STATS if stats else 0) | (_lib.TAIL if tail else 0)
            | (_lib.FCORR if fcorr else 0))


class Engine:
    """One engine: R >= 1 periodic systems sharing one LJ force field."""

    def __init__(self, n_types=1, skin=0.3, precision="mixed", stream=None, device=-1):
        prec = {"mixed": _lib.PREC_MIXED, "fp64": _lib.PREC_FP64}[precision]
        opt = _lib.Options(int(n_types), float(skin), prec,
                           C.c_void_p(stream) if stream else None, int(device))
        h = C.c_void_p()
        check(lib().sre_create(C.byref(h), C.byref(opt)))
        self._h = h
        self.n_types = int(n_types)
        self.skin = float(skin)
        self.precision = precision
        self.n_atoms = 0
        self.n_replicas = 0
        self._pos_ref = None

    def close(self):
        if getattr(self, "_h", None):
            lib().sre_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- description -------------------------------------------------------
    def set_system(self, n_atoms, boxes):
        n = np.ascontiguousarray(np.atleast_1d(n_atoms), dtype=np.int32)
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.float32).reshape(-1, 3))
        assert b.shape[0] == n.size
        check(lib().sre_set_system(self._h, n.size, _ptr(n, C.c_int32), _ptr(b, C.c_float)))
        self.n_atoms = int(n.sum())
        self.n_replicas = int(n.size)
        self._rc_layout = [None] * self.n_replicas

    def set_types(self, types):
        t = np.ascontiguousarray(types, dtype=np.int32)
        assert t.size == self.n_atoms
        check(lib().sre_set_types(self._h, _ptr(t, C.c_int32)))

    def set_pair_params(self, eps, sigma, shift=None):
        nt = self.n_types
        e = np.ascontiguousarray(np.asarray(eps, np.float32).reshape(nt, nt))
        s = np.ascontiguousarray(np.asarray(sigma, np.float32).reshape(nt, nt))
        sh = (np.ascontiguousarray(np.asarray(shift, np.float32).reshape(nt, nt))
              if shift is not None else None)
        check(lib().sre_set_pair_params(self._h, _ptr(e, C.c_float), _ptr(s, C.c_float),
                                        _ptr(sh, C.c_float) if sh is not None else None))

    def set_rc_table(self, table, replica=0):
        t = np.ascontiguousarray(np.asarray(table, np.float32))
        if t.ndim == 1:
            t = t.reshape(1, 1, -1)
        nx, ny, nz = t.shape
        check(lib().sre_set_rc_table(self._h, int(replica), nx, ny, nz, _ptr(t, C.c_float)))
        self._rc_layout[int(replica)] = t.shape

    @property
    def rc_size(self):
        """Total number of rc lattice values over all replicas."""
        return int(sum(int(np.prod(s)) for s in self._rc_layout if s is not None))

    def set_force_grid(self, f, replica=0):
        """Grid force correction (nx, ny, nz, 3) added with fcorr=True; None removes it."""
        if f is None:
            check(lib().sre_set_force_grid(self._h, int(replica), 0, 0, 0, None))
            return
        a = np.ascontiguousarray(np.asarray(f, np.float32))
        assert a.ndim == 4 and a.shape[3] == 3
        nx, ny, nz = a.shape[:3]
        check(lib().sre_set_force_grid(self._h, int(replica), nx, ny, nz,
                                       _ptr(a, C.c_float)))

    def set_density_profile(self, rho, replica=0):
        if rho is None:
            check(lib().sre_set_density_profile(self._h, int(replica), 0, None))
            return
        r = np.ascontiguousarray(np.asarray(rho, np.float64).reshape(self.n_types, -1))
        check(lib().sre_set_density_profile(self._h, int(replica), r.shape[1],
                                            _ptr(r, C.c_double)))

    def set_tail_mode(self, exact=False):
        check(lib().sre_set_tail_mode(self._h, _lib.TAIL_EXACT if exact else _lib.TAIL_BIN))

    # -- positions ---------------------------------------------------------
    def upload_positions(self, xyz):
        x = np.ascontiguousarray(np.asarray(xyz, np.float32))
        assert x.shape[0] == self.n_atoms and x.shape[1] in (3, 4)
        check(lib().sre_upload_positions(self._h, _ptr(x, C.c_float), x.shape[1]))

    def bind_positions(self, tensor_or_ptr, stride=4):
        """Binds a device array: a torch CUDA tensor (N, stride) or a raw pointer."""
        if hasattr(tensor_or_ptr, "data_ptr"):
            t = tensor_or_ptr
            assert t.is_cuda and t.dtype.itemsize == 4 and t.is_contiguous()
            stride = t.shape[1]
            self._pos_ref = t
            ptr = t.data_ptr()
        else:
            ptr = int(tensor_or_ptr)
        check(lib().sre_bind_positions(self._h, C.c_void_p(ptr), int(stride)))

    # -- hot path ----------------------------------------------------------
    def rebuild(self):
        check(lib().sre_rebuild(self._h))

    def update_positions(self, check_skin=True):
        need = C.c_int(0)
        check(lib().sre_update_positions(self._h, C.byref(need) if check_skin else None))
        return bool(need.value)

    def compute(self, stats=False, tail=False, fcorr=False):
        check(lib().sre_compute(self._h, _flags(stats, tail, fcorr)))

    def step(self, stats=False, tail=False, fcorr=False):
        rebuilt = C.c_int(0)
        check(lib().sre_step(self._h, _flags(stats, tail, fcorr), C.byref(rebuilt)))
        return bool(rebuilt.value)

    # -- results -----------------------------------------------------------
    def forces(self):
        out = np.empty((self.n_atoms, 3), np.float64)
        check(lib().sre_get_forces(self._h, _ptr(out, C.c_double), 0))
        return out

    def forces_to_device(self, tensor):
        assert tensor.is_cuda and tensor.dtype.itemsize == 8
        check(lib().sre_get_forces(self._h, C.cast(C.c_void_p(tensor.data_ptr()),
                                                   C.POINTER(C.c_double)), 1))

    def add_forces_f32_soa(self, fx, fy, fz):
        check(lib().sre_add_forces_f32_soa(self._h, C.c_void_p(fx.data_ptr()),
                                           C.c_void_p(fy.data_ptr()),
                                           C.c_void_p(fz.data_ptr())))

    def forces_fixed(self):
        out = np.empty((self.n_atoms, 3), np.int64)
        check(lib().sre_get_forces_fixed(self._h, _ptr(out, C.c_int64)))
        return out

    def stats(self):
        out = np.empty((self.n_replicas, 4), np.float64)
        check(lib().sre_get_stats(self._h, _ptr(out, C.c_double)))
        return out

    def atom_stats(self):
        out = np.empty((self.n_atoms, 4), np.float64)
        check(lib().sre_get_atom_stats(self._h, _ptr(out, C.c_double)))
        return out

    def neighbor_count(self, in_rc=True):
        """(list entries, pairs inside rc at the current positions, longest
        list).  The in-rc count runs a counting kernel; in_rc=False skips it
        (returns -1 there)."""
        nl, ni, mx = C.c_int64(), C.c_int64(-1), C.c_int32()
        check(lib().sre_neighbor_count(self._h, C.byref(nl),
                                       C.byref(ni) if in_rc else None, C.byref(mx)))
        return nl.value, ni.value, mx.value

    def export_pairs(self):
        off = np.empty(self.n_atoms + 1, np.int64)
        check(lib().sre_export_pairs(self._h, _ptr(off, C.c_int64), None))
        j = np.empty(max(int(off[-1]), 1), np.int32)
        check(lib().sre_export_pairs(self._h, _ptr(off, C.c_int64), _ptr(j, C.c_int32)))
        return off, j[:int(off[-1])]

    def rcut(self):
        out = np.empty(self.n_atoms, np.float32)
        check(lib().sre_get_rcut(self._h, _ptr(out, C.c_float)))
        return out

    def assign_rcut(self):
        """AssignRCut::assign: rc_i from the tables at the current positions,
        without a list rebuild; from then on rc_i changes only at these calls
        (the reference's timing) instead of at every rebuild."""
        check(lib().sre_assign_rcut(self._h))

    def adjoint(self, dLdE=None, dLdF=None, tail=False, n_rc=None):
        """Gradient of sum_r dLdE[r] E_r + sum_i dLdF_i . F_i.

        dLdF: a CUDA float64 tensor (N, 3) in caller order, or None.
        Returns (g_eps (nt, nt), g_sigma (nt, nt), g_rc (total rc lattice size,)).
        """
        nt = self.n_types
        ge = np.zeros(nt * nt)
        gs = np.zeros(nt * nt)
        grc = np.zeros(max(int(n_rc or 0), 1))
        dE = np.ascontiguousarray(np.broadcast_to(
            np.asarray(0.0 if dLdE is None else dLdE, np.float64), (self.n_replicas,)))
        ptr = None
        if dLdF is not None:
            assert dLdF.is_cuda and dLdF.dtype.itemsize == 8 and dLdF.is_contiguous()
            ptr = C.c_void_p(dLdF.data_ptr())
        check(lib().sre_adjoint(self._h, _ptr(dE, C.c_double), ptr,
                                _lib.TAIL if tail else 0, _ptr(ge, C.c_double),
                                _ptr(gs, C.c_double),
                                _ptr(grc, C.c_double) if n_rc else None))
        return ge.reshape(nt, nt), gs.reshape(nt, nt), (grc if n_rc else None)

    # -- tuning, graphs, timing ---------------------------------------------
    def tuning(self):
        t = _lib.Tuning()
        check(lib().sre_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.Tuning()
        check(lib().sre_get_tuning(self._h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        check(lib().sre_set_tuning(self._h, C.byref(t)))

    def autotune(self, stats=True, tail=False, cache_path=None):
        check(lib().sre_autotune(self._h, _flags(stats, tail),
                                 cache_path.encode() if cache_path else None))
        return self.tuning()

    def graph_run(self, n_steps, stats=False, tail=False, perturb=False, fcorr=False):
        """Replays [perturb +] update + compute n_steps times as one hipGraph."""
        f = _flags(stats, tail, fcorr) | (_lib.PERTURB if perturb else 0)
        check(lib().sre_graph_run(self._h, f, int(n_steps)))

    def run_segments(self, n_segments, steps, stats=False, tail=False, perturb=False,
                     fcorr=False):
        """n_segments x (graph_run(steps) + rebuild) with one synchronisation;
        returns (list length built after each segment, largest displacement
        of any segment)."""
        nl = np.zeros(max(int(n_segments), 1), np.int64)
        d = C.c_float()
        fl = _flags(stats, tail, fcorr) | (_lib.PERTURB if perturb else 0)
        check(lib().sre_run_segments(self._h, fl, int(n_segments), int(steps),
                                     _ptr(nl, C.c_int64), C.byref(d)))
        return nl[:int(n_segments)], d.value

    def set_perturbation(self, sigma, seed, first_step=0, anchored=False):
        """Perturbation applied by graph_run(perturb=True): a random walk
        (x += sigma N(0,1) per replay) or, anchored, vibration about the
        positions bound now (x = x0 + sigma N(0,1))."""
        check(lib().sre_set_perturbation(self._h, float(sigma), int(seed), int(first_step),
                                         1 if anchored else 0))

    def check_skin_flag(self):
        check(lib().sre_check_skin_flag(self._h))

    def last_rebuild_displacement(self):
        """Max displacement between the previous rebuild and the last one."""
        v = C.c_float()
        check(lib().sre_last_rebuild_displacement(self._h, C.byref(v)))
        return v.value

    def perturb(self, sigma, seed, step):
        check(lib().sre_perturb_positions(self._h, float(sigma), int(seed), int(step)))

    def time_compute(self, n_iter=10, stats=False, tail=False):
        ms = C.c_double()
        check(lib().sre_time_compute(self._h, _flags(stats, tail), int(n_iter), C.byref(ms)))
        return ms.value

    def time_rebuild(self, n_iter=3):
        ms = C.c_double()
        check(lib().sre_time_rebuild(self._h, int(n_iter), C.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().sre_synchronize(self._h))

    # -- MD integrator -------------------------------------------------------
    def set_masses(self, mass):
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32),
                                                 (self.n_types,)))
        check(lib().sre_set_masses(self._h, _ptr(m, C.c_float)))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, np.float32).reshape(self.n_atoms, 3)
        check(lib().sre_set_velocities(self._h, _ptr(v, C.c_float)))

    def velocities(self):
        out = np.empty((self.n_atoms, 3), np.float32)
        check(lib().sre_get_velocities(self._h, _ptr(out, C.c_float)))
        return out

    def set_integrator(self, dt, thermostat=None, ref_T=0.0, tau_T=0.0, remove_com=True):
        """Velocity Verlet with time step dt; thermostat None (NVE) or "nhc2"
        (two-link Nose-Hoover chain at ref_T, period tau_T); remove_com removes
        the centre-of-mass velocity at the start of every step."""
        th = {None: _lib.THERMO_NONE, "nve": _lib.THERMO_NONE,
              "nhc2": _lib.THERMO_NHC2}[thermostat]
        check(lib().sre_set_integrator(self._h, float(dt), th, float(ref_T), float(tau_T),
                                       1 if remove_com else 0))

    def md_run(self, n_steps, rebuild_every=10, stats=False, tail=False, fcorr=False):
        """n_steps MD steps; returns the number of list rebuilds.
        rebuild_every: a fixed interval (> 0), 0 = check the displacement every
        step (the reference's rule), "adaptive" = speculative graph stretches
        with rollback."""
        if rebuild_every == "adaptive":
            rebuild_every = -1
        nr = C.c_int(0)
        check(lib().sre_md_run(self._h, int(n_steps), int(rebuild_every),
                               _flags(stats, tail, fcorr), C.byref(nr)))
        return nr.value

    def md_state(self):
        s = _lib.MDState()
        check(lib().sre_get_md_state(self._h, C.byref(s)))
        return s.as_dict()

    def set_md_state(self, **kw):
        s = _lib.MDState()
        check(lib().sre_get_md_state(self._h, C.byref(s)))
        for k, v in kw.items():
            setattr(s, k, v)
        check(lib().sre_set_md_state(self._h, C.byref(s)))


def forces_static(xyz, box, rc_z, skin=0.3, types=None, eps=None, sigma=None,
                  n_types=1, stream=None, forces=None):
    """sre_forces_static: one force evaluation (rebuild + mixed-precision
    compute) with no device allocation by the library.  xyz: CUDA float32
    tensor (N, 3 or 4); types: CUDA int32 tensor (N,) or None; forces: CUDA
    float64 tensor (N, 3) to fill (allocated here if None).  Returns
    (forces, stats[4] = energy, virial xx, yy, zz)."""
    import torch
    assert xyz.is_cuda and xyz.dtype == torch.float32 and xyz.is_contiguous()
    n = xyz.shape[0]
    if types is not None:
        assert types.is_cuda and types.dtype == torch.int32 and types.is_contiguous()
    if forces is None:
        forces = torch.empty((n, 3), dtype=torch.float64, device=xyz.device)
    assert forces.is_cuda and forces.dtype == torch.float64 and forces.is_contiguous()
    nt = int(n_types)
    e = np.ascontiguousarray(np.ones((nt, nt)) if eps is None else eps, np.float32)
    s = np.ascontiguousarray(np.ones((nt, nt)) if sigma is None else sigma, np.float32)
    b = np.ascontiguousarray(np.asarray(box, np.float32).ravel()[:3])
    rc = np.ascontiguousarray(np.asarray(rc_z, np.float32).ravel())
    st = np.zeros(4)
    check(lib().sre_forces_static(n, C.c_void_p(xyz.data_ptr()), xyz.shape[1],
                                  C.c_void_p(types.data_ptr()) if types is not None else None,
                                  _ptr(b, C.c_float), nt, _ptr(e, C.c_float), _ptr(s, C.c_float),
                                  rc.size, _ptr(rc, C.c_float), float(skin),
                                  C.c_void_p(stream) if stream else None,
                                  C.c_void_p(forces.data_ptr()), _ptr(st, C.c_double)))
    return forces, st


def gen_fcc(ncell, a, jitter, seed):
    n = 4 * ncell ** 3
    out = np.empty((n, 4), np.float32)
    check(lib().sre_gen_fcc(int(ncell), float(a), float(jitter), int(seed),
                            _ptr(out, C.c_float)))
    return out


def gen_slab(n, box, h, rho_l, dmin, seed):
    out = np.empty((n, 4), np.float32)
    check(lib().sre_gen_slab(int(n), float(box[0]), float(box[1]), float(box[2]), float(h),
                             float(rho_l), float(dmin), int(seed), _ptr(out, C.c_float)))
    return out
