"""ctypes binding of libsre.so (the C ABI declared in include/sre.h).

The library must exist (build with ``python md.error_amd/build.py`` or
``__graft_entry__.build()``); there is no fallback implementation.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsre.so")

# status codes (include/sre.h)
OK = 0
ERR_LIST_OVERFLOW = 1
ERR_CELL = 2
ERR_OUT_OF_BOX = 3
ERR_INVALID = 4
ERR_HIP = 5
ERR_BOX_TOO_SMALL = 6
ERR_FORCE_RANGE = 7
ERR_SKIN = 8
ERR_NOT_READY = 9
ERR_IO = 10
END_OF_FILE = 11

PREC_MIXED = 0
PREC_FP64 = 1
STATS = 1
TAIL = 2
PERTURB = 4
FCORR = 8
TAIL_BIN = 0
TAIL_EXACT = 1
THERMO_NONE = 0
THERMO_NHC2 = 1


class Options(C.Structure):
    _fields_ = [("n_types", C.c_int), ("skin", C.c_float), ("precision", C.c_int),
                ("stream", C.c_void_p), ("device", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("force_lanes", C.c_int), ("build_lanes", C.c_int), ("block", C.c_int),
                ("cell_xy", C.c_float), ("cell_z", C.c_float), ("build_mode", C.c_int),
                ("xcd_remap", C.c_int), ("force_unroll", C.c_int), ("persistent", C.c_int),
                ("sort_mode", C.c_int), ("force_mode", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MDState(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("xi1", C.c_double), ("xi2", C.c_double),
                ("vxi1", C.c_double), ("vxi2", C.c_double), ("nhc_energy", C.c_double),
                ("Q1", C.c_double), ("Q2", C.c_double), ("adapt_steps", C.c_int),
                ("rollbacks", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every exported entry point: name -> (restype, argtypes)
_vp = C.c_void_p
_f32p, _f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
_i32p, _i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
_st = C.c_int
SIGNATURES = {
    "sre_create": (_st, [C.POINTER(_vp), C.POINTER(Options)]),
    "sre_destroy": (None, [_vp]),
    "sre_last_error": (C.c_char_p, []),
    "sre_version": (C.c_char_p, []),
    "sre_set_system": (_st, [_vp, C.c_int, _i32p, _f32p]),
    "sre_set_types": (_st, [_vp, _i32p]),
    "sre_set_pair_params": (_st, [_vp, _f32p, _f32p, _f32p]),
    "sre_set_rc_table": (_st, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f32p]),
    "sre_set_density_profile": (_st, [_vp, C.c_int, C.c_int, _f64p]),
    "sre_set_tail_mode": (_st, [_vp, C.c_int]),
    "sre_upload_positions": (_st, [_vp, _f32p, C.c_int]),
    "sre_bind_positions": (_st, [_vp, _vp, C.c_int]),
    "sre_rebuild": (_st, [_vp]),
    "sre_update_positions": (_st, [_vp, C.POINTER(C.c_int)]),
    "sre_compute": (_st, [_vp, C.c_uint]),
    "sre_step": (_st, [_vp, C.c_uint, C.POINTER(C.c_int)]),
    "sre_get_forces": (_st, [_vp, _f64p, C.c_int]),
    "sre_add_forces_f32_soa": (_st, [_vp, _vp, _vp, _vp]),
    "sre_get_forces_fixed": (_st, [_vp, _i64p]),
    "sre_get_stats": (_st, [_vp, _f64p]),
    "sre_get_atom_stats": (_st, [_vp, _f64p]),
    "sre_neighbor_count": (_st, [_vp, _i64p, _i64p, _i32p]),
    "sre_export_pairs": (_st, [_vp, _i64p, _i32p]),
    "sre_get_rcut": (_st, [_vp, _f32p]),
    "sre_assign_rcut": (_st, [_vp]),
    "sre_adjoint": (_st, [_vp, _f64p, _vp, C.c_uint, _f64p, _f64p, _f64p]),
    "sre_get_tuning": (_st, [_vp, C.POINTER(Tuning)]),
    "sre_set_tuning": (_st, [_vp, C.POINTER(Tuning)]),
    "sre_autotune": (_st, [_vp, C.c_uint, C.c_char_p]),
    "sre_graph_run": (_st, [_vp, C.c_uint, C.c_int]),
    "sre_run_segments": (_st, [_vp, C.c_uint, C.c_int, C.c_int, _i64p, _f32p]),
    "sre_set_perturbation": (_st, [_vp, C.c_float, C.c_uint64, C.c_uint64, C.c_int]),
    "sre_check_skin_flag": (_st, [_vp]),
    "sre_last_rebuild_displacement": (_st, [_vp, _f32p]),
    "sre_perturb_positions": (_st, [_vp, C.c_float, C.c_uint64, C.c_uint64]),
    "sre_time_compute": (_st, [_vp, C.c_uint, C.c_int, _f64p]),
    "sre_time_rebuild": (_st, [_vp, C.c_int, _f64p]),
    "sre_synchronize": (_st, [_vp]),
    "sre_set_masses": (_st, [_vp, _f32p]),
    "sre_set_velocities": (_st, [_vp, _f32p]),
    "sre_get_velocities": (_st, [_vp, _f32p]),
    "sre_set_integrator": (_st, [_vp, C.c_double, C.c_int, C.c_double, C.c_double,
                                 C.c_int]),
    "sre_md_run": (_st, [_vp, C.c_int, C.c_int, C.c_uint, C.POINTER(C.c_int)]),
    "sre_get_md_state": (_st, [_vp, C.POINTER(MDState)]),
    "sre_set_md_state": (_st, [_vp, C.POINTER(MDState)]),
    "sre_arc_create": (_st, [C.POINTER(_vp), _f64p, C.c_double, C.c_int, _vp]),
    "sre_arc_destroy": (None, [_vp]),
    "sre_arc_dims": (_st, [_vp, _i32p, C.POINTER(C.c_int)]),
    "sre_arc_set_candidates": (_st, [_vp, C.c_double, C.c_double, C.c_double]),
    "sre_arc_rc_list": (_st, [_vp, _f64p]),
    "sre_arc_clear": (_st, [_vp]),
    "sre_arc_deposit": (_st, [_vp, _vp, C.c_int, C.c_int]),
    "sre_arc_deposit_engine": (_st, [_vp, _vp, C.c_int]),
    "sre_arc_set_density": (_st, [_vp, _f64p]),
    "sre_arc_density": (_st, [_vp, _f64p]),
    "sre_arc_cal_error": (_st, [_vp, C.c_int]),
    "sre_arc_cal_rcut": (_st, [_vp, C.c_double]),
    "sre_arc_refine": (_st, [_vp]),
    "sre_arc_result": (_st, [_vp, _f32p, _f64p, _i32p]),
    "sre_arc_set_index": (_st, [_vp, _i32p]),
    "sre_arc_errors": (_st, [_vp, _f64p, _f64p]),
    "sre_arc_kernels": (_st, [_vp, _f64p, _f64p]),
    "sre_arc_apply": (_st, [_vp, _vp, C.c_int]),
    "sre_arc_pressure": (_st, [_vp, _f64p]),
    "sre_arc_force_grid": (_st, [_vp, _f32p]),
    "sre_arc_apply_force_correction": (_st, [_vp, _vp, C.c_int]),
    "sre_set_force_grid": (_st, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f32p]),
    "sre_arc_pressure_integrals": (_st, [_vp, _f64p, _f64p]),
    "sre_forces_static": (_st, [C.c_int, _vp, C.c_int, _vp, _f32p, C.c_int, _f32p, _f32p,
                                C.c_int, _f32p, C.c_float, _vp, _vp, _f64p]),
    "sre_xtc_open": (_st, [C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "sre_xtc_write": (_st, [_vp, C.c_int, C.c_int, C.c_float, _f32p, _f32p, C.c_float]),
    "sre_xtc_read": (_st, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float), _f32p,
                           _f32p, C.POINTER(C.c_float)]),
    "sre_xtc_natoms": (_st, [C.c_char_p, C.POINTER(C.c_int)]),
    "sre_xtc_close": (_st, [_vp]),
    "sre_gen_fcc": (_st, [C.c_int, C.c_double, C.c_double, C.c_uint64, _f32p]),
    "sre_gen_slab": (_st, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                           C.c_double, C.c_double, C.c_uint64, _f32p]),
}

_LIB = None


def _init_torch_first():
    """libsre.so runs on the system HIP runtime; PyTorch for ROCm carries a
    runtime of its own, and in a process that holds both, PyTorch finds no
    device if the system runtime initialised first.  So where PyTorch is
    importable its runtime starts before the library is loaded."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Loads libsre.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python md.error_amd/build.py`")
        _init_torch_first()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB
