"""CPU checks of the boundary and the host logic: the C-ABI library loads and
exports every entry point declared in include/sre.h, the synthetic workload
generators are deterministic and meet their geometric contracts, and the
engine refuses to run without a GPU (no CPU fallback exists)."""
import ctypes
import os
import re

import numpy as np
import pytest

from sre import _lib, configs, gen_fcc, gen_slab

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "sre.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sre_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 30
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_last_error_without_gpu_calls():
    L = _lib.lib()
    assert L.sre_version().decode().startswith("sre")
    assert isinstance(L.sre_last_error(), bytes)


def test_invalid_arguments_are_rejected_on_host():
    L = _lib.lib()
    assert L.sre_create(None, None) == _lib.ERR_INVALID
    opt = _lib.Options(0, 0.3, 0, None, -1)
    h = ctypes.c_void_p()
    assert L.sre_create(ctypes.byref(h), ctypes.byref(opt)) == _lib.ERR_INVALID
    assert b"n_types" in L.sre_last_error()


def test_engine_needs_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from sre import Engine, MDExcptCuda
    with pytest.raises(MDExcptCuda):
        Engine()


def test_fcc_generator():
    x = gen_fcc(4, 1.5, 0.05, 3)
    assert x.shape == (256, 4) and x.dtype == np.float32
    assert np.all(x[:, :3] >= 0) and np.all(x[:, :3] < 6.0)
    x2 = gen_fcc(4, 1.5, 0.05, 3)
    assert np.array_equal(x, x2)
    x0 = gen_fcc(4, 1.5, 0.0, 3)
    d = np.abs(x[:, :3] - x0[:, :3])
    d = np.minimum(d, 6.0 - d)
    assert d.max() <= 0.05 + 1e-6
    # the unjittered lattice is FCC: every site has 12 nearest neighbours at a/sqrt(2)
    dv = x0[None, :, :3] - x0[:, None, :3]
    dv -= np.round(dv / 6.0) * 6.0
    r = np.sqrt((dv ** 2).sum(-1))
    assert np.all(((np.abs(r - 1.5 / np.sqrt(2)) < 1e-5).sum(1)) == 12)


def test_slab_generator_contract():
    box = (10.0, 10.0, 30.0)
    h = configs.slab_half_width(1500, box)
    x = gen_slab(1500, box, h, 0.73, 0.85, 11)
    assert np.array_equal(x, gen_slab(1500, box, h, 0.73, 0.85, 11))
    assert not np.array_equal(x, gen_slab(1500, box, h, 0.73, 0.85, 12))
    b = np.array(box, np.float32)
    assert np.all(x[:, :3] >= 0) and np.all(x[:, :3] < b)
    dv = x[None, :, :3].astype(np.float64) - x[:, None, :3]
    dv -= np.round(dv / b) * b
    r2 = (dv ** 2).sum(-1) + np.eye(len(x)) * 1e9
    assert r2.min() >= 0.85 ** 2 * (1 - 1e-6)
    n_liq = int(round(0.73 * 10 * 10 * 2 * h))
    zc = x[:, 2] - 15.0
    assert np.sum(np.abs(zc) < h) >= n_liq
    # the vapour holds the rest at ~ rho_v
    rho_v = (1500 - n_liq) / (10 * 10 * (30 - 2 * h))
    assert abs(rho_v - 0.02) < 0.01


def test_slab_geometry_resolution():
    assert abs(configs.slab_half_width(64000, (30, 30, 120)) - 48.388) < 1e-3
    t = configs.rc_profile(256, 120.0, 48.388, 2.5, 5.0).ravel()
    assert t.shape == (256,) and t[128] == np.float32(2.5) and t[0] == np.float32(5.0)
    assert np.all(np.diff(t[:128]) <= 0) and np.all(np.diff(t[128:]) >= 0)


def test_density_histogram_counts_atoms():
    x = gen_fcc(3, 2.0, 0.0, 1)
    rho = configs.density_histogram(x, np.zeros(len(x), np.int32), np.array([6, 6, 6.0]), 3, 1)
    assert abs(rho.sum() * (6 * 6 * 2) - len(x)) < 1e-9


def test_replica_workload_sizes():
    w = configs.replicas(count=6)
    assert len(w.n_atoms) == 6 and w.xyz.shape[0] == sum(w.n_atoms)
    assert all(2000 <= n <= 32000 for n in w.n_atoms)
    for r in range(6):
        L = w.boxes[r]
        assert 3.0 <= L[2] / L[0] <= 5.0
        # minimum image stays unambiguous for rc_max + skin
        assert 2 * (4.0 + w.skin) < L[0]
