"""GPU parity of the hot path against the FP64 oracle, through the C ABI.

Pair sets (neighbor lists, rc assignment) must be bit-exact; forces, energy
and virial of the mixed-precision path must meet the tolerances below against
the FP64 oracle; the FP64 GPU form must agree to round-off.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine, configs

pytestmark = pytest.mark.gpu

# Mixed precision (FP32 pair math): max_i |F_i - F_i^64| / RMS_i |F_i^64|
TOL_F_MIXED = 1e-5
TOL_E_MIXED = 1e-6        # |E - E^64| / sum_i |E_i^64|
TOL_F_FP64 = 1e-12
TOL_E_FP64 = 1e-12


def engine_for(w, precision="mixed", tail=False, **tuning):
    e = Engine(n_types=w.n_types, skin=w.skin, precision=precision)
    if tuning:
        e.set_tuning(**tuning)
    w.setup(e, tail=tail)
    e.rebuild()
    e.compute(stats=True, tail=tail)
    return e


def oracle_for(w, r=0, tail=False, exact=False):
    sl = w.replica_slice(r)
    xyz, types, box = w.xyz[sl], w.types[sl], w.boxes[r]
    rc, _ = O.assign_rc(xyz, box, w.rc_tables[r])
    off, j = O.build_pairs(xyz, box, rc, w.skin)
    tl = None
    if tail:
        tl = O.tail_atoms(xyz, box, types, w.eps, w.sigma, w.rc_tables[r].ravel(), w.rho[r],
                          exact=exact)
    F, E, W, nin = O.forces(xyz, box, types, w.eps, w.sigma, rc, off, j, tail=tl)
    return dict(rc=rc, off=off, j=j, F=F, E=E, W=W, nin=nin)


def force_err(F, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    d = np.sqrt(((F - ref) ** 2).sum(1))
    return d.max() / rms, np.sqrt((d ** 2).mean()) / rms


@pytest.fixture(scope="module")
def fcc():
    return configs.plain_fcc()


@pytest.fixture(scope="module")
def slab16():
    # a 16000-atom slab of the same geometry family (quick to check)
    return configs.slab(seed=11, n=16000, box=(20.0, 20.0, 75.0), nz=128, name="slab_16k")


@pytest.fixture(scope="module")
def slab64():
    return configs.slab()


@pytest.mark.parametrize("build_mode", [0, 1])
@pytest.mark.parametrize("wname", ["fcc", "slab16", "slab64"])
def test_pairs_and_rc_bit_exact(torch_cuda, wname, build_mode, request):
    """Both list builders (per-atom search, 64-atom tiles) give the oracle's pairs."""
    w = request.getfixturevalue(wname)
    e = engine_for(w, build_mode=build_mode)
    ref = oracle_for(w)
    assert np.array_equal(e.rcut(), ref["rc"])
    off, j = e.export_pairs()
    assert np.array_equal(off, ref["off"])
    assert np.array_equal(j, ref["j"])
    nl, nin, mx = e.neighbor_count()
    assert nl == ref["off"][-1]
    assert nin == ref["nin"]


@pytest.mark.parametrize("wname", ["fcc", "slab16", "slab64"])
def test_forces_energy_virial_mixed(torch_cuda, wname, request):
    w = request.getfixturevalue(wname)
    e = engine_for(w)
    ref = oracle_for(w)
    F = e.forces()
    emax, erms = force_err(F, ref["F"])
    print(f"{w.name}: max rel force err {emax:.3e}, rms {erms:.3e}")
    assert emax <= TOL_F_MIXED
    st = e.stats()[0]
    E64 = ref["E"].sum()
    scale = np.abs(ref["E"]).sum()
    assert abs(st[0] - E64) / scale <= TOL_E_MIXED
    W64 = ref["W"].sum(0)
    wscale = np.abs(ref["W"]).sum(0)
    assert np.all(np.abs(st[1:] - W64) / wscale <= TOL_E_MIXED)
    a = e.atom_stats()
    assert np.abs(a[:, 0] - ref["E"]).max() <= 1e-4 * np.abs(ref["E"]).max()


@pytest.mark.parametrize("wname", ["fcc", "slab16"])
def test_fp64_form_matches_oracle(torch_cuda, wname, request):
    w = request.getfixturevalue(wname)
    e = engine_for(w, precision="fp64")
    ref = oracle_for(w)
    F = e.forces()
    emax, _ = force_err(F, ref["F"])
    assert emax <= TOL_F_FP64
    st = e.stats()[0]
    assert abs(st[0] - ref["E"].sum()) / np.abs(ref["E"]).sum() <= TOL_E_FP64


def test_forces_bitwise_independent_of_tuning(torch_cuda, slab16):
    base = None
    # (lanes, build lanes, cells xy, z, block, build mode, XCD map, unroll,
    #  persistent wavefronts per SIMD)
    cases = [(4, 8, 0.5, 0.5, 256, 0, 0, 4, 0), (1, 4, 0.35, 0.25, 64, 0, 1, 1, 0),
             (2, 16, 0.7, 1.0, 128, 1, 0, 2, 0), (8, 2, 1.0, 0.35, 128, 1, 1, 1, 0),
             (16, 1, 0.5, 0.5, 256, 1, 1, 2, 0), (4, 8, 1.0, 0.35, 64, 1, 1, 4, 0),
             (4, 4, 1.0, 0.35, 256, 0, 1, 2, 0), (8, 4, 1.0, 0.35, 256, 0, 1, 1, 0),
             (4, 4, 1.0, 0.35, 256, 0, 1, 2, 1), (2, 2, 0.7, 0.35, 128, 0, 0, 1, 3),
             (8, 8, 1.0, 0.5, 64, 0, 1, 4, 16)]
    for fl, bl, cxy, cz, blk, bm, xr, fu, pw in cases:
        e = engine_for(slab16, force_lanes=fl, build_lanes=bl, cell_xy=cxy, cell_z=cz,
                       block=blk, build_mode=bm, xcd_remap=xr, force_unroll=fu,
                       persistent=pw)
        f = e.forces_fixed()
        if base is None:
            base = f
        else:
            assert np.array_equal(f, base), (fl, bl, cxy, cz, blk, bm, xr, fu, pw)


def test_fp64_pipelined_equals_plain(torch_cuda, slab16):
    """The FP64 form in the mixed kernel's pipelined structure (force_unroll 1)
    sums each lane's pairs in the same order as the plain FP64 kernel, so
    forces and statistics agree bit for bit."""
    out = []
    for fu in (4, 1):
        e = engine_for(slab16, precision="fp64", force_unroll=fu)
        out.append((e.forces(), e.stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
