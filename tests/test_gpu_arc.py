"""GPU parity of the adaptive-cutoff producer (sre_arc_*) with the oracle.

Kernel tables, the error of every candidate (and its parts), the chosen and
widened candidate per cell, the deposited density, and the full pipeline on
the reference's own grid (151 x 21 x 21, 19 candidates) against both the
oracle and the reference's slab-run output.
"""
import os
import warnings

import numpy as np
import pytest
from scipy import integrate

from oracle import arc as A
from oracle import oracle as O
from sre import Arc, Engine

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(autouse=True)
def _quiet():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        yield


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def small(torch_cuda):
    box = (12.0, 9.0, 15.0)
    arc = Arc(box, 1.0)
    rcs = arc.set_candidates(3.0, 5.0, 0.5)
    assert arc.dims == A.grid_dims(box, 1.0) == (13, 9, 15)
    assert np.allclose(rcs, A.rc_list(3.0, 5.0, 0.5))
    S1, G, kv = A.kernels(box, arc.dims, rcs)
    rng = np.random.default_rng(0)
    x, y, z = np.meshgrid(*[np.arange(n) for n in arc.dims], indexing="ij")
    rho = (0.4 + 0.3 * np.tanh(np.sin(2 * np.pi * z / arc.dims[2]) * 3)
           + 0.05 * rng.random(arc.dims))
    return dict(box=box, arc=arc, rcs=rcs, S1=S1, G=G, kv=kv, rho=rho)


def test_kernel_tables(small):
    S1, G = small["arc"].kernels()
    for c in range(len(small["rcs"])):
        assert rel(S1[c], small["S1"][c]) < 1e-10
        assert rel(G[c], small["G"][c]) < 1e-10


def test_errors_and_parts(small):
    arc = small["arc"]
    arc.set_density(small["rho"])
    assert np.array_equal(arc.density(), small["rho"])
    arc.cal_error(keep_parts=True)
    err, parts = arc.errors(parts=True)
    ref, e1, e2 = A.cal_error(small["rho"], small["box"], small["S1"], small["G"],
                              small["kv"], parts=True)
    assert rel(err, ref) < 1e-10
    assert rel(parts[..., 0], e1) < 1e-10
    for a in range(3):
        assert rel(parts[..., 1 + a], e2[:, a]) < 1e-9


def test_rcut_and_refine(small):
    arc = small["arc"]
    arc.set_density(small["rho"])
    arc.cal_error()
    ref = A.cal_error(small["rho"], small["box"], small["S1"], small["G"], small["kv"])
    for prec in (0.004, 0.01, 0.03):
        arc.cal_rcut(prec)
        want = A.cal_rcut(ref, prec)
        rc, err, idx = arc.result()
        assert np.array_equal(idx, want)
        for n in (1, 2):
            arc.refine()
            want = A.refine(want)
            rc, err, idx = arc.result()
            assert np.array_equal(idx, want)
        assert np.array_equal(rc, small["rcs"][want].astype(np.float32))
        assert rel(err, np.take_along_axis(ref, want[None], 0)[0]) < 1e-10


def test_deposit_matches_oracle(torch_cuda):
    g = np.load(os.path.join(GOLDEN, "conf16000.npz"))
    xyz, box = g["xyz"], g["box"].astype(np.float64)
    arc = Arc(box, 1.0)
    pos = torch_cuda.from_numpy(np.c_[xyz, np.zeros(len(xyz), np.float32)]).cuda()
    arc.deposit(pos)
    arc.deposit(pos)
    counts = np.zeros(arc.dims)
    A.deposit(counts, xyz, box, arc.dims)
    A.deposit(counts, xyz, box, arc.dims)
    got, want = arc.density(), A.density(counts, 2, box, arc.dims)
    assert np.array_equal(got, want), (np.count_nonzero(got != want),
                                       np.abs(got - want).max())
    arc.clear()
    with pytest.raises(Exception):
        arc.density()


def test_reference_grid_pipeline(torch_cuda):
    """The reference's own grid (151 x 21 x 21, rc 3.0..7.5 step 0.25, target
    0.0045, two refinements), fed its run-averaged density along x: the same
    candidates as the oracle's 1-D computation in every cell, and the
    reference's last rc update in >= 85 % of the liquid/interface bins."""
    fx = np.load(os.path.join(GOLDEN, "arc_t085.npz"))
    d, ex = fx["a_density"], fx["error_x"]
    box = (150.0, 21.0, 21.0)
    arc = Arc(box, 1.0)
    rcs = arc.set_candidates(3.0, 7.5, 0.25)
    nx = arc.dims[0]
    xc = (np.arange(nx) + 0.5) * 150.0 / nx
    prof = np.interp(xc, d[:, 0], d[:, 1], period=150.0)
    arc.set_density(np.broadcast_to(prof[:, None, None], arc.dims))
    arc.cal_error()
    arc.cal_rcut(0.0045)
    arc.refine(2)
    rc, err, idx = arc.result()
    dims1 = (nx, 1, 1)
    S1, G, kv = A.kernels(box, dims1, rcs)
    ref = A.cal_error(prof.reshape(dims1), box, S1, G, kv)
    want = A.refine(A.refine(A.cal_rcut(ref, 0.0045)))
    assert np.array_equal(idx, np.broadcast_to(want, arc.dims))
    e_col = np.take_along_axis(ref, want[None], 0)[0, :, 0, 0]
    assert rel(err[:, 3, 7], e_col) < 1e-9
    dense = prof > 0.05
    assert np.mean(rc[dense, 0, 0] == ex[dense, 2]) >= 0.85


def test_apply_installs_rc_lattice(torch_cuda):
    g = np.load(os.path.join(GOLDEN, "conf16000.npz"))
    xyz, box = g["xyz"], g["box"]
    e = Engine(n_types=1, skin=0.3)
    e.set_system([len(xyz)], box.reshape(1, 3))
    e.set_types(np.zeros(len(xyz), np.int32))
    e.set_pair_params(np.ones((1, 1)), np.ones((1, 1)))
    pos = torch_cuda.from_numpy(np.c_[xyz, np.zeros(len(xyz), np.float32)]).cuda()
    torch_cuda.cuda.synchronize()
    e.bind_positions(pos)
    arc = Arc(box.astype(np.float64), 1.0)
    arc.set_candidates(3.0, 6.0, 0.5)
    for _ in range(3):
        arc.deposit(e)
    arc.cal_error()
    arc.cal_rcut(0.01)
    arc.refine(1)
    rc_tab, _, _ = arc.result()
    arc.apply(e)
    e.rebuild()
    rc_atoms = e.rcut()
    want, _ = O.assign_rc(xyz, box, rc_tab)
    assert np.array_equal(rc_atoms, want)
    assert rc_tab.min() < rc_tab.max()          # the slab is inhomogeneous


def test_pressure_correction(small):
    arc = small["arc"]
    box, rcs = small["box"], small["rcs"]
    I1, I5 = arc.pressure_integrals()
    # spot-check the integrals against the oracle's exact-Bessel quadrature
    kx, ky, kz = A.k_grid(box, arc.dims)
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    for cell in [(1, 0, 0), (3, 2, 5), (6, 4, 7)]:
        for c in (0, len(rcs) - 1):
            assert abs(I1[(c,) + cell] / A._int_bessel(0.5, k[cell], rcs[c], 30.0) - 1) < 1e-10
            assert abs(I5[(c,) + cell] / A._int_bessel(2.5, k[cell], rcs[c], 30.0) - 1) < 1e-9
    K = A.pressure_kernels(box, arc.dims, rcs)
    arc.set_density(small["rho"])
    arc.cal_error()
    for prec, n_ref in ((0.01, 0), (0.004, 2)):
        arc.cal_rcut(prec)
        arc.refine(n_ref)
        _, _, idx = arc.result()
        want = A.pressure_correction(small["rho"], box, K, idx.astype(np.int64))
        got = arc.pressure()
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_reference_pressure_correction_class(torch_cuda):
    from sre import reference_api as R
    g = np.load(os.path.join(GOLDEN, "conf16000.npz"))
    box = g["box"]
    sys = R.MDSystem()
    sys.initConfig(g["xyz"], box)
    sys.initTopology()
    sys.initDeviceData()
    dp = R.DensityProfile_PiecewiseConst()
    dp.reinit(*box, 1.0)
    arc = R.AdaptRCut()
    arc.reinit(3.0, 4.0, 0.5, dp)        # few candidates: the oracle's kernels are slow
    pc = R.PressureCorrection(arc, dp)
    dp.deposite(sys)
    dp.calculate()
    arc.calError(dp)
    arc.calRCut(0.01)
    arc.refineRCut()
    pc.correction(arc, dp)
    K = A.pressure_kernels(box, dp.arc.dims, arc.getRcList())
    want = A.pressure_correction(dp.profile, box, K, arc.getProfileIndex().astype(np.int64))
    got = np.array([pc.pxx, pc.pyy, pc.pzz])
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()
    assert got[0] != got[1]                      # the configuration is not planar-symmetric


def test_force_grid_and_correction(torch_cuda):
    """One candidate (uniform rc): the device's mean error force grid equals
    the oracle's, and SRE_FCORR adds each atom's cell value (reference FP32
    cell rule) to the pair forces, nothing else changing."""
    from sre import configs
    w = configs.slab(seed=31, n=12000, box=(20.0, 20.0, 60.0), nz=60, name="slab_12k")
    box = tuple(float(v) for v in w.boxes[0])
    arc = Arc(box, 1.0)
    arc.set_candidates(4.0, 4.0, 1.0)
    pos = torch_cuda.from_numpy(w.xyz.copy()).cuda()
    for _ in range(2):
        arc.deposit(pos)
    arc.cal_error(keep_parts=True)
    arc.cal_rcut(1.0)
    grid = arc.force_grid()
    rho = arc.density()
    S1, G, kv = A.kernels(box, arc.dims, np.array([4.0]))
    _, _, e2 = A.cal_error(rho, box, S1, G, kv, parts=True)
    want = np.moveaxis(e2[0], 0, -1).astype(np.float32)
    assert np.abs(grid - want).max() < 1e-6 * np.abs(want).max()
    # a transverse component has no kx = 0 (ky = 0) Fourier part, so its
    # average along its own axis vanishes; the z profile carries the slab
    zmax = np.abs(grid[..., 2].mean(axis=(0, 1))).max()
    assert np.abs(grid[..., 0].mean(axis=0)).max() < 1e-5 * zmax
    assert np.abs(grid[..., 1].mean(axis=1)).max() < 1e-5 * zmax
    assert zmax > 0.01

    e = Engine(n_types=1, skin=w.skin)
    w.setup(e, tail=False)
    torch_cuda.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.compute(stats=True)
    f0, s0 = e.forces_fixed(), e.stats()
    arc.apply_force_correction(e)
    e.compute(stats=True, fcorr=True)
    f1, s1 = e.forces_fixed(), e.stats()
    assert np.array_equal(s0, s1)
    # oracle cell rule (AssignForceCorr::assign)
    xyz = w.xyz[:, :3]
    L = np.asarray(box, np.float32)
    cells = []
    for d in range(3):
        x = xyz[:, d].astype(np.float32)
        x = np.where(x < 0, (x + L[d]).astype(np.float32),
                     np.where(x >= L[d], (x - L[d]).astype(np.float32), x))
        c = ((x * np.float32(arc.dims[d])).astype(np.float32) / L[d]).astype(np.int64)
        cells.append(np.clip(c, 0, arc.dims[d] - 1))
    add = grid[cells[0], cells[1], cells[2]].astype(np.float64)
    fix = np.round(add * 2.0 ** 32).astype(np.int64)
    assert np.array_equal(f1 - f0, fix)
    e.compute(stats=True)                    # without the flag: the pair forces again
    assert np.array_equal(e.forces_fixed(), f0)


def test_reference_force_corr_classes(torch_cuda):
    from sre import reference_api as R
    g = np.load(os.path.join(GOLDEN, "conf16000.npz"))
    xyz, box = g["xyz"], g["box"]
    sys = R.MDSystem()
    sys.initConfig(xyz, box)
    nb = R.SystemNonBondedInteraction()
    nb.add(0, 0, R.LennardJones6_12Parameter(1.0, 1.0, 0.0, 5.0))
    sys.initTopology()
    sys.initDeviceData()
    dp = R.DensityProfile_PiecewiseConst()
    dp.reinit(*box, 1.0)
    fc = R.ForceCorr()
    fc.reinit(5.0, dp)
    afc = R.AssignForceCorr()
    afc.reinit(sys, fc)
    afc.zero()
    afc.assign(sys)
    assign = R.AssignRCut()
    assign.reinit(sys, (1, 1, 1))
    assign.uniform(5.0)
    assign.assign(sys)
    nlist = R.NeighborList(nb, sys, 5.3, 0.3)
    nlist.rebuild(sys)
    inter = R.InteractionEngine(sys)
    inter.registNonBondedInteraction(nb)
    inter.clearInteraction(sys)
    inter.applyNonBondedInteraction(sys, nlist)
    f_zero = sys.forcx.cpu().numpy().copy()
    dp.deposite(sys)
    dp.calculate()
    fc.calError(dp)
    afc.getForceCorr(fc)
    afc.assign(sys)
    inter.clearInteraction(sys)
    inter.applyNonBondedInteraction(sys, nlist)
    f_corr = sys.forcx.cpu().numpy()
    diff = f_corr - f_zero
    assert np.abs(fc.getForceCorrX()).max() > 0
    assert 0 < np.abs(diff).max() <= 1.01 * np.abs(fc.getForceCorrX()).max() + 1e-6
