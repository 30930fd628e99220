"""The reference's driver sequence (H/app/lj.nhc-xtc.cu:72-113, 169-188) through
the reference-named API, on the reference's own 16000-atom slab configuration
(conf.16000.gro), and the torch.autograd bridge of the adjoint."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sre import reference_api as R

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_driver_sequence_on_conf16000(torch_cuda):
    d = np.load(os.path.join(GOLD, "conf16000.npz"))
    xyz, box = d["xyz"], d["box"]
    rcmax, skin, nz = 5.0, 0.49, 150
    sys = R.MDSystem()
    sys.initConfig(xyz, box)
    ljparam = R.LennardJones6_12Parameter(1.0, 1.0, 0.0, rcmax)
    nb = R.SystemNonBondedInteraction()
    nb.add(0, 0, ljparam)
    sys.initTopology()
    sys.initDeviceData()
    # an adaptive profile along x (the slab normal of this configuration)
    table = np.where(np.abs((np.arange(nz) + 0.5) - 75) < 20, 3.5, 5.0).astype(np.float32)
    arc = R.AssignRCut()
    arc.reinit(sys, (nz, 1, 1))
    arc.getRCut(table)
    arc.assign(sys)
    clist = R.CellList(sys, rcmax + skin)
    nlist = R.NeighborList(nb, sys, rcmax + skin, skin)
    sys.normalizeDeviceData()
    clist.rebuild(sys)
    nlist.rebuild(sys, clist)
    inter = R.InteractionEngine(sys)
    inter.registNonBondedInteraction(nb)
    st = R.MDStatistic(sys)
    st.clearDevice()
    inter.clearInteraction(sys)
    inter.applyNonBondedInteraction(sys, nlist, st)
    st.updateHost()

    rc, _ = O.assign_rc(xyz, box, table.reshape(nz, 1, 1))
    off, j = O.build_pairs(xyz, box, rc, skin)
    F, E, W, _ = O.forces(xyz, box, None, np.ones((1, 1)), np.ones((1, 1)), rc, off, j)
    assert nlist.calSumNeighbor() == off[-1]
    Fg = np.stack([sys.forcx.cpu().numpy(), sys.forcy.cpu().numpy(),
                   sys.forcz.cpu().numpy()], 1).astype(np.float64)
    rms = np.sqrt((F ** 2).sum(1).mean())
    # float32 accumulation targets (the reference's own force arrays are FP32)
    assert np.sqrt(((Fg - F) ** 2).sum(1)).max() / rms < 1e-5
    assert abs(st.nonBondedEnergy() - E.sum()) / np.abs(E).sum() < 1e-6
    np.testing.assert_allclose(st.virial(), W.sum(0), rtol=1e-6)
    # the skin test of the reference loop: nothing moved
    assert R.Displacement_max(sys).calMaxDisplacemant(sys) is False


def test_autograd_bridge_matches_finite_differences(torch_cuda):
    import torch
    from sre import Engine, configs
    from sre.autograd import lj_energy_forces
    w = configs.slab(seed=41, n=6000, box=(16.0, 16.0, 50.0), nz=48, binary=True,
                     name="binary_6k")
    e = Engine(n_types=2, skin=w.skin, precision="fp64")
    w.setup(e, tail=True)
    e.rebuild()
    eps = torch.tensor(w.eps, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(w.sigma, dtype=torch.float64, requires_grad=True)
    E, F = lj_energy_forces(e, eps, sig, None, True)
    Ft = torch.from_numpy(np.random.default_rng(2).normal(size=(w.n, 3))).cuda() + F.detach()
    loss = 0.1 * E.sum() + ((F - Ft) ** 2).sum()
    loss.backward()

    def L(ep, sg):
        with torch.no_grad():
            E2, F2 = lj_energy_forces(e, ep, sg, None, True)
            return float(0.1 * E2.sum() + ((F2 - Ft) ** 2).sum())

    # a small step: the loss is quadratic in F and F ~ sigma^12, so the
    # third-order term of the central difference is large at h = 2^-12
    h = 2.0 ** -17
    for a, b in [(0, 0), (0, 1), (1, 1)]:
        for which, grad in (("eps", eps.grad), ("sig", sig.grad)):
            vals = []
            for sgn in (1, -1):
                ep, sg = eps.detach().clone(), sig.detach().clone()
                p = ep if which == "eps" else sg
                p[a, b] += sgn * h
                p[b, a] = p[a, b]
                # parameters travel as FP32: step by what FP32 represents
                vals.append((L(ep, sg), float(np.float32(p[a, b].item()))))
            fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
            # the forward is a function of the symmetrised matrices: the
            # shared parameter's derivative is split over both entries
            g = float(grad[a, b] + grad[b, a]) if a != b else float(grad[a, a])
            assert abs(fd - g) <= 2e-4 * max(1.0, abs(g)), (which, a, b, fd, g)


def test_reference_adaptive_cutoff_sequence(torch_cuda, tmp_path):
    """The adaptive part of the reference driver (H/app/lj.nhc-xtc.cu:84-100,
    228-245) through the reference-named API: density grid, rc candidates,
    deposits, calError / calRCut / refineRCut x2, getRCut into AssignRCut,
    then the force path with the new lattice; and save_rc / load_rc."""
    from oracle import arc as A
    d = np.load(os.path.join(GOLD, "conf16000.npz"))
    xyz, box = d["xyz"], d["box"]
    sys = R.MDSystem()
    sys.initConfig(xyz, box)
    nb = R.SystemNonBondedInteraction()
    nb.add(0, 0, R.LennardJones6_12Parameter(1.0, 1.0, 0.0, 7.5))
    sys.initTopology()
    sys.initDeviceData()
    dp = R.DensityProfile_PiecewiseConst()
    dp.reinit(*box, 1.0)
    arc = R.AdaptRCut()
    arc.reinit(3.0, 7.5, 0.25, dp)
    assign = R.AssignRCut()
    assign.reinit(sys, arc)
    assign.uniform(5.0)
    assign.assign(sys)
    for _ in range(4):
        dp.deposite(sys.coord, sys.numAtom)
    dp.calculate()
    arc.calError(dp)
    arc.calRCut(0.0045)
    for _ in range(2):
        arc.refineRCut()
    assign.getRCut(arc)
    assign.assign(sys)
    nlist = R.NeighborList(nb, sys, 7.5 + 0.49, 0.49)
    nlist.rebuild(sys)
    inter = R.InteractionEngine(sys)
    inter.registNonBondedInteraction(nb)
    st = R.MDStatistic(sys)
    inter.clearInteraction(sys)
    inter.applyNonBondedInteraction(sys, nlist, st)
    st.updateHost()

    # the oracle's choice from the same density
    dims = (dp.getNx(), dp.getNy(), dp.getNz())
    counts = np.zeros(dims)
    for _ in range(4):
        A.deposit(counts, xyz, box, dims)
    rho = A.density(counts, 4, box, dims)
    assert np.array_equal(dp.profile, rho)
    rcs = A.rc_list(3.0, 7.5, 0.25)
    table = arc.getRCut()
    assert set(np.unique(table)) <= set(rcs)
    rc_atoms = sys.engine.rcut()
    want, _ = O.assign_rc(xyz, box, table.astype(np.float32))
    assert np.array_equal(rc_atoms, want)
    off, j = O.build_pairs(xyz, box, want, 0.49)
    F, E, _, _ = O.forces(xyz, box, None, np.ones((1, 1)), np.ones((1, 1)), want, off, j)
    assert abs(st.nonBondedEnergy() - E.sum()) / np.abs(E).sum() < 1e-6
    # save_rc / load_rc round trip (H/src/AdaptRCut.cu:627-726 text format)
    path = str(tmp_path / "rcut.save")
    arc.save_rc(path)
    arc2 = R.AdaptRCut()
    dp2 = R.DensityProfile_PiecewiseConst()
    dp2.reinit(*box, 1.0)
    arc2.reinit(3.0, 7.5, 0.25, dp2)
    arc2.load_rc(path)
    assert np.array_equal(arc2.getProfileIndex(), arc.getProfileIndex())
    assert np.array_equal(arc2.getRCut(), table)
    # trajectories and configuration as the driver writes them
    from sre import io
    assign.init_write(str(tmp_path / "rcut.rtj"))
    dp.init_write(str(tmp_path / "density.dtj"))
    for t in (1.0, 2.0):
        assign.write(t)
        dp.write(t)
    assign.end_write()
    dp.end_write()
    _, d1, times, fr = io.read_grid_trajectory(str(tmp_path / "rcut.rtj"), np.float32)
    assert d1 == dims and np.array_equal(times, [1.0, 2.0])
    assert np.array_equal(fr[1], table.astype(np.float32))
    _, _, _, fr = io.read_grid_trajectory(str(tmp_path / "density.dtj"), np.float64)
    assert np.array_equal(fr[0], rho)
    sys.writeHostDataGro(str(tmp_path / "confout.gro"), 10, 0.05)
    x2, b2 = io.read_gro(str(tmp_path / "confout.gro"))
    assert np.abs(x2 - xyz).max() < 5.1e-4 and np.array_equal(b2, box)
    # the driver's xtc trajectory (H/app/lj.nhc-xtc.cu:155-158, 260, 275)
    sys.initWriteXtc(str(tmp_path / "traj.xtc"))
    for i in range(3):
        sys.writeHostDataXtc(i, 0.005 * i)
    sys.endWriteXtc()
    frames = io.read_xtc(str(tmp_path / "traj.xtc"))
    assert [f[0] for f in frames] == [0, 1, 2]
    for step, t, b, x3, prec in frames:
        assert prec == 1000.0 and np.array_equal(np.diag(b), box.astype(np.float32))
        # half a lattice step, plus the FP32 rounding of the decoding
        assert np.abs(x3 - xyz).max() <= 0.5e-3 + 3e-7 * np.abs(xyz).max()
