"""The reference's object API for the short-range path, over one Engine.

Class and method names follow the reference's C++ interface so that a driver
written against it (H/app/lj.nhc-xtc.cu:72-113, 169-188) reads the same:

    sys = MDSystem(); sys.initConfig("conf.gro")
    ljparam = LennardJones6_12Parameter(1., 1., 0., rcmax)
    sysNbInter = SystemNonBondedInteraction(); sysNbInter.add(0, 0, ljparam)
    sys.initTopology(types); sys.initDeviceData()
    assign = AssignRCut(); assign.reinit(sys, (1, 1, nz)); assign.getRCut(table)
    assign.assign(sys)
    clist = CellList(sys, rlist); nlist = NeighborList(sysNbInter, sys, rlist, skin)
    nlist.rebuild(sys, clist)
    inter = InteractionEngine(sys); inter.registNonBondedInteraction(sysNbInter)
    st = MDStatistic(sys); inter.clearInteraction(sys)
    inter.applyNonBondedInteraction(sys, nlist, st)
    st.updateHost(); st.nonBondedEnergy()

Semantics kept: forces accumulate (+=) into FP32 SoA device arrays
``sys.forcx/forcy/forcz`` after ``clearInteraction`` zeroes them
(H/src/InteractionEngine_interface.cu:218-225, 1270-1273); energy and the
virial diagonal accumulate into the statistic (H/include/Statistic_interface.h).
Differences: the cell list needs no separate rebuild (binning is part of
``NeighborList.rebuild``), ``normalizeDeviceData`` is a no-op (wrapping is
exact), and rc assignment happens at every list rebuild.
"""
import numpy as np

from . import io as _io
from .arc import Arc
from .engine import Engine, NotReady


def read_gro(path):
    """GROMACS .gro: positions (columns 21-44) and the box (last line)."""
    return _io.read_gro(path)


class LennardJones6_12Parameter:
    """{epsilon, sigma, shift, rcut} (H/include/NonBondedInteraction.h:19-42)."""

    def __init__(self, epsilon=1.0, sigma=1.0, shift=0.0, rcut=2.5):
        self.reinit(epsilon, sigma, shift, rcut)

    def reinit(self, epsilon, sigma, shift, rcut):
        self.epsilon, self.sigma, self.shift, self.rc = (float(epsilon), float(sigma),
                                                        float(shift), float(rcut))

    def rcut(self):
        return self.rc


class SystemNonBondedInteraction:
    """Type-pair table (H/src/SystemNonBondedInteraction.cu:125-330)."""

    def __init__(self):
        self.table = {}

    def add(self, i, j, param):
        self.table[(min(i, j), max(i, j))] = param

    def numAtomTypes(self):
        return 1 + max(max(k) for k in self.table) if self.table else 0

    def maxRcut(self):
        return max(p.rcut() for p in self.table.values())

    def matrices(self):
        nt = self.numAtomTypes()
        eps = np.zeros((nt, nt), np.float32)
        sig = np.ones((nt, nt), np.float32)
        sh = np.zeros((nt, nt), np.float32)
        for (a, b), p in self.table.items():
            for x, y in ((a, b), (b, a)):
                eps[x, y], sig[x, y], sh[x, y] = p.epsilon, p.sigma, p.shift
        return eps, sig, sh


class MDSystem:
    """Positions, types and box; owns the engine (H/include/MDSystem_interface.h)."""

    def __init__(self):
        self.xyz = None
        self.box = None
        self.types = None
        self.engine = None
        self._skin = 0.3
        self._rc = None
        self._nb = None
        self.forcx = self.forcy = self.forcz = None

    @property
    def numAtom(self):
        return 0 if self.xyz is None else len(self.xyz)

    def initConfig(self, filename_or_xyz, box=None):
        if isinstance(filename_or_xyz, str):
            self.xyz, self.box, self.velo = _io.read_gro(filename_or_xyz, velocities=True)
        else:
            self.xyz = np.ascontiguousarray(filename_or_xyz, np.float32)[:, :3]
            self.box = np.asarray(box, np.float32)

    def initTopology(self, types=None):
        self.types = (np.zeros(self.numAtom, np.int32) if types is None
                      else np.ascontiguousarray(types, np.int32))

    def initDeviceData(self):
        import torch
        self._pos = torch.zeros((self.numAtom, 4), dtype=torch.float32, device="cuda")
        self._pos[:, :3] = torch.from_numpy(self.xyz).cuda()
        self.forcx = torch.zeros(self.numAtom, dtype=torch.float32, device="cuda")
        self.forcy = torch.zeros_like(self.forcx)
        self.forcz = torch.zeros_like(self.forcx)
        torch.cuda.synchronize()

    @property
    def coord(self):
        """Device positions (N, 4) float32, like sys.ddata.coord."""
        return self._pos

    def normalizeDeviceData(self):
        pass

    def writeHostDataGro(self, filename, step=0, time=0.0, timer=None):
        """Positions from the device (and velocities if set) as .gro
        (H/src/MDSystem_interface.cu:174-214)."""
        xyz = self._pos[:, :3].cpu().numpy()
        vel = getattr(self, "velo", None)
        if self.engine is not None:
            try:                       # the engine's MD velocities, if it has any
                vel = self.engine.velocities()
            except NotReady:
                pass
        _io.write_gro(filename, xyz, self.box, vel, step, time)

    def initWriteXtc(self, filename, prec=1000.0):
        """xdrfile_open(filename, "w") and the rectangular box matrix
        (H/src/MDSystem_interface.cu:243-263)."""
        self._xtc = _io.XtcFile(filename, "w")
        self._xtc_prec = float(prec)

    def writeHostDataXtc(self, step, time, timer=None):
        """write_xtc of the device positions (H/src/MDSystem_interface.cu:265-284)."""
        xyz = self._pos[:, :3].cpu().numpy()
        self._xtc.write(step, time, self.box, xyz, self._xtc_prec)

    def endWriteXtc(self):
        """xdrfile_close (H/src/MDSystem_interface.cu:286-290)."""
        self._xtc.close()
        self._xtc = None

    def _engine(self, n_types, skin):
        if self.engine is None:
            self.engine = Engine(n_types=n_types, skin=skin)
            self.engine.set_system([self.numAtom], self.box.reshape(1, 3))
            self.engine.set_types(self.types)
            self.engine.bind_positions(self._pos)
        return self.engine


class DensityProfile_PiecewiseConst:
    """Density on a grid of odd edges n = int(L/refh) (H/include/DensityProfile.h:14-59),
    accumulated on the device (deposite) and averaged by calculate()."""

    def __init__(self):
        self.arc = None
        self.profile = None

    def reinit(self, bx, by, bz, refh):
        self.arc = Arc((bx, by, bz), refh)
        self.nx, self.ny, self.nz = self.arc.dims
        self.profile = None

    def getNx(self):
        return self.nx

    def getNy(self):
        return self.ny

    def getNz(self):
        return self.nz

    def getBox(self):
        return list(self.arc.box)

    def clearData(self):
        self.arc.clear()

    def deposite(self, coord, numAtom=None):
        """coord: sys.coord (device (N, 4) float32) or an MDSystem."""
        if isinstance(coord, MDSystem):
            coord = coord.coord
        self.arc.deposit(coord if numAtom is None else coord[:numAtom])

    def calculate(self):
        self.profile = self.arc.density()

    def getProfile(self, i, j=None, k=None):
        if j is None:
            return self.profile.ravel()[i]
        return self.profile[i, j, k]

    def print_x(self, path):
        hx = self.arc.box[0] / self.nx
        with open(path, "w") as f:
            for i in range(self.nx):
                f.write("%f %f\n" % ((i + 0.5) * hx, self.profile[i, 0, 0]))

    def save(self, path):
        _io.save_density(path, self.arc.box, self.profile)

    def load(self, path):
        box, rho = _io.load_density(path)
        if tuple(rho.shape) != tuple(self.arc.dims):
            raise ValueError(f"{path}: grid {rho.shape} differs from {self.arc.dims}")
        self.arc.set_density(rho)
        self.profile = rho

    def init_write(self, path):
        self._traj = _io.GridTrajectory(path, self.arc.box, self.arc.dims, np.float64)

    def write(self, time):
        self._traj.write(time, self.profile)

    def end_write(self):
        self._traj.close()


class AdaptRCut:
    """Per-cell cutoff from the estimated force error (H/include/AdaptRCut.h:31-108),
    computed on the device over the grid of a DensityProfile_PiecewiseConst."""

    def __init__(self):
        self.dp = None
        self.rcut = None
        self.rcutIndex = None
        self.result_error = None

    def reinit(self, rcmin, rcmax, rcstep, dp):
        self.dp = dp
        self.rcList = dp.arc.set_candidates(rcmin, rcmax, rcstep)

    def getNx(self):
        return self.dp.nx

    def getNy(self):
        return self.dp.ny

    def getNz(self):
        return self.dp.nz

    def getRcList(self):
        return self.rcList

    def calError(self, dp=None):
        self.dp.arc.cal_error()

    def calRCut(self, prec):
        self.dp.arc.cal_rcut(prec)
        self._fetch()

    def refineRCut(self):
        self.dp.arc.refine()
        self._fetch()

    def _fetch(self):
        rc, err, idx = self.dp.arc.result()
        self.rcutIndex = idx
        self.rcut = self.rcList[idx]
        self.result_error = err

    def getRCut(self):
        return self.rcut

    def getProfileIndex(self):
        return self.rcutIndex

    def print_x(self, path):
        hx = self.dp.arc.box[0] / self.dp.nx
        with open(path, "w") as f:
            for i in range(self.dp.nx):
                f.write("%f %e %f\n" % ((i + 0.5) * hx, self.result_error[i, 0, 0],
                                        self.rcut[i, 0, 0]))

    def print_rc_avg(self, path):
        hx = self.dp.arc.box[0] / self.dp.nx
        avg = self.rcut.mean(axis=(1, 2))
        with open(path, "w") as f:
            for i in range(self.dp.nx):
                f.write("%f %e\n" % ((i + 0.5) * hx, avg[i]))

    def print_error_avg(self, dp, path):
        """sqrt(sum_yz err^2 rho / sum_yz rho) along x."""
        hx = self.dp.arc.box[0] / self.dp.nx
        rho = dp.profile
        num = (self.result_error ** 2 * rho).sum(axis=(1, 2))
        den = rho.sum(axis=(1, 2))
        with open(path, "w") as f:
            for i in range(self.dp.nx):
                f.write("%f %e\n" % ((i + 0.5) * hx, np.sqrt(num[i] / den[i])))

    def save_rc(self, path):
        """Text state of H/src/AdaptRCut.cu:698-726: candidates, box, grid, indices."""
        _io.write_rc_save(path, self.rcList, self.dp.arc.box, self.dp.arc.dims,
                         self.rcutIndex)

    def load_rc(self, path):
        """Reads save_rc's text back; the candidates must be equally spaced."""
        rcs, box, dims, idx = _io.read_rc_save(path)
        nrc = len(rcs)
        box = [float(v) for v in box]
        if self.dp is None or tuple(self.dp.arc.dims) != dims:
            raise ValueError(f"{path}: grid {dims} differs from the density grid")
        step = rcs[1] - rcs[0] if nrc > 1 else 1.0
        self.rcList = self.dp.arc.set_candidates(rcs[0], rcs[-1] + 0.5 * step, step)
        if len(self.rcList) != nrc or np.abs(self.rcList - rcs).max() > 1e-5:
            raise ValueError(f"{path}: candidates are not equally spaced")
        self.dp.arc.set_index(idx)
        self._fetch()
        self.box = box


class ForceCorr:
    """Mean error force of the pairs beyond a uniform rc on the density grid
    (FC/include/ForceCorr.h; reinit/calError FC/src/ForceCorr.cu:73-162), on
    the device: a one-candidate AdaptRCut grid of its own."""

    def __init__(self):
        self.arc = None
        self.f = None

    def reinit(self, rc, dp):
        self.dp = dp
        self.arc = Arc(dp.arc.box, 1.0)
        if self.arc.dims != dp.arc.dims:        # refh is not kept: rebuild from dims
            raise ValueError("ForceCorr needs the density grid of refh 1")
        self.arc.set_candidates(rc, rc, 1.0)

    def calError(self, dp=None):
        dp = dp or self.dp
        if dp.profile is None:
            dp.calculate()
        self.arc.set_density(dp.profile)
        self.arc.cal_error(keep_parts=True)
        self.arc.cal_rcut(1.0)                   # one candidate: index 0 everywhere
        self.f = self.arc.force_grid()

    def getForceCorrX(self):
        return self.f[..., 0]

    def getForceCorrY(self):
        return self.f[..., 1]

    def getForceCorrZ(self):
        return self.f[..., 2]

    def getNx(self):
        return self.arc.dims[0]

    def getNy(self):
        return self.arc.dims[1]

    def getNz(self):
        return self.arc.dims[2]


class AssignForceCorr:
    """Per-atom force correction from a ForceCorr grid (FC/include/AssignForceCorr.h;
    FC/src/AssignForceCorr.cu:15-145): applied by every force evaluation."""

    def __init__(self):
        self.grid = None

    def reinit(self, sys, fc, NThread=96):
        self.sys = sys
        self.dims = (fc.getNx(), fc.getNy(), fc.getNz())

    def getForceCorr(self, fc):
        self.grid = np.array(fc.f, np.float32)

    def zero(self):
        self.grid = np.zeros(self.dims + (3,), np.float32)

    def assign(self, sys):
        sys._fgrid = self.grid          # uploaded by the next force evaluation


class PressureCorrection:
    """Pressure-tensor correction of the pairs beyond each cell's rc
    (H/include/PressureCorrection.h:54-99), on the device over the AdaptRCut grid."""

    def __init__(self, arc=None, dp=None):
        self.pxx = self.pyy = self.pzz = 0.0
        if arc is not None:
            self.reinit(arc, dp)

    def reinit(self, arc, dp=None):
        self.arc = arc

    def correction(self, arc=None, dp=None):
        self.pxx, self.pyy, self.pzz = (float(v) for v in (arc or self.arc).dp.arc.pressure())


class AssignRCut:
    """rc lattice assigned per atom (H/include/AssignRCut.h:29-37)."""

    def __init__(self):
        self.dims = (1, 1, 1)
        self.table = None

    def reinit(self, sys, dims_or_arc, NThread=96):
        if isinstance(dims_or_arc, AdaptRCut):
            dims_or_arc = (dims_or_arc.getNx(), dims_or_arc.getNy(), dims_or_arc.getNz())
        self.dims = tuple(int(v) for v in dims_or_arc)
        self.sys = sys

    def getRCut(self, table):
        """table: an rc lattice or an AdaptRCut (its current result, as FP32)."""
        if isinstance(table, AdaptRCut):
            table = table.getRCut()
        self.table = np.asarray(table, np.float32).reshape(self.dims)
        self.sys._rc = self.table

    def uniform(self, rc):
        self.getRCut(np.full(self.dims, rc, np.float32))

    def init_write(self, path):
        self._traj = _io.GridTrajectory(path, self.sys.box, self.dims, np.float32)

    def write(self, time):
        self._traj.write(time, self.table)

    def end_write(self):
        self._traj.close()

    def assign(self, sys):
        sys._rc = self.table
        if sys.engine is not None:
            sys.engine.set_rc_table(self.table)


class CellList:
    """Kept for interface parity; binning runs inside NeighborList.rebuild."""

    def __init__(self, sys, cellSize, NTcell=96, NTatom=96, bdir=7, divide=1):
        self.cellSize = float(cellSize)

    def rebuild(self, sys, timer=None):
        pass


class NeighborList:
    """Verlet list with the per-atom radius rc_i + skin (H/include/NeighborList_interface.h)."""

    def __init__(self, sysNbInter, sys, rlist, rlistExten, NThread=96, expansion=4.0):
        self.nb = sysNbInter
        self.skin = float(rlistExten)
        self.sys = sys
        sys._skin = self.skin

    def rebuild(self, sys, clist=None, timer=None):
        e = sys._engine(max(1, self.nb.numAtomTypes()), self.skin)
        eps, sig, sh = self.nb.matrices()
        e.set_pair_params(eps, sig, sh)
        e.set_rc_table(sys._rc)
        e.rebuild()

    def calSumNeighbor(self):
        return self.sys.engine.neighbor_count(in_rc=False)[0]


class MDStatistic:
    """Energy and virial diagonal (H/include/Statistic_interface.h:30-117)."""

    def __init__(self, sys):
        self.sys = sys
        self.clearDevice()

    def clearDevice(self):
        self.ddata = np.zeros(4)
        self.hdata = np.zeros(4)

    def _add(self, stats):
        self.ddata += stats

    def updateHost(self):
        self.hdata = self.ddata.copy()

    def nonBondedEnergy(self):
        return float(self.hdata[0])

    def virial(self):
        return self.hdata[1:].copy()

    def pressureXX(self, box, kinetic_xx=0.0):
        V = float(np.prod(box))
        return 2.0 / V * (kinetic_xx - 0.5 * self.hdata[1])


class InteractionEngine:
    """Force/energy/virial of the registered LJ table (H/include/InteractionEngine_interface.h)."""

    def __init__(self, sys, NThread=96):
        self.sys = sys
        self.tail = False

    def registNonBondedInteraction(self, sysNbInter):
        self.nb = sysNbInter

    def clearInteraction(self, sys):
        import torch
        sys.forcx.zero_()
        sys.forcy.zero_()
        sys.forcz.zero_()
        # the engine adds on its own stream: the zeros must land first
        torch.cuda.current_stream().synchronize()

    def applyNonBondedInteraction(self, sys, nlist, st=None, excl=None, timer=None):
        e = sys.engine
        e.update_positions(check_skin=False)
        fgrid = getattr(sys, "_fgrid", None)
        if fgrid is not None and getattr(sys, "_fgrid_sent", None) is not fgrid:
            e.set_force_grid(fgrid)
            sys._fgrid_sent = fgrid
        e.compute(stats=st is not None, tail=self.tail, fcorr=fgrid is not None)
        e.add_forces_f32_soa(sys.forcx, sys.forcy, sys.forcz)
        e.synchronize()          # sys.forc* are read on torch's stream afterwards
        if st is not None:
            st._add(e.stats()[0])


class Displacement_max:
    """Largest displacement since recordCoord (H/include/Displacement_interface.h:15-43)."""

    def __init__(self, sys, NThread=96):
        self.sys = sys

    def recordCoord(self, sys):
        pass  # the engine records positions at every list rebuild

    def calMaxDisplacemant(self, sys, timer=None):
        """Returns True when max |dr| > skin/2 (the reference compares on the host)."""
        return sys.engine.update_positions(check_skin=True)
