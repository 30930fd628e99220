"""On-disk formats of the reference's MD driver (H = shortRange.inhomo/template/
tools/simulation.hetero.cutoff), written and read byte for byte alike:

  .gro        GromacsFileManager::writeGroFile (H/src/GromacsFileManager.cpp:95-121)
              via MDSystem::writeHostDataGro (H/src/MDSystem_interface.cu:174-214):
              a title line
              "# at time = %f, step = %d", the atom count, one
              "%5d%5s%5s%5d%8.3f%8.3f%8.3f%8.4f%8.4f%8.4f" line per atom
              (positions then velocities), the box "%f %f %f".
  rcut.rtj    AssignRCut::init_write / write (H/src/AssignRCut.cu:145-177):
              box (3 float64), grid (3 int32), then per frame the time
              (float32) and the rc lattice (float32 x nx ny nz).
  density.dtj DensityProfile_PiecewiseConst::init_write / write
              (H/src/DensityProfile.cu:273-306): the same header, frames of
              time (float32) + density (float64 x nx ny nz).
  density.save DensityProfile_PiecewiseConst::save / load (:225-258): text
              "%f %f %f  %d %d %d\\n" then "%e " per cell.
  rcut.save   AdaptRCut::save_rc (H/src/AdaptRCut.cu:697-726): text
              "%d " (candidate count), "%f " per candidate, "\\n%f %f %f\\n"
              (box), "%d %d %d\\n" (grid), "%d " per cell (candidate index,
              cells in the order iz + nz (iy + ny ix)), "\\n".
Native byte order, as the reference's fwrite.
  .xtc        MDSystem::initWriteXtc / writeHostDataXtc / endWriteXtc
              (H/src/MDSystem_interface.cu:243-290) over libxdrfile's
              write_xtc: GROMACS compressed-coordinate frames (XDR header,
              coordinates rounded to 1/precision, bit-packed), written and
              read by the library (csrc/xtc.cpp) -- XtcFile below.
"""
import ctypes as C

import numpy as np


def write_rc_save(path, rcs, box, dims, idx):
    """AdaptRCut::save_rc: candidates, box, grid and per-cell candidate index."""
    idx = np.asarray(idx).ravel()
    dims = tuple(int(v) for v in dims)
    if idx.size != int(np.prod(dims)):
        raise ValueError("index count does not match the grid")
    parts = ["%d " % len(rcs), "".join("%f " % float(r) for r in rcs),
             "\n%f %f %f\n" % tuple(float(v) for v in box[:3]),
             "%d %d %d\n" % dims,
             " ".join(str(int(v)) for v in idx), " \n" if idx.size else "\n"]
    with open(path, "w") as f:
        f.write("".join(parts))


def read_rc_save(path):
    """Inverse of write_rc_save -> (rcs float64, box float64[3], dims, idx int32[nx, ny, nz])."""
    with open(path) as f:
        tok = f.read().split()
    nrc = int(tok[0])
    rcs = np.array([float(v) for v in tok[1:1 + nrc]])
    box = np.array([float(v) for v in tok[1 + nrc:4 + nrc]])
    dims = tuple(int(v) for v in tok[4 + nrc:7 + nrc])
    n = int(np.prod(dims))
    body = tok[7 + nrc:]
    if len(body) != n:
        raise ValueError(f"{path}: {len(body)} cell indices for a {dims} grid")
    idx = np.array(body, dtype=np.int64).astype(np.int32).reshape(dims)
    return rcs, box, dims, idx


def read_gro(path, velocities=False):
    """GROMACS .gro: positions (columns 21-44), optional velocities (45-68),
    the box (last line)."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1].split()[0])
    xyz = np.empty((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    for i in range(n):
        ln = lines[2 + i]
        xyz[i] = [float(ln[20:28]), float(ln[28:36]), float(ln[36:44])]
        if velocities and len(ln) >= 68:
            vel[i] = [float(ln[44:52]), float(ln[52:60]), float(ln[60:68])]
    box = np.array([float(v) for v in lines[2 + n].split()[:3]], np.float32)
    return (xyz, box, vel) if velocities else (xyz, box)


def write_gro(path, xyz, box, vel=None, step=0, time=0.0, resname="LJ", atomname="A"):
    """MDSystem::writeHostDataGro; residue and atom indices count from 1 and
    wrap at 100000 as the fixed columns do."""
    xyz = np.asarray(xyz, np.float64)
    n = len(xyz)
    vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, np.float64)
    out = ["# at time = %f, step = %d" % (time, step), "\n%d\n" % n]
    for i in range(n):
        k = (i + 1) % 100000
        out.append("%5d%5s%5s%5d%8.3f%8.3f%8.3f%8.4f%8.4f%8.4f\n" % (
            k, resname, atomname, k, xyz[i, 0], xyz[i, 1], xyz[i, 2],
            vel[i, 0], vel[i, 1], vel[i, 2]))
    out.append("%f %f %f\n" % tuple(float(v) for v in box[:3]))
    with open(path, "w") as f:
        f.write("".join(out))


class GridTrajectory:
    """Writer of rcut.rtj (frame dtype float32) or density.dtj (float64)."""

    def __init__(self, path, box, dims, dtype):
        self.dims = tuple(int(v) for v in dims)
        self.dtype = np.dtype(dtype)
        self.f = open(path, "wb")
        self.f.write(np.asarray(box, np.float64).reshape(3).tobytes())
        self.f.write(np.asarray(self.dims, np.int32).tobytes())

    def write(self, time, grid):
        g = np.ascontiguousarray(np.asarray(grid, self.dtype).reshape(self.dims))
        self.f.write(np.float32(time).tobytes())
        self.f.write(g.tobytes())

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_grid_trajectory(path, dtype):
    """(box, dims, times, frames) of an .rtj (float32) / .dtj (float64) file."""
    dt = np.dtype(dtype)
    raw = open(path, "rb").read()
    box = np.frombuffer(raw, np.float64, 3, 0)
    dims = tuple(int(v) for v in np.frombuffer(raw, np.int32, 3, 24))
    nele = int(np.prod(dims))
    frame = 4 + nele * dt.itemsize
    body = raw[36:]
    if len(body) % frame:
        raise ValueError(f"{path}: truncated frame")
    times, frames = [], []
    for o in range(0, len(body), frame):
        times.append(float(np.frombuffer(body, np.float32, 1, o)[0]))
        frames.append(np.frombuffer(body, dt, nele, o + 4).reshape(dims))
    return box.copy(), dims, np.array(times, np.float32), np.array(frames)


def save_density(path, box, rho):
    """DensityProfile_PiecewiseConst::save."""
    rho = np.asarray(rho, np.float64)
    with open(path, "w") as f:
        f.write("%f %f %f  " % tuple(float(v) for v in box[:3]))
        f.write("%d %d %d\n" % rho.shape)
        f.write("".join("%e " % v for v in rho.ravel()))


def load_density(path):
    """DensityProfile_PiecewiseConst::load -> (box, rho[nx, ny, nz])."""
    tok = open(path).read().split()
    box = np.array([float(v) for v in tok[:3]])
    dims = tuple(int(v) for v in tok[3:6])
    rho = np.array([float(v) for v in tok[6:6 + int(np.prod(dims))]]).reshape(dims)
    return box, rho


class XtcFile:
    """An xtc trajectory (mode "w", "a" or "r").

    write(step, time, box, xyz, precision=1000.0): box (3,) for a
    rectangular box or (3, 3); xyz (N, 3) or (N, >=3) (extra columns
    ignored).  read() returns (step, time, box (3, 3), xyz (N, 3),
    precision) or None after the last frame.  Iterating yields the frames."""

    def __init__(self, path, mode="w", natoms=None):
        from ._lib import lib
        from .engine import check
        self._lib, self._check = lib(), check
        h = C.c_void_p()
        check(self._lib.sre_xtc_open(str(path).encode(), mode.encode(), C.byref(h)))
        self._h = h
        self.mode = mode
        self.natoms = natoms
        if mode == "r" and natoms is None:
            n = C.c_int()
            check(self._lib.sre_xtc_natoms(str(path).encode(), C.byref(n)))
            self.natoms = n.value

    @staticmethod
    def _box(box):
        b = np.asarray(box, np.float32)
        if b.shape == (3,):
            b = np.diag(b)
        return np.ascontiguousarray(b.reshape(3, 3), np.float32)

    def write(self, step, time, box, xyz, precision=1000.0):
        x = np.ascontiguousarray(np.asarray(xyz, np.float32)[:, :3])
        b = self._box(box)
        self._check(self._lib.sre_xtc_write(
            self._h, len(x), int(step), float(time),
            b.ctypes.data_as(C.POINTER(C.c_float)),
            x.ctypes.data_as(C.POINTER(C.c_float)), float(precision)))

    def read(self):
        from ._lib import END_OF_FILE
        n = int(self.natoms)
        x = np.empty((n, 3), np.float32)
        b = np.empty((3, 3), np.float32)
        step, t, prec = C.c_int(), C.c_float(), C.c_float()
        st = self._lib.sre_xtc_read(self._h, n, C.byref(step), C.byref(t),
                                    b.ctypes.data_as(C.POINTER(C.c_float)),
                                    x.ctypes.data_as(C.POINTER(C.c_float)), C.byref(prec))
        if st == END_OF_FILE:
            return None
        self._check(st)
        return step.value, t.value, b, x, prec.value

    def __iter__(self):
        while True:
            f = self.read()
            if f is None:
                return
            yield f

    def close(self):
        if self._h:
            self._check(self._lib.sre_xtc_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_xtc(path):
    """All frames of an xtc file: list of (step, time, box, xyz, precision)."""
    with XtcFile(path, "r") as f:
        return list(f)
