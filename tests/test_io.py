"""File formats of the reference's driver (sre/io.py): byte layouts and round
trips (CPU)."""
import struct

import numpy as np

from sre import io


def test_gro_write_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    xyz = rng.uniform(0, 30, (12, 3)).astype(np.float32)
    vel = rng.normal(size=(12, 3)).astype(np.float32)
    box = np.array([30.0, 30.0, 120.0], np.float32)
    p = tmp_path / "c.gro"
    io.write_gro(p, xyz, box, vel, step=7, time=0.035)
    lines = p.read_text().splitlines()
    assert lines[0] == "# at time = 0.035000, step = 7"
    assert lines[1] == "12"
    for i, ln in enumerate(lines[2:14]):
        assert len(ln) == 68
        assert int(ln[0:5]) == i + 1 and int(ln[15:20]) == i + 1
        assert ln[5:10] == "   LJ" and ln[10:15] == "    A"
        np.testing.assert_allclose([float(ln[20:28]), float(ln[28:36]), float(ln[36:44])],
                                   xyz[i], atol=5e-4)
        np.testing.assert_allclose([float(ln[44:52]), float(ln[52:60]), float(ln[60:68])],
                                   vel[i], atol=5e-5)
    assert lines[14] == "30.000000 30.000000 120.000000"
    x2, b2, v2 = io.read_gro(p, velocities=True)
    np.testing.assert_allclose(x2, xyz, atol=5e-4)
    np.testing.assert_allclose(v2, vel, atol=5e-5)
    assert np.array_equal(b2, box)


def test_rtj_dtj_byte_layout(tmp_path):
    box, dims = (150.0, 21.0, 21.0), (5, 3, 2)
    rng = np.random.default_rng(1)
    frames32 = rng.uniform(3, 7, (2,) + dims).astype(np.float32)
    frames64 = rng.uniform(0, 1, (2,) + dims)
    for name, dt, frames in (("rcut.rtj", np.float32, frames32),
                             ("density.dtj", np.float64, frames64)):
        p = tmp_path / name
        with io.GridTrajectory(p, box, dims, dt) as t:
            t.write(10.0, frames[0])
            t.write(20.5, frames[1])
        raw = p.read_bytes()
        head = struct.pack("=3d3i", *box, *dims)
        code = "f" if dt == np.float32 else "d"
        body = b"".join(struct.pack("=f", tm) + struct.pack("=%d%s" % (frames[k].size, code),
                                                            *frames[k].ravel())
                        for k, tm in enumerate((10.0, 20.5)))
        assert raw == head + body
        b, d, times, fr = io.read_grid_trajectory(p, dt)
        assert tuple(b) == box and d == dims
        assert np.array_equal(times, [10.0, 20.5]) and np.array_equal(fr, frames)


def test_density_save_load(tmp_path):
    rho = np.random.default_rng(2).uniform(0, 0.8, (7, 3, 5))
    p = tmp_path / "density.save"
    io.save_density(p, (30.0, 30.0, 120.0), rho)
    first = p.read_text().split("\n")[0]
    assert first == "30.000000 30.000000 120.000000  7 3 5"
    box, r2 = io.load_density(p)
    assert np.array_equal(box, [30.0, 30.0, 120.0])
    np.testing.assert_allclose(r2, rho, rtol=1e-6)
