"""File formats of the reference's driver (sre/io.py): byte layouts and round
trips (CPU)."""
import hashlib
import os
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


GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def test_reference_rcut_save_byte_for_byte(tmp_path):
    """AdaptRCut::save_rc's text (H/src/AdaptRCut.cu:697-726): the reference's
    own rcut.save (cluster collision, 11 candidates, 128 x 64 x 128 grid)
    re-emitted from its parsed content equals the original bytes, and reads
    back to the same content."""
    d = np.load(f"{GOLD}/formats.npz")
    dims = tuple(int(v) for v in d["rc_dims"])
    p = tmp_path / "rcut.save"
    io.write_rc_save(p, d["rc_rcs"], d["rc_box"], dims, d["rc_idx"].reshape(dims))
    assert p.stat().st_size == int(d["rc_bytes"])
    assert _sha(p) == str(d["rc_sha256"])
    rcs, box, dims2, idx = io.read_rc_save(p)
    assert dims2 == dims == (128, 64, 128)
    assert np.array_equal(rcs, d["rc_rcs"]) and np.array_equal(box, d["rc_box"])
    assert np.array_equal(idx.ravel(), d["rc_idx"])


def test_reference_confout_gro_byte_for_byte(tmp_path):
    """MDSystem::writeHostDataGro: the reference's confout.gro of the T* = 0.85
    slab (16000 atoms with velocities) re-emitted from its parsed atoms equals
    the original bytes."""
    d = np.load(f"{GOLD}/formats.npz")
    s = np.load(f"{GOLD}/slab_t085.npz")
    title = str(d["gro_title"])
    assert title.startswith("# at time = ")
    time = float(title.split("=")[1].split(",")[0])
    step = int(title.split("=")[2])
    p = tmp_path / "confout.gro"
    io.write_gro(p, s["xyz"], s["box"], s["vel"], step=step, time=time,
                 resname=str(d["gro_resname"]), atomname=str(d["gro_atomname"]))
    assert p.stat().st_size == int(d["gro_bytes"])
    assert _sha(p) == str(d["gro_sha256"])
    x2, b2, v2 = io.read_gro(p, velocities=True)
    assert np.array_equal(x2, s["xyz"]) and np.array_equal(v2, s["vel"])
