"""The uniform-cutoff MD driver (sre.apps.lj_nhc) on the reference's
equilibrated 16000-atom slab (T* = 0.85, tests/golden/slab_t085.npz)."""
import io as _io
import os

import numpy as np
import pytest

from sre import apps, io

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_lj_nhc_driver_short_run(torch_cuda, tmp_path):
    fx = np.load(os.path.join(HERE, "golden", "slab_t085.npz"))
    conf = tmp_path / "conf.gro"
    io.write_gro(conf, fx["xyz"], fx["box"], fx["vel"], resname="lj", atomname="lj")
    frames = []
    out = _io.StringIO()
    rows = apps.lj_nhc(str(conf), 600, rcut=5.0, ref_T=0.85, out=out, conf_feq=200,
                       on_frame=lambda s, t, x, e: frames.append((s, x)),
                       confstart=str(tmp_path / "confstart.gro"),
                       confout=str(tmp_path / "confout.gro"),
                       xtc=str(tmp_path / "traj.xtc"))
    text = out.getvalue().splitlines()
    assert text[1].startswith("#*     1") and text[2].startswith("#* nstep")
    assert rows.shape == (6, 12)
    assert np.array_equal(rows[:, 0], [100, 200, 300, 400, 500, 600])
    for ln, r in zip(text[3:], rows):
        assert int(ln.split()[0]) == r[0]
    # the equilibrated slab stays at its temperature and density
    assert np.all(np.abs(rows[:, 4] - 0.85) < 0.03)
    assert np.all(rows[:, 2] / 16000 < -4.0)
    assert [s for s, _ in frames] == [0, 200, 400, 600]
    x0, box = fx["xyz"].astype(np.float64), fx["box"]
    d = frames[-1][1] - np.mod(x0, box)
    d -= np.round(d / box) * box
    assert 0.05 < np.sqrt((d ** 2).sum(1).mean()) < 3.0
    # confout holds the final frame at the .gro precision
    xo, bo, vo = io.read_gro(tmp_path / "confout.gro", velocities=True)
    assert np.array_equal(bo, box)
    assert np.abs(xo - frames[-1][1]).max() < 6e-4
    xs, _ = io.read_gro(tmp_path / "confstart.gro")
    assert np.abs(xs - np.mod(x0, box)).max() < 6e-4
    # the xtc trajectory holds the same frames at precision 1000
    traj = io.read_xtc(str(tmp_path / "traj.xtc"))
    assert [f[0] for f in traj] == [0, 200, 400, 600]
    assert np.allclose([f[1] for f in traj], [0.0, 1.0, 2.0, 3.0])
    for (s, x), f in zip(frames, traj):
        assert np.array_equal(np.diag(f[2]), np.asarray(box, np.float32))
        # half a lattice step, plus the FP32 rounding of x and of the decoding
        assert np.abs(f[3] - x).max() <= 0.5e-3 + 3e-7 * np.abs(x).max()


def test_lj_nhc_force_corrected_driver(torch_cuda, tmp_path):
    """The force-corrected driver adds no force before its first correction
    update: a run whose first update comes at step 200 matches, bit for bit up
    to step 200, one whose first update comes at its end (same density
    sampling, so the same stretches), and departs from it afterwards."""
    fx = np.load(os.path.join(HERE, "golden", "slab_t085.npz"))
    conf = tmp_path / "conf.gro"
    io.write_gro(conf, fx["xyz"], fx["box"], fx["vel"], resname="lj", atomname="lj")
    runs = {}
    for upd in (200, 400):
        frames = {}
        apps.lj_nhc(str(conf), 400, rcut=5.0, ref_T=0.85, out=_io.StringIO(), conf_feq=200,
                    on_frame=lambda s, t, x, e: frames.__setitem__(s, x), confstart=None,
                    confout=None, fcorr=True, density_feq=40, update_fc_feq=upd)
        runs[upd] = frames
    assert np.array_equal(runs[200][200], runs[400][200])
    d = np.abs(runs[200][400] - runs[400][400])
    assert d.max() > 1e-4


def test_lj_nhc_adapt_driver(torch_cuda, tmp_path):
    """The adaptive-cutoff driver: fifteen thermo columns, the pressure
    correction of the installed cutoffs from the first update on, a cutoff
    lattice made of candidates that is wider at the interfaces than in the
    bulk liquid, and its rc trajectory file."""
    fx = np.load(os.path.join(HERE, "golden", "slab_t085.npz"))
    conf = tmp_path / "conf.gro"
    io.write_gro(conf, fx["xyz"], fx["box"], fx["vel"], resname="lj", atomname="lj")
    rtj = tmp_path / "rcut.rtj"
    rows, rc, err = apps.lj_nhc_adapt(
        str(conf), 800, rcut=5.0, ref_T=0.85, target_prec=0.0045, update_rc_feq=400,
        n_refine=2, out=_io.StringIO(), conf_feq=400, confstart=None, confout=None,
        rc_trajectory=str(rtj))
    assert rows.shape == (8, 15)
    assert np.all(rows[:4, 11:14] == 0.0)
    assert np.all(rows[4:, 13] > 0.0)         # the correction raises the tension
    cands = np.arange(3.0, 10.01, 0.5)
    assert rc.shape == (151, 21, 21)
    assert np.all(np.isin(rc, cands.astype(np.float32)))
    prof = rc.mean((1, 2))
    assert prof[70:80].max() < prof[45:60].max()   # bulk liquid vs the left interface
    assert np.all(err[70:80] <= 0.0045 * 1.0001)
    box, dims, times, frames = io.read_grid_trajectory(rtj, np.float32)
    assert dims == (151, 21, 21) and np.allclose(times, [2.0, 4.0])
    # a frame holds the lattice installed at or before its step
    assert np.all(np.isin(frames[0], cands.astype(np.float32)))
    assert np.array_equal(frames[1], rc)
