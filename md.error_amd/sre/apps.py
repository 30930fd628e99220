"""The reference's uniform-cutoff MD drivers over the engine.

`lj_nhc` follows R/simulation.uniform.cutoff/app/lj.nhc.cu and, with
`fcorr=True`, R/simulation.fcorr/app/lj.nhc-xtc.cu (R =
shortRange.inhomo/template/tools), with the parameters their run scripts
inject (shortRange.inhomo/template/parameters.sh: dt 0.005, skin 0.49,
tau_T 1, refh 1, density sampling every 40 steps, force-correction update
every 20000): one LJ species (eps = sigma = 1, no energy shift) with a
uniform cutoff, velocity Verlet + two-link Nose-Hoover chain +
centre-of-mass removal, a thermo line every `thermo_feq` steps in the
drivers' column layout, and confstart.gro / confout.gro.  The force-corrected
driver samples the density on the refh grid, and every `update_fc_feq` steps
turns the profile sampled since the last update into the mean error force of
the cutoff (the long-range force correction), which every later force
evaluation adds at each atom's grid cell; before the first update the
correction is zero (R/simulation.fcorr/app/lj.nhc-xtc.cu:147-222).

The stretches between thermo lines, samples and updates run on the device
(`sre_md_run`, adaptive rebuild stretches); the drivers' trajectory file
(xtc) is replaced by a callback, on_frame(step, time, xyz, engine), called
every `conf_feq` steps with the host positions and the engine (whose bound
positions an Arc can sample directly).  Differences that leave the ensemble unchanged: the centre-of-mass
velocity is removed every step (the uniform driver: every 10th), the
correction force is looked up at every force evaluation (the driver assigns
it every 20 steps), and the kinetic tensor of the thermo line is taken after
the step's second chain half-step.

Command line, as the driver's:  python -m sre.apps conf.gro nstep [rcut refT]
"""
import sys

import numpy as np

from . import analysis, io
from .arc import Arc
from .engine import Engine

HEADER = (
    "#*     1     2           3         4            5       6                7"
    "          8          9         10        11\n"
    "#* nstep  time  nonBondedE  kineticE  temperature  totalE  NHC_Hamiltonian"
    " pressureXX pressureYY pressureZZ s_tension\n")


def thermo_line(step, dt, n, box, stats, kin_diag, nhc_energy, n_list):
    """One line of the driver's output (lj.nhc.cu:158-171) and its numbers."""
    p = analysis.pressure_diag(stats, kin_diag, box)
    k = float(np.sum(kin_diag))
    e = float(stats[0])
    row = (step, step * dt, e, k, k * 2.0 / 3.0 / (n - 3.0), e + k, e + k + nhc_energy,
           p[0], p[1], p[2], analysis.tension(p, box[0]), float(n_list))
    text = "%09d %07e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.2e\n" % row
    return text, row


def lj_nhc(conf, nstep, rcut=10.0, ref_T=0.70, tau_T=1.0, dt=0.005, skin=0.49,
           thermo_feq=100, conf_feq=2000, out=sys.stdout, on_frame=None,
           confstart="confstart.gro", confout="confout.gro", autotune=True,
           fcorr=False, refh=1.0, density_feq=40, update_fc_feq=20000):
    """Run the driver; returns the thermo rows (float64 array, one row per
    thermo line, columns as printed, the neighbour count last)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("lj_nhc runs on the GPU")
    xyz, box, vel = io.read_gro(conf, velocities=True)
    n = len(xyz)
    box = np.asarray(box, np.float32)
    e = Engine(n_types=1, skin=skin, precision="mixed")
    e.set_system([n], [box])
    e.set_types(np.zeros(n, np.int32))
    one = np.ones((1, 1), np.float32)
    e.set_pair_params(one, one, np.zeros((1, 1), np.float32))
    e.set_rc_table(np.full((1, 1, 1), rcut, np.float32))
    pos4 = np.zeros((n, 4), np.float32)
    pos4[:, :3] = np.mod(xyz, box)            # normalizeDeviceData
    pos = torch.from_numpy(pos4).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.set_masses([1.0])
    e.set_velocities(vel)
    e.set_integrator(dt, "nhc2", ref_T, tau_T, remove_com=True)
    e.rebuild()
    if autotune:
        e.autotune(stats=True)

    def positions():
        return pos.cpu().numpy()[:, :3].astype(np.float64)

    if confstart:
        io.write_gro(confstart, positions(), box, e.velocities(), step=0, time=0.0,
                     resname="lj", atomname="lj")
    out.write("# prepare ok, start to run\n")
    out.write(HEADER)
    if on_frame is not None:
        on_frame(0, 0.0, positions(), e)
    fc = None
    if fcorr:
        fc = Arc(box, refh)
        fc.set_candidates(rcut, rcut, 1.0)
    fc_on = False
    rows = []
    done = 0
    while done < nstep:
        todo = min(thermo_feq - done % thermo_feq, nstep - done)
        if on_frame is not None:
            todo = min(todo, conf_feq - done % conf_feq)
        if fc is not None:
            todo = min(todo, density_feq - done % density_feq,
                       update_fc_feq - done % update_fc_feq)
        e.md_run(todo, rebuild_every="adaptive", stats=True, fcorr=fc_on)
        done += todo
        if fc is not None and done % density_feq == 0:
            fc.deposit(e)
        if fc is not None and done % update_fc_feq == 0:
            fc.cal_error(keep_parts=True)
            fc.cal_rcut(1.0)
            fc.apply_force_correction(e)
            fc_on = True
            if done != nstep:
                fc.clear()
        if done % thermo_feq == 0:
            v = e.velocities().astype(np.float64)
            kin = 0.5 * (v * v).sum(0)
            s = e.md_state()
            nl, _, _ = e.neighbor_count(in_rc=False)
            text, row = thermo_line(done, dt, n, box, e.stats()[0], kin, s["nhc_energy"], nl)
            out.write(text)
            out.flush()
            rows.append(row)
        if on_frame is not None and done % conf_feq == 0:
            on_frame(done, done * dt, positions(), e)
    if confout:
        io.write_gro(confout, positions(), box, e.velocities(), step=nstep,
                     time=nstep * dt, resname="lj", atomname="lj")
    if fc is not None:
        fc.close()
    e.close()
    return np.array(rows, np.float64).reshape(-1, 12)


def main(argv):
    if len(argv) not in (3, 5):
        print("Usage:\n%s conf.gro nstep [rcut refT]" % argv[0])
        return 1
    kw = {}
    if len(argv) == 5:
        kw = dict(rcut=float(argv[3]), ref_T=float(argv[4]))
    lj_nhc(argv[1], int(argv[2]), **kw)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
