"""The reference's MD driver programs over the engine (R =
shortRange.inhomo/template/tools, parameters as its run scripts inject them,
shortRange.inhomo/template/parameters.sh: dt 0.005, skin 0.49, tau_T 1,
refh 1, density sampling every 40 steps):

  lj_nhc(fcorr=False)  R/simulation.uniform.cutoff/app/lj.nhc.cu — one LJ
                       species (eps = sigma = 1, no energy shift), a uniform
                       cutoff, velocity Verlet + two-link Nose-Hoover chain +
                       centre-of-mass removal.
  lj_nhc(fcorr=True)   R/simulation.fcorr/app/lj.nhc-xtc.cu — the same plus
                       the long-range force correction: the density sampled
                       on the refh grid since the last update becomes, every
                       `update_fc_feq` steps, the mean error force of the
                       cutoff, which later force evaluations add at each
                       atom's grid cell (zero before the first update).
  lj_nhc_adapt         R/simulation.hetero.cutoff/app/lj.nhc-xtc.cu — the
                       adaptive cutoff: a uniform start, then every
                       `update_rc_feq` steps the sampled density gives every
                       candidate's error, each cell takes the smallest
                       candidate within `target_prec`, the choice is widened
                       `n_refine` times, installed, and the pressure
                       correction of the new cutoffs is printed with the
                       thermo line.

Each prints a thermo line every `thermo_feq` steps in its program's column
layout and writes confstart.gro / confout.gro.  The stretches between thermo
lines, samples and updates run on the device (`sre_md_run`, adaptive rebuild
stretches).  With xtc=path the programs' xtc trajectory is written (a frame
at step 0 and every `conf_feq` steps, precision 1000, as writeHostDataXtc);
a callback on_frame(step, time, xyz, engine) is called at the same steps
with the host positions and the engine (whose bound positions an Arc can
sample directly).  Differences that leave the ensemble unchanged: the centre-of-mass
velocity is removed every step (the uniform program: every 10th), the
correction force and the per-atom cutoff are looked up from the grids at
every force evaluation / list rebuild (the programs assign them every 20-40
steps), and the kinetic tensor of the thermo line is taken after the step's
second chain half-step.

Command line:  python -m sre.apps conf.gro nstep [rcut refT]
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


def thermo_line(step, dt, n, box, stats, kin_diag, nhc_energy, n_list, pc=None):
    """One line of the programs' output and its numbers: lj.nhc.cu:158-171
    (twelve columns, 7 digits) or, given the pressure correction pc, the
    adaptive program's fifteen (hetero.cutoff lj.nhc-xtc.cu:198-222: 5 digits,
    pc_xx, pc_yy and the correction of the tension before the count)."""
    p = analysis.pressure_diag(stats, kin_diag, box)
    k = float(np.sum(kin_diag))
    e = float(stats[0])
    row = (step, step * dt, e, k, k * 2.0 / 3.0 / (n - 3.0), e + k, e + k + nhc_energy,
           p[0], p[1], p[2], analysis.tension(p, box[0]))
    if pc is None:
        row = row + (float(n_list),)
        text = "%09d %07e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.2e\n" % row
    else:
        row = row + (float(pc[0]), float(pc[1]), analysis.tension(pc, box[0]),
                     float(n_list))
        text = ("%09d %05e %.5e %.5e %.5e %.5e %.5e %.5e %.5e %.5e %.5e %.5e %.5e %.5e"
                " %.2e\n") % row
    return text, row


class _Run:
    """Engine, bound positions and the loop shared by the drivers."""

    def __init__(self, conf, rcut, ref_T, tau_T, dt, skin, autotune):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("the MD drivers run on the GPU")
        xyz, box, vel = io.read_gro(conf, velocities=True)
        self.n = len(xyz)
        self.box = np.asarray(box, np.float32)
        self.dt = dt
        e = Engine(n_types=1, skin=skin, precision="mixed")
        e.set_system([self.n], [self.box])
        e.set_types(np.zeros(self.n, np.int32))
        one = np.ones((1, 1), np.float32)
        e.set_pair_params(one, one, np.zeros((1, 1), np.float32))
        e.set_rc_table(np.full((1, 1, 1), rcut, np.float32))
        pos4 = np.zeros((self.n, 4), np.float32)
        pos4[:, :3] = np.mod(xyz, self.box)            # normalizeDeviceData
        self.pos = torch.from_numpy(pos4).cuda()
        torch.cuda.synchronize()
        e.bind_positions(self.pos)
        e.set_masses([1.0])
        e.set_velocities(vel)
        e.set_integrator(dt, "nhc2", ref_T, tau_T, remove_com=True)
        e.rebuild()
        if autotune:
            e.autotune(stats=True)
        self.e = e

    def positions(self):
        return self.pos.cpu().numpy()[:, :3].astype(np.float64)

    def write_conf(self, path, step):
        io.write_gro(path, self.positions(), self.box, self.e.velocities(), step=step,
                     time=step * self.dt, resname="lj", atomname="lj")

    def frames(self, xtc, on_frame):
        """The per-frame callback: the xtc frame (if a path is given), then
        on_frame; returns (callback or None, file or None)."""
        if not xtc:
            return on_frame, None
        f = io.XtcFile(xtc, "w")

        def frame(step, time, xyz, e):
            f.write(step, time, self.box, xyz.astype(np.float32), 1000.0)
            if on_frame is not None:
                on_frame(step, time, xyz, e)
        return frame, f

    def loop(self, nstep, thermo_feq, conf_feq, out, on_frame, hooks, pc=lambda: None,
             fcorr=lambda: False):
        """hooks: (period, fn(step)) called after every period-th step, after
        that step's thermo line (the programs' order)."""
        e = self.e
        if on_frame is not None:
            on_frame(0, 0.0, self.positions(), e)
        periods = [thermo_feq] + [p for p, _ in hooks]
        if on_frame is not None:
            periods.append(conf_feq)
        rows = []
        done = 0
        while done < nstep:
            todo = min([nstep - done] + [p - done % p for p in periods])
            e.md_run(todo, rebuild_every="adaptive", stats=True, fcorr=fcorr())
            done += todo
            if done % thermo_feq == 0:
                v = e.velocities().astype(np.float64)
                kin = 0.5 * (v * v).sum(0)
                nl, _, _ = e.neighbor_count(in_rc=False)
                text, row = thermo_line(done, self.dt, self.n, self.box, e.stats()[0], kin,
                                        e.md_state()["nhc_energy"], nl, pc())
                out.write(text)
                out.flush()
                rows.append(row)
            for p, fn in hooks:
                if done % p == 0:
                    fn(done)
            if on_frame is not None and done % conf_feq == 0:
                on_frame(done, done * self.dt, self.positions(), e)
        return np.array(rows, np.float64)


def lj_nhc(conf, nstep, rcut=10.0, ref_T=0.70, tau_T=1.0, dt=0.005, skin=0.49,
           thermo_feq=100, conf_feq=2000, out=sys.stdout, on_frame=None,
           confstart="confstart.gro", confout="confout.gro", autotune=True,
           fcorr=False, refh=1.0, density_feq=40, update_fc_feq=20000, xtc=None):
    """Run the uniform-cutoff (or force-corrected) program; returns the thermo
    rows (float64, one row per line, columns as printed)."""
    r = _Run(conf, rcut, ref_T, tau_T, dt, skin, autotune)
    on_frame, xf = r.frames(xtc, on_frame)
    if confstart:
        r.write_conf(confstart, 0)
    out.write("# prepare ok, start to run\n")
    out.write(HEADER)
    hooks, fc, on = [], None, [False]
    if fcorr:
        fc = Arc(r.box, refh)
        fc.set_candidates(rcut, rcut, 1.0)

        def update(step):
            fc.cal_error(keep_parts=True)
            fc.cal_rcut(1.0)
            fc.apply_force_correction(r.e)
            on[0] = True
            if step != nstep:
                fc.clear()

        hooks = [(density_feq, lambda step: fc.deposit(r.e)), (update_fc_feq, update)]
    rows = r.loop(nstep, thermo_feq, conf_feq, out, on_frame, hooks,
                  fcorr=lambda: on[0])
    if xf is not None:
        xf.close()
    if confout:
        r.write_conf(confout, nstep)
    if fc is not None:
        fc.close()
    r.e.close()
    return rows.reshape(-1, 12)


def lj_nhc_adapt(conf, nstep, rcut=5.0, ref_T=1.10, tau_T=1.0, dt=0.005, skin=0.49,
                 rcmin=3.0, rcmax=10.0, rcstep=0.5, target_prec=0.020, refh=1.0,
                 density_feq=40, update_rc_feq=2000, n_refine=0, thermo_feq=100,
                 conf_feq=2000, out=sys.stdout, on_frame=None, confstart="confstart.gro",
                 confout="confout.gro", autotune=True, rc_trajectory=None,
                 on_update=None, assign_feq=None, xtc=None):
    """Run the adaptive-cutoff program (defaults as its source; the run
    scripts set rcutUpdateFeq 20000 and rcutNumRefine 2).  Returns the thermo
    rows (15 columns as printed) and the last installed rc lattice and its
    estimated error.  rc_trajectory: a path for the rc lattice trajectory
    (rcut.rtj layout), one frame per `conf_feq` steps; on_update(step, rc,
    err) is called after each cutoff update.  assign_feq: None assigns each
    atom's cutoff at every list rebuild (lists always cover rc_i + skin); a
    number reassigns every assign_feq steps without a rebuild, as the program
    does (rcutAssignFeq 40, H/app/lj.nhc-xtc.cu:184-188)."""
    r = _Run(conf, rcut, ref_T, tau_T, dt, skin, autotune)
    on_frame, xf = r.frames(xtc, on_frame)
    arc = Arc(r.box, refh)
    arc.set_candidates(rcmin, rcmax, rcstep)
    if confstart:
        r.write_conf(confstart, 0)
    out.write("# prepare ok, start to run\n")
    pc = [np.zeros(3)]
    last = [np.full(arc.dims, rcut, np.float32), None]
    rtj = None
    if rc_trajectory:
        rtj = io.GridTrajectory(rc_trajectory, r.box, arc.dims, np.float32)

    def update(step):
        arc.cal_error()
        arc.cal_rcut(target_prec)
        arc.refine(n_refine)
        arc.apply(r.e)
        pc[0] = arc.pressure()
        rc, err, _ = arc.result()
        last[0], last[1] = rc, err
        if on_update is not None:
            on_update(step, rc, err)
        if step != nstep:
            arc.clear()

    hooks = [(density_feq, lambda step: arc.deposit(r.e)), (update_rc_feq, update)]
    if assign_feq:
        r.e.assign_rcut()
        hooks.append((assign_feq, lambda step: r.e.assign_rcut()))
    if rtj is not None:
        hooks.append((conf_feq, lambda step: rtj.write(step * dt, last[0])))
    rows = r.loop(nstep, thermo_feq, conf_feq, out, on_frame, hooks, pc=lambda: pc[0])
    if xf is not None:
        xf.close()
    if rtj is not None:
        rtj.close()
    if confout:
        r.write_conf(confout, nstep)
    arc.close()
    r.e.close()
    return rows.reshape(-1, 15), last[0], last[1]


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
