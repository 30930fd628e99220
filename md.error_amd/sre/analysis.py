"""Observables of a liquid-vapour slab, as the reference's analysis tools
compute them (R = shortRange.inhomo/template/tools):

  BlockAverage   R/cal.density.profile/src/BlockAverage.cpp:4-54 — the mean
                 over the first nblock * (n // nblock) samples and the error
                 of that mean from the spread of the nblock block means,
                 sqrt(sum (b_k - mean)^2 / (nblock (nblock - 1))).
  slab_densities R/cal.density.profile/app/density.cpp:84-112 — per frame, the
                 atoms whose x (wrapped once into [0, Lx)) lies in
                 [l1, l2] (liquid) or in [0, g1] or [g2, Lx) (gas), divided by
                 the region volumes.
  ErrorProfile   R/cal.real.error/src/ErrorProfile.cpp:6-122 — the RMS of
                 a per-atom force error over the atoms of each x slab of
                 width ~refh (an odd number of slabs), as print_x_avg writes
                 it: sqrt(sum |dF|^2 / count) over the slab.
  tension        column 11 of the MD driver's thermo line,
                 (P_xx - (P_yy + P_zz) / 2) * Lx / 2
                 (R/simulation.uniform.cutoff/app/lj.nhc.cu:158-171), averaged
                 from a start time on (R/sh.scripts/tension.uniform.sh:5-22,
                 which quotes twice the error of the mean).
"""
import numpy as np


class BlockAverage:
    def __init__(self, data=None, nblock=16):
        if data is not None:
            self.process(data, nblock)

    def process(self, data, nblock=16):
        v = np.asarray(data, np.float64).ravel()
        per = v.size // nblock
        if per < 1:
            raise ValueError(f"{v.size} samples cannot fill {nblock} blocks")
        used = v[:per * nblock]
        self.n_used = used.size
        self.avg = used.mean()
        blocks = used.reshape(nblock, per).mean(1)
        self.error = float(np.sqrt(((blocks - self.avg) ** 2).sum() / (nblock * (nblock - 1))))
        return self


def slab_densities(x, box, l1, l2, g1, g2):
    """(gas, liquid) number densities of one frame; x holds the atoms'
    x coordinates (the slab normal is x)."""
    lx, ly, lz = (float(v) for v in box[:3])
    t = np.asarray(x, np.float64).copy()
    t[t >= lx] -= lx
    t[t < 0] += lx
    vol_g = ly * lz * (g1 + lx - g2)
    vol_l = ly * lz * (l2 - l1)
    gas = np.count_nonzero((t <= g1) | (t >= g2))
    liq = np.count_nonzero((t >= l1) & (t <= l2) & ~((t <= g1) | (t >= g2)))
    return gas / vol_g, liq / vol_l


def pressure_diag(stats, kinetic_diag, box):
    """P_aa = 2 / V (K_aa - W_aa / 2) from the engine's (E, W_xx, W_yy, W_zz)
    and the kinetic tensor diagonal (R/../simulation.hetero.cutoff/include/
    Statistic_interface.h:92-117)."""
    vol = float(box[0]) * float(box[1]) * float(box[2])
    w = np.asarray(stats, np.float64)[1:4]
    return 2.0 / vol * (np.asarray(kinetic_diag, np.float64) - 0.5 * w)


def tension(p, lx):
    return (p[0] - 0.5 * (p[1] + p[2])) * 0.5 * float(lx)


class ErrorProfile:
    """RMS force error per x slab (the reference's real-error profile)."""

    def __init__(self, box, refh):
        self.lx = float(box[0])
        nx = int(self.lx / refh)
        nx += 1 - nx % 2
        self.nx = nx
        self.hx = self.lx / nx
        self.sum = np.zeros(nx)
        self.count = np.zeros(nx, np.int64)

    def deposit(self, x, dF):
        t = np.asarray(x, np.float64).copy()
        t[t >= self.lx] -= self.lx
        t[t < 0] += self.lx
        ix = np.minimum((t / self.hx).astype(np.int64), self.nx - 1)
        np.add.at(self.sum, ix, (np.asarray(dF, np.float64) ** 2).sum(1))
        np.add.at(self.count, ix, 1)

    def x_avg(self):
        """(slab centres, RMS error) as print_x_avg."""
        xs = (np.arange(self.nx) + 0.5) * self.hx
        rms = np.sqrt(np.divide(self.sum, self.count, out=np.zeros(self.nx),
                                where=self.count > 0))
        return xs, rms

