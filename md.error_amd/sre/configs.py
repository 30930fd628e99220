"""Seeded synthetic workloads of the short-range force path.

  plain_fcc      homogeneous LJ fluid, 4000 atoms, jittered FCC, rho 0.8442,
                 uniform rc 2.5, skin 0.3 (the uniform-cutoff "plain method")
  slab           64000-atom liquid-vapour slab in 30 x 30 x 120, rho 0.73/0.02,
                 rc(z) in 256 bins from 2.5 (bulk liquid) to 5.0 (interfaces)
  binary_slab    the slab with A/B species 80:20, eps 1.0/1.5/0.5, sigma 1.0/0.8/0.88;
                 perturbed_params gives its force-matching target's parameters
  replicas       256 slabs, N log-uniform in [2000, 32000], aspect 3-5, 64-bin rc(z)
                 from 2.5 to 4.0
  large_slab     2,048,000 binary atoms in 96 x 96 x 384, rc(z) 2.5 -> 6.0 in 1024 bins

Resolution of the slab geometry: a liquid of density rho_l filling |z - Lz/2| < h
with rho_v elsewhere must hold exactly N atoms, so h follows from mass balance,
h = (N/(Lx Ly) - rho_v Lz) / (2 (rho_l - rho_v))   (48.39 for the 64k slab).
Positions are uniform in each region and placed by random sequential addition
with a minimum pair distance dmin = 0.85 sigma (no relaxation step needed).
The rc(z) table is rc_min for |z - Lz/2| <= h - w, rises linearly to rc_max at
|z - Lz/2| = h + w (w = 2.5), and is rc_max in the vapour; values are taken at
bin centres.  The density profile fed to the tail correction is the histogram of
the generated positions (the reference's DensityProfile_PiecewiseConst::deposite,
H/src/DensityProfile.cu:69-90).
"""
from dataclasses import dataclass, field

import numpy as np

from .engine import gen_fcc, gen_slab

RHO_L, RHO_V = 0.73, 0.02
DMIN = 0.85
SKIN = 0.3


@dataclass
class Workload:
    name: str
    xyz: np.ndarray                 # (N, 4) float32, replicas concatenated
    types: np.ndarray               # (N,) int32
    n_atoms: list                   # per replica
    boxes: np.ndarray               # (R, 3) float32
    rc_tables: list                 # per replica, (nx, ny, nz) float32
    eps: np.ndarray                 # (nt, nt) float32
    sigma: np.ndarray               # (nt, nt) float32
    skin: float = SKIN
    rho: list = field(default_factory=list)   # per replica (nt, nz) float64 or None
    notes: str = ""

    @property
    def n(self):
        return int(sum(self.n_atoms))

    @property
    def n_types(self):
        return int(self.eps.shape[0])

    def replica_slice(self, r):
        off = int(sum(self.n_atoms[:r]))
        return slice(off, off + int(self.n_atoms[r]))

    def setup(self, engine, tail=True):
        """Loads this workload into an Engine (positions uploaded)."""
        engine.set_system(self.n_atoms, self.boxes)
        engine.set_types(self.types)
        engine.set_pair_params(self.eps, self.sigma)
        for r, t in enumerate(self.rc_tables):
            engine.set_rc_table(t, replica=r)
            if tail and self.rho and self.rho[r] is not None:
                engine.set_density_profile(self.rho[r], replica=r)
        engine.upload_positions(self.xyz)
        return engine


def slab_half_width(n, box, rho_l=RHO_L, rho_v=RHO_V):
    lx, ly, lz = box
    return (n / (lx * ly) - rho_v * lz) / (2.0 * (rho_l - rho_v))


def rc_profile(nz, lz, h, rc_min, rc_max, w=2.5):
    zc = (np.arange(nz) + 0.5) * lz / nz
    a = np.abs(zc - 0.5 * lz)
    t = np.clip((a - (h - w)) / (2.0 * w), 0.0, 1.0)
    return (rc_min + t * (rc_max - rc_min)).astype(np.float32).reshape(1, 1, nz)


def density_histogram(xyz, types, box, nz, nt):
    lz = float(box[2])
    z = np.mod(xyz[:, 2].astype(np.float64), lz)
    b = np.minimum((z / lz * nz).astype(np.int64), nz - 1)
    vol = float(box[0]) * float(box[1]) * lz / nz
    rho = np.zeros((nt, nz))
    for t in range(nt):
        rho[t] = np.bincount(b[types == t], minlength=nz) / vol
    return rho


def _binary_params():
    eps = np.array([[1.0, 1.5], [1.5, 0.5]], np.float32)
    sig = np.array([[1.0, 0.8], [0.8, 0.88]], np.float32)
    return eps, sig


def plain_fcc(seed=1, ncell=10, density=0.8442, jitter=0.05, rc=2.5):
    a = (4.0 / density) ** (1.0 / 3.0)
    L = np.float32(ncell * a)
    xyz = gen_fcc(ncell, float(L) / ncell, jitter, seed)
    n = xyz.shape[0]
    return Workload(
        name="plain_fcc_4k", xyz=xyz, types=np.zeros(n, np.int32), n_atoms=[n],
        boxes=np.array([[L, L, L]], np.float32),
        rc_tables=[np.full((1, 1, 1), rc, np.float32)],
        eps=np.ones((1, 1), np.float32), sigma=np.ones((1, 1), np.float32),
        rho=[None],
        notes=f"FCC {ncell}^3 cells, a={a:.5f}, jitter U(+-{jitter}), uniform rc {rc}")


def slab(seed=2, n=64000, box=(30.0, 30.0, 120.0), nz=256, rc_min=2.5, rc_max=5.0,
         binary=False, name=None):
    box = np.asarray(box, np.float32)
    h = slab_half_width(n, box)
    xyz = gen_slab(n, box, h, RHO_L, DMIN, seed)
    if binary:
        rng = np.random.default_rng(seed + 1000)
        types = (rng.random(n) >= 0.8).astype(np.int32)   # 80 % A (type 0), 20 % B
        eps, sig = _binary_params()
    else:
        types = np.zeros(n, np.int32)
        eps, sig = np.ones((1, 1), np.float32), np.ones((1, 1), np.float32)
    nt = eps.shape[0]
    return Workload(
        name=name or ("binary_slab_64k" if binary else "slab_64k"), xyz=xyz, types=types,
        n_atoms=[n], boxes=box.reshape(1, 3),
        rc_tables=[rc_profile(nz, float(box[2]), h, rc_min, rc_max)],
        eps=eps, sigma=sig, rho=[density_histogram(xyz, types, box, nz, nt)],
        notes=f"h={h:.3f} (mass balance), dmin={DMIN}, rc {rc_min}->{rc_max} in {nz} bins")


def binary_slab(seed=3, n=64000):
    return slab(seed=seed, n=n, binary=True)


def perturbed_params(w, rel=0.02, seed=11):
    """The force-matching target's parameters of the differentiation
    workload: every type-pair parameter (n_types (n_types + 1) / 2 epsilons
    and as many sigmas, the pair (a, b) shared by both matrix entries)
    multiplied by 1 + rel N(0, 1), seeded.  Returns symmetric float32
    (eps, sigma) matrices; the target forces are the FP64 forces of the
    workload's positions at these parameters."""
    rng = np.random.default_rng(seed)
    nt = w.n_types
    eps = np.asarray(w.eps, np.float64).copy()
    sig = np.asarray(w.sigma, np.float64).copy()
    for m in (eps, sig):
        for a in range(nt):
            for b in range(a, nt):
                m[a, b] = m[b, a] = m[a, b] * (1.0 + rel * rng.normal())
    return eps.astype(np.float32), sig.astype(np.float32)


def replicas(seed=4, count=256, n_lo=2000, n_hi=32000, nz=64, rc_min=2.5, rc_max=4.0,
             f_liquid=0.807):
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count))).astype(np.int64)
    aspect = rng.uniform(3.0, 5.0, count)
    rho_bar = f_liquid * RHO_L + (1 - f_liquid) * RHO_V
    xyz, types, boxes, tables, rhos = [], [], [], [], []
    for r in range(count):
        n = int(ns[r])
        L = (n / (aspect[r] * rho_bar)) ** (1.0 / 3.0)
        box = np.array([L, L, aspect[r] * L], np.float32)
        h = slab_half_width(n, box)
        x = gen_slab(n, box, h, RHO_L, DMIN, seed * 100003 + r)
        t = np.zeros(n, np.int32)
        xyz.append(x)
        types.append(t)
        boxes.append(box)
        tables.append(rc_profile(nz, float(box[2]), h, rc_min, rc_max))
        rhos.append(density_histogram(x, t, box, nz, 1))
    return Workload(
        name="replicas_256", xyz=np.concatenate(xyz), types=np.concatenate(types),
        n_atoms=[int(v) for v in ns], boxes=np.stack(boxes), rc_tables=tables,
        eps=np.ones((1, 1), np.float32), sigma=np.ones((1, 1), np.float32), rho=rhos,
        notes=f"{count} replicas, N log-uniform [{n_lo},{n_hi}], aspect U[3,5], "
              f"rc {rc_min}->{rc_max} in {nz} bins")


def large_slab(seed=5, n=2048000, box=(96.0, 96.0, 384.0), nz=1024):
    return slab(seed=seed, n=n, box=box, nz=nz, rc_min=2.5, rc_max=6.0, binary=True,
                name="large_slab_2M")


WORKLOADS = {
    "plain_fcc_4k": plain_fcc,
    "slab_64k": slab,
    "binary_slab_64k": binary_slab,
    "replicas_256": replicas,
    "large_slab_2M": large_slab,
}
