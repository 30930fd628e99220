"""Converts the reference's configuration files into compact fixtures.

    python tests/golden/make_fixtures.py

Reads (as data) the .gro inputs the reference ships and writes their
positions and boxes as .npz files next to this script:
  lj27.npz        shortRange.inhomo/template/tools/cal.real.error/examples/lj.27.gro
                  (27 LJ atoms in a 6^3 box, the reference's real-error example)
  conf16000.npz   shortRange.inhomo/template/confs/conf.16000.gro
                  (16000-atom liquid-vapour slab, 150 x 21 x 21, T* = 0.85; the
                  system of PRE Table I)
  conf10000.npz   shortRange.inhomo/template/confs/N00010000-T0.7-100x020x020.gro
  arc_t085.npz    the adaptive-cutoff outputs of the reference's 16000-atom slab run
                  (T* = 0.85, target precision 0.0045; shortRange.inhomo/alltext/
                  pubilsh.pre/resubmit/fig/t0.85-n16000-adapt-e0.0045/): the
                  run-averaged density along x (a.density.x.out), the rc averaged
                  over y, z (a.rcut.x.out), and the estimated error and rc of the
                  y = z = 0 column at the last update (error.x.out: x, error, rc),
                  the run-averaged estimated error along x (a.error.x.out) and
                  the run-averaged real error, the RMS over each 2-wide x slab
                  of the force beyond each atom's cutoff (a.real.x.out);
                  and of its force-corrected run at uniform rc 7.5
                  (…/t0.85-n16000-fcorr-rc07.5-feq0200/): that run's averaged
                  density (a.density.x.out) and force correction of the
                  y = z = 0 column (afc.x.out: x, fcx, fcy, fcz)
  slab_t085.npz   the last configuration, with velocities, of the reference's
                  16000-atom slab run at T* = 0.85 and uniform rc 7.5
                  (…/fig/t0.85-n16000-rc07.5uni/confout.gro, time 15000), and the
                  run-averaged observables of that system
                  (…/fig/converge/t0.85.{liquid,gas,tension}.{uni,adapt,fcorr}.out:
                  rc, value, error, large-rc reference value, its error; the
                  tension tables carry the tail correction in column 4), and
                  the same tables at T* = 0.70, 1.10 and 1.20 (keys t<T>_...)
GROMACS .gro layout: title line, atom count, one fixed-column line per atom
(positions in columns 21-44, 8 characters each), box line.
"""
import os
import sys

import numpy as np

REF = "/root/reference/shortRange.inhomo/template"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "lj27.npz": "tools/cal.real.error/examples/lj.27.gro",
    "conf16000.npz": "confs/conf.16000.gro",
    "conf10000.npz": "confs/N00010000-T0.7-100x020x020.gro",
}


def read_gro(path):
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1].split()[0])
    xyz = np.empty((n, 3), np.float32)
    for i in range(n):
        ln = lines[2 + i]
        xyz[i] = [float(ln[20:28]), float(ln[28:36]), float(ln[36:44])]
    box = np.array([float(v) for v in lines[2 + n].split()[:3]], np.float32)
    return xyz, box


FIG = "/root/reference/shortRange.inhomo/alltext/pubilsh.pre/resubmit/fig/"
ARC_DIR = FIG + "t0.85-n16000-adapt-e0.0045"
FC_DIR = FIG + "t0.85-n16000-fcorr-rc07.5-feq0200"


def arc_fixture():
    tables = {}
    for key, d, name in (("a_density", ARC_DIR, "a.density.x.out"),
                         ("a_rcut", ARC_DIR, "a.rcut.x.out"),
                         ("error_x", ARC_DIR, "error.x.out"),
                         ("a_error", ARC_DIR, "a.error.x.out"),
                         ("a_real", ARC_DIR, "a.real.x.out"),
                         ("fc_density", FC_DIR, "a.density.x.out"),
                         ("afc", FC_DIR, "afc.x.out")):
        src = os.path.join(d, name)
        if not os.path.exists(src):
            print(f"missing {src}", file=sys.stderr)
            return
        tables[key] = np.loadtxt(src)
    np.savez_compressed(os.path.join(HERE, "arc_t085.npz"), **tables)
    print("arc_t085.npz", {k: v.shape for k, v in tables.items()})


def read_gro_v(path):
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[1].split()[0])
    xyz = np.empty((n, 3), np.float32)
    vel = np.empty((n, 3), np.float32)
    for i in range(n):
        ln = lines[2 + i]
        xyz[i] = [float(ln[20:28]), float(ln[28:36]), float(ln[36:44])]
        vel[i] = [float(ln[44:52]), float(ln[52:60]), float(ln[60:68])]
    box = np.array([float(v) for v in lines[2 + n].split()[:3]], np.float32)
    return xyz, vel, box


def slab_fixture():
    src = FIG + "t0.85-n16000-rc07.5uni/confout.gro"
    if not os.path.exists(src):
        print(f"missing {src}", file=sys.stderr)
        return
    xyz, vel, box = read_gro_v(src)
    tables = {}
    for obs in ("liquid", "gas", "tension"):
        for kind in ("uni", "adapt", "fcorr"):
            f = FIG + f"converge/t0.85.{obs}.{kind}.out"
            # commented lines are runs left out of the published figure
            tables[f"{obs}_{kind}"] = np.loadtxt(f, comments="#", ndmin=2)
            for temp in ("0.70", "1.10", "1.20"):
                f = FIG + f"converge/t{temp}.{obs}.{kind}.out"
                tables[f"t{temp}_{obs}_{kind}"] = np.loadtxt(f, comments="#", ndmin=2)
    np.savez_compressed(os.path.join(HERE, "slab_t085.npz"), xyz=xyz, vel=vel, box=box,
                        **tables)
    print("slab_t085.npz", xyz.shape, box, {k: v.shape for k, v in tables.items()})


RCSAVE = ("/root/reference/shortRange.inhomo/alltext/writeDown/fig/"
          "adaptColl-x0.60-u2.20-rc5.0-63/plots/rcut.save")


def formats_fixture():
    """formats.npz: what pins the file writers byte for byte without shipping
    the files: the parsed content of the reference's rcut.save (AdaptRCut::
    save_rc output of its cluster-collision run, a 128 x 64 x 128 grid) and
    the header of the T* = 0.85 slab's confout.gro (its atoms are in
    slab_t085.npz), each with the SHA-256 and size of the original bytes."""
    import hashlib
    out = {}
    if os.path.exists(RCSAVE):
        raw = open(RCSAVE, "rb").read()
        tok = raw.decode().split()
        nrc = int(tok[0])
        dims = [int(v) for v in tok[4 + nrc:7 + nrc]]
        out.update(rc_rcs=np.array([float(v) for v in tok[1:1 + nrc]]),
                   rc_box=np.array([float(v) for v in tok[1 + nrc:4 + nrc]]),
                   rc_dims=np.array(dims, np.int32),
                   rc_idx=np.array(tok[7 + nrc:], dtype=np.int64).astype(np.int8),
                   rc_sha256=hashlib.sha256(raw).hexdigest(), rc_bytes=len(raw))
    else:
        print(f"missing {RCSAVE}", file=sys.stderr)
    src = FIG + "t0.85-n16000-rc07.5uni/confout.gro"
    if os.path.exists(src):
        raw = open(src, "rb").read()
        lines = raw.decode().splitlines()
        first = lines[2]
        out.update(gro_title=lines[0], gro_resname=first[5:10].strip(),
                   gro_atomname=first[10:15].strip(),
                   gro_sha256=hashlib.sha256(raw).hexdigest(), gro_bytes=len(raw))
    np.savez_compressed(os.path.join(HERE, "formats.npz"), **out)
    print("formats.npz", sorted(out))


def main():
    arc_fixture()
    slab_fixture()
    formats_fixture()
    for out, rel in FILES.items():
        src = os.path.join(REF, rel)
        if not os.path.exists(src):
            print(f"missing {src}", file=sys.stderr)
            continue
        xyz, box = read_gro(src)
        np.savez_compressed(os.path.join(HERE, out), xyz=xyz, box=box, source=rel)
        print(out, xyz.shape, box)


if __name__ == "__main__":
    main()
