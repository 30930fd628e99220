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


def main():
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
