"""Binning at a rebuild: the counting sort by cell (sort_mode 2; the default,
sort_mode 0, takes it up to 15 360 cells) and the LSD radix sort of the cell
key (sort_mode 1) must give the same cell order -- stable by caller index within a cell -- and so the same list and
bitwise the same results.  The FP64 form sums each atom's pairs in list
order, so its forces change in the last bits if any atom's place differs;
the fixed-point forces, statistics, cutoffs and pair sets are compared too.
Cases: a slab, several replicas, a binary mixture, cells far larger than
the default (hundreds of atoms per cell: long rank loops), more cells than
one block scans (the separate scan and scatter launches), and positions
moved between rebuilds (atoms crossing cells)."""
import numpy as np
import pytest

from sre import Engine, configs

pytestmark = pytest.mark.gpu


def results(w, prec, sort_mode, tuning=None, moves=2):
    import torch
    e = Engine(n_types=w.n_types, skin=w.skin, precision=prec)
    e.set_tuning(sort_mode=sort_mode, **(tuning or {}))
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    out = []
    e.rebuild()
    e.set_perturbation(0.02, 5, 0, anchored=False)
    for m in range(moves + 1):
        if m:
            e.graph_run(4, stats=False, tail=False, perturb=True)
            e.rebuild()
        e.compute(stats=True, tail=True)
        F = e.forces_fixed() if prec == "mixed" else e.forces()
        off, j = e.export_pairs()
        out.append((F, e.stats().copy(), e.rcut().copy(), off.copy(), j.copy(),
                    pos.cpu().numpy().copy()))
    e.close()
    return out


def same(a, b, tag):
    for k, (x, y) in enumerate(zip(a, b)):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v, err_msg=f"{tag} rebuild {k}")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_counting_sort_matches_radix_sort_slab(torch_cuda, prec):
    w = configs.slab(seed=3, n=12000, box=(18.0, 18.0, 64.0), nz=96, name="slab_12k")
    same(results(w, prec, 2), results(w, prec, 1), w.name)


def test_counting_sort_matches_radix_sort_replicas_and_mixture(torch_cuda):
    for w in (configs.replicas(seed=9, count=5, n_lo=1500, n_hi=5000, nz=32),
              configs.slab(seed=2, n=10000, box=(18.0, 18.0, 56.0), nz=64, binary=True,
                           name="binary_slab_10k")):
        same(results(w, "fp64", 2), results(w, "fp64", 1), w.name)


def test_counting_sort_matches_radix_sort_large_cells(torch_cuda):
    """Cells of 2.5 x 2.5 x 2 (rc_min + skin): several hundred atoms per
    cell in the liquid."""
    w = configs.slab(seed=4, n=9000, box=(16.0, 16.0, 50.0), nz=64, name="slab_9k")
    t = dict(cell_xy=2.5, cell_z=2.0)
    same(results(w, "fp64", 2, t, moves=1), results(w, "fp64", 1, t, moves=1), w.name)


def test_counting_sort_matches_radix_sort_many_cells(torch_cuda):
    """Cells of 0.5 x 0.5 x 0.1 (rc_min + skin): ~33 000 cells, above the
    16 384 that one block scans, so the scan and the scatter are separate
    launches."""
    w = configs.slab(seed=5, n=12000, box=(18.0, 18.0, 64.0), nz=96, name="slab_12k")
    t = dict(cell_xy=0.5, cell_z=0.1)
    same(results(w, "fp64", 2, t), results(w, "fp64", 1, t), w.name)


def test_rebuild_figures_with_many_partials(torch_cuda):
    """1.2M atoms: more than 16 384 list-build partials and skin partials, so
    the rebuild's figures (total list length, longest list, largest
    displacement since the previous rebuild) come from the two-stage
    reduction; checked against the per-atom list lengths and the positions."""
    import ctypes as C
    import torch
    from sre.engine import _ptr, check, lib
    w = configs.slab(seed=8, n=1_200_000, box=(80.0, 80.0, 300.0), nz=512,
                     name="slab_1.2M")
    e = Engine(n_types=w.n_types, skin=w.skin)
    w.setup(e, tail=False)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    x0 = pos.cpu().numpy()[:, :3].astype(np.float64)
    e.set_perturbation(0.01, 3, 0, anchored=False)
    e.graph_run(1, stats=False, tail=False, perturb=True)   # one step: its
    e.rebuild()                                             # displacement is the max
    x1 = pos.cpu().numpy()[:, :3].astype(np.float64)
    box = w.boxes[0].astype(np.float64)
    d = x1 - x0
    d -= box * np.round(d / box)
    disp = float(np.sqrt((d * d).sum(1).max()))
    assert 0.0 < disp < 0.5 * w.skin
    assert abs(e.last_rebuild_displacement() - disp) <= 1e-4 * disp + 1e-5
    nl, _, mx = e.neighbor_count(in_rc=False)
    off = np.empty(w.xyz.shape[0] + 1, np.int64)
    check(lib().sre_export_pairs(e._h, _ptr(off, C.c_int64), None))
    cnt = np.diff(off)
    assert nl == int(off[-1]) and mx == int(cnt.max())
    e.close()
