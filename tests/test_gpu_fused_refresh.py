"""Graph steps in perturbation mode fold the refresh of each next step into
the force launch (a second coordinate buffer; the step counter in two
slots).  A graph of one step is not folded, so K single-step graphs are the
unfolded reference: the forces, statistics, bound positions and skin test of
K-step graphs must be bitwise the same, for every lane count (the refresh
split over 4 lanes per atom, or done by one lane), both precisions, one or
several replicas, anchored or free perturbation, and across rebuilds."""
import numpy as np
import pytest

from sre import Engine, configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab():
    return configs.slab(seed=41, n=12000, box=(18.0, 18.0, 64.0), nz=96, name="slab_12k")


@pytest.fixture(scope="module")
def reps():
    return configs.replicas(seed=7, count=6, n_lo=2000, n_hi=6000, nz=32)


def run(w, prec, lanes, blocks, anchored, unfused, unroll=2):
    """blocks: graph lengths between rebuilds; returns a record per block."""
    import torch
    e = Engine(n_types=w.n_types, skin=w.skin, precision=prec)
    e.set_tuning(force_lanes=lanes, build_lanes=lanes, force_unroll=unroll)
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 17, 3, anchored=anchored)
    out = []
    for k in blocks:
        if unfused:
            for _ in range(k):
                e.graph_run(1, stats=True, tail=True, perturb=True)
        else:
            e.graph_run(k, stats=True, tail=True, perturb=True)
        F = e.forces_fixed() if prec == "mixed" else e.forces()
        out.append((F, e.stats().copy(), e.atom_stats().copy(), pos.cpu().numpy().copy()))
        e.rebuild()
        out.append(e.last_rebuild_displacement())
    e.close()
    return out


# (precision, lanes per atom, unroll): the FP64 form's plain kernel at unroll
# 2 and its pipelined one at unroll 1
@pytest.mark.parametrize("prec,lanes,unroll", [("mixed", 4, 2), ("mixed", 2, 2),
                                               ("mixed", 8, 4), ("mixed", 1, 1),
                                               ("fp64", 8, 2), ("fp64", 4, 1)])
def test_folded_refresh_equals_unfolded(torch_cuda, slab, prec, lanes, unroll):
    blocks = [6, 7, 2]
    a = run(slab, prec, lanes, blocks, True, unfused=True, unroll=unroll)
    b = run(slab, prec, lanes, blocks, True, unfused=False, unroll=unroll)
    for k, (x, y) in enumerate(zip(a, b)):
        if isinstance(x, tuple):
            for u, v in zip(x, y):
                np.testing.assert_array_equal(u, v, err_msg=f"block {k}")
        else:
            assert x == y, k
    assert 0.0 < b[1] <= 0.5 * slab.skin


def test_folded_refresh_free_walk_and_replicas(torch_cuda, slab, reps):
    """A free (unanchored) perturbation reads the positions the previous
    launch wrote; several replicas take their geometry from the table."""
    for w, anchored in ((slab, False), (reps, True), (reps, False)):
        a = run(w, "mixed", 4, [5, 3], anchored, unfused=True)
        b = run(w, "mixed", 4, [5, 3], anchored, unfused=False)
        for x, y in zip(a, b):
            if isinstance(x, tuple):
                for u, v in zip(x, y):
                    np.testing.assert_array_equal(u, v, err_msg=w.name)
            else:
                assert x == y, w.name


def test_folded_refresh_skin_flag(torch_cuda, slab):
    """The skin test of folded steps: a free walk of many steps trips it, and
    the flag agrees with the unfolded steps' displacement."""
    from sre.engine import SkinExceeded
    import torch
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=False)
    pos = torch.from_numpy(slab.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 123, 0, anchored=False)
    e.graph_run(3, stats=False, tail=False, perturb=True)
    e.check_skin_flag()               # 3 steps of 0.01: within skin/2 = 0.15
    e.graph_run(400, stats=False, tail=False, perturb=True)
    with pytest.raises(SkinExceeded):
        e.check_skin_flag()
