"""The cell-window form of the mixed force evaluation (sre_tuning.force_mode
= 1: no neighbor list; one wavefront per 16-atom tile of a cell column, four
lanes per atom, the candidates of the tile's cell window staged in LDS)
evaluates the same pairs with the same per-pair arithmetic as the list
kernel, so its fixed-point forces are bitwise the list kernel's; energy and
virial sum the same FP32 terms in another order.  Checked at the rebuild
positions and after steps that move the atoms (within the skin), on a
slab, a binary mixture (per-type parameters) and several replicas."""
import numpy as np
import pytest

from sre import Engine, configs
from sre.engine import InvalidArgument

pytestmark = pytest.mark.gpu


def run(w, force_mode, steps=(0, 3, 4)):
    import torch
    e = Engine(n_types=w.n_types, skin=w.skin)
    e.set_tuning(force_mode=force_mode)
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 9, 0, anchored=False)
    out = []
    for k in steps:
        if k:
            e.graph_run(k, stats=True, tail=True, perturb=True)
        else:
            e.compute(stats=True, tail=True)
        out.append((e.forces_fixed(), e.stats().copy(), pos.cpu().numpy().copy()))
    e.close()
    return out


@pytest.mark.parametrize("name", ["slab", "binary", "replicas"])
def test_cell_window_forces_equal_list_forces(torch_cuda, name):
    w = {"slab": lambda: configs.slab(seed=12, n=12000, box=(18.0, 18.0, 64.0), nz=96,
                                      name="slab_12k"),
         "binary": lambda: configs.slab(seed=13, n=10000, box=(18.0, 18.0, 56.0), nz=64,
                                        binary=True, name="binary_slab_10k"),
         "replicas": lambda: configs.replicas(seed=14, count=5, n_lo=1500, n_hi=5000,
                                              nz=32)}[name]()
    a, b = run(w, 0), run(w, 1)
    for k, ((fa, sa, xa), (fb, sb, xb)) in enumerate(zip(a, b)):
        np.testing.assert_array_equal(xa, xb, err_msg=f"positions {k}")
        np.testing.assert_array_equal(fa, fb, err_msg=f"forces {k}")
        np.testing.assert_allclose(sb, sa, rtol=1e-5, atol=1e-5 * np.abs(sa).max(),
                                   err_msg=f"energy/virial {k}")


def test_cell_window_form_is_mixed_only(torch_cuda):
    import torch
    w = configs.slab(seed=12, n=6000, box=(16.0, 16.0, 40.0), nz=64, name="slab_6k")
    e = Engine(n_types=w.n_types, skin=w.skin, precision="fp64")
    e.set_tuning(force_mode=1)
    w.setup(e, tail=False)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    with pytest.raises(InvalidArgument):
        e.compute(stats=False, tail=False)
    e.close()
