"""The allocation-free entry point (sre_forces_static) through the C ABI,
checked against the engine path (sre_rebuild + sre_compute + sre_get_forces
+ sre_get_stats) on the same inputs: fixed-point forces, hence the exported
doubles, are bit-identical, and so are the energy and virial (same kernels,
same list order, same reduction order).  Two calls in a row, the second with
a smaller system, since every call on a GPU shares one scratch array."""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine, configs
from sre.engine import InvalidArgument, forces_static

pytestmark = pytest.mark.gpu


def engine_path(w):
    e = Engine(n_types=w.n_types, skin=w.skin)
    w.setup(e, tail=False)
    e.rebuild()
    e.compute(stats=True)
    F, st = e.forces(), e.stats()[0]
    e.close()
    return F, st


def static_path(w, torch, with_types):
    x = torch.from_numpy(w.xyz.copy()).cuda()
    t = torch.from_numpy(w.types.astype(np.int32)).cuda() if with_types else None
    F, st = forces_static(x, w.boxes[0], w.rc_tables[0].ravel(), skin=w.skin, types=t,
                          eps=w.eps, sigma=w.sigma, n_types=w.n_types)
    torch.cuda.synchronize()
    return F.cpu().numpy(), st


def test_static_entry_point_equals_engine_path(torch_cuda):
    torch = torch_cuda
    big = configs.slab(seed=51, n=12000, box=(18.0, 18.0, 64.0), nz=96, binary=True,
                       name="binary_12k")
    small = configs.slab(seed=52, n=6000, box=(16.0, 16.0, 50.0), nz=48, name="slab_6k")
    for w, with_types in ((big, True), (small, False)):
        print(f"sre_forces_static: {w.name}, N={w.n}, types given: {with_types}", flush=True)
        Fs, ss = static_path(w, torch, with_types)
        Fe, se = engine_path(w)
        assert np.array_equal(Fs, Fe), w.name
        assert np.array_equal(ss, se), (w.name, ss, se)
        # and the oracle agrees to the mixed path's tolerance
        box = w.boxes[0]
        rc, _ = O.assign_rc(w.xyz, box, w.rc_tables[0])
        off, j = O.build_pairs(w.xyz, box, rc, w.skin)
        F64, E64, W64, _ = O.forces(w.xyz, box, w.types, w.eps, w.sigma, rc, off, j)
        rms = np.sqrt((F64 ** 2).sum(1).mean())
        assert np.sqrt(((Fs - F64) ** 2).sum(1)).max() / rms <= 1e-5
        assert abs(ss[0] - E64.sum()) / np.abs(E64).sum() <= 1e-6


def test_static_entry_point_rejects_sizes_beyond_its_limits(torch_cuda):
    torch = torch_cuda
    x = torch.zeros((16385, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(InvalidArgument, match="static limit"):
        forces_static(x, [40.0, 40.0, 40.0], [2.5])
