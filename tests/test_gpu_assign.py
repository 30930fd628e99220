"""sre_assign_rcut: the reference's AssignRCut::assign timing (cutoffs
reassigned between rebuilds, kept across rebuilds), through the C ABI, against
the oracle on a 16000-atom slab."""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine, configs
from sre.engine import InvalidArgument

pytestmark = pytest.mark.gpu

TOL_F = 1e-5     # max |F - F_oracle| / RMS |F_oracle|, as the parity tests


@pytest.fixture(scope="module")
def slab16():
    return configs.slab(seed=11, n=16000, box=(20.0, 20.0, 75.0), nz=128, name="slab_16k")


def force_err(F, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    return np.sqrt(((F - ref) ** 2).sum(1)).max() / rms


def test_assign_between_rebuilds_matches_oracle(torch_cuda, slab16):
    w = slab16
    box = w.boxes[0]
    t1 = w.rc_tables[0]
    # cutoffs up by 0.5 (more than the skin) in even z bins, down in odd ones
    sign = np.where(np.arange(t1.shape[2]) % 2 == 0, 0.5, -0.5).astype(np.float32)
    t2 = np.ascontiguousarray(t1 + sign)
    e = Engine(n_types=w.n_types, skin=w.skin)
    w.setup(e, tail=False)
    e.rebuild()
    rc1, _ = O.assign_rc(w.xyz, box, t1)
    rc2, _ = O.assign_rc(w.xyz, box, t2)
    assert np.array_equal(e.rcut(), rc1)
    e.assign_rcut()
    assert np.array_equal(e.rcut(), rc1)
    # a new table: the rebuild keeps the assigned cutoffs (lists for rc1)
    e.set_rc_table(t2)
    e.rebuild()
    assert np.array_equal(e.rcut(), rc1)
    off1, j1 = O.build_pairs(w.xyz, box, rc1, w.skin)
    off, j = e.export_pairs()
    assert np.array_equal(off, off1) and np.array_equal(j, j1)
    # reassigned without a rebuild: rc2 on the rc1 lists, as the reference
    e.assign_rcut()
    assert np.array_equal(e.rcut(), rc2)
    e.compute(stats=True)
    F, E, _, _ = O.forces(w.xyz, box, w.types, w.eps, w.sigma, rc2, off1, j1)
    assert force_err(e.forces(), F) <= TOL_F
    # and some pairs within rc2 are indeed absent from those lists
    off2, j2 = O.build_pairs(w.xyz, box, rc2, w.skin)
    F2, _, _, _ = O.forces(w.xyz, box, w.types, w.eps, w.sigma, rc2, off2, j2)
    assert off2[-1] > off1[-1]
    assert np.abs(F - F2).max() > 1e-3
    # the next rebuild lists rc2 + skin
    e.rebuild()
    off, j = e.export_pairs()
    assert np.array_equal(off, off2) and np.array_equal(j, j2)
    e.compute(stats=True)
    assert force_err(e.forces(), F2) <= TOL_F
    with pytest.raises(InvalidArgument):
        e.adjoint(dLdE=1.0, n_rc=t2.size)
