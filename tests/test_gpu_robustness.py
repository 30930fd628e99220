"""GPU checks of the engine's error and recovery paths, through the C ABI:
neighbor-list overflow (the reference's mdErrorShortNeighborList,
H/src/NeighborList.cu:413-418) regrown by sre_rebuild and reported by
sre_run_segments, recovery after a force-range report, and the integrator
being discarded by sre_set_system.  The FP64 oracle is the checker."""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine
from sre.engine import ForceRange, NotReady, ShortNeighborList, SkinExceeded

pytestmark = pytest.mark.gpu

L = 20.0
RC = 2.5
SKIN = 0.3


def lattice_gas(seed=5, m=12, jitter=0.02):
    """m^3 atoms on a jittered simple cubic lattice filling an L^3 box."""
    rng = np.random.default_rng(seed)
    a = L / m
    g = (np.arange(m) + 0.5) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = x + rng.uniform(-jitter, jitter, x.shape)
    return np.concatenate([x, np.zeros((len(x), 1))], 1).astype(np.float32)


def compressed(x, radius=6.0, factor=0.6):
    """The atoms within `radius` of the box centre pulled towards it by
    `factor`: a cluster at density 0.216 / factor^3 = 1.0 (nearest neighbours
    a * factor = 1.0 sigma apart) inside an empty shell."""
    y = x.copy()
    c = np.full(3, L / 2, np.float32)
    d = y[:, :3] - c
    inside = (d ** 2).sum(1) < radius ** 2
    y[inside, :3] = c + factor * d[inside]
    return y


def initial_capacity(n, lanes):
    """The engine's first list-capacity guess: twice the shell count at the
    mean density plus 32, rounded up to the force kernel's lanes per atom."""
    rl = RC + SKIN
    g = int(4.18879 * rl ** 3 * (n / L ** 3) * 2.0 + 32.0)
    return -(-g // lanes) * lanes


def make_engine(x, precision="mixed"):
    e = Engine(n_types=1, skin=SKIN, precision=precision)
    e.set_system([len(x)], np.array([[L, L, L]], np.float32))
    e.set_types(np.zeros(len(x), np.int32))
    e.set_pair_params(np.ones((1, 1)), np.ones((1, 1)))
    e.set_rc_table(np.full((1, 1, 1), RC, np.float32))
    return e


def oracle(x):
    box = np.array([L, L, L])
    rc, _ = O.assign_rc(x, box, np.full((1, 1, 1), RC, np.float32))
    off, j = O.build_pairs(x, box, rc, SKIN)
    F, E, W, _ = O.forces(x, box, np.zeros(len(x), np.int32), np.ones((1, 1)),
                          np.ones((1, 1)), rc, off, j)
    return off, j, F, E


def check_against_oracle(e, x, tol):
    off, j, F, E = oracle(x)
    goff, gj = e.export_pairs()
    assert np.array_equal(goff, off) and np.array_equal(gj, j)
    rms = np.sqrt((F ** 2).sum(1).mean())
    err = np.sqrt(((e.forces() - F) ** 2).sum(1)).max() / rms
    assert err <= tol, err
    st = e.stats()[0]
    assert abs(st[0] - E.sum()) / np.abs(E).sum() <= 1e-6


def test_configuration_overflows_first_capacity():
    """CPU-side check of the set-up: the compressed cluster's longest list
    exceeds the capacity the engine sizes from the gas density."""
    x = lattice_gas()
    off, _, _, _ = oracle(compressed(x))
    off0, _, _, _ = oracle(x)
    cap = initial_capacity(len(x), 4)
    assert np.diff(off0).max() <= cap < np.diff(off).max()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_rebuild_regrows_list_after_overflow(torch_cuda, prec):
    """Lists built for a gas; the next rebuild meets a cluster whose lists
    overflow that capacity: sre_rebuild regrows and builds again, and the
    lists and forces equal the oracle's."""
    x = lattice_gas()
    e = make_engine(x, prec)
    e.upload_positions(x)
    e.rebuild()
    _, _, mx0 = e.neighbor_count(in_rc=False)
    cap = initial_capacity(len(x), e.tuning()["force_lanes"])
    assert mx0 <= cap
    y = compressed(x)
    e.upload_positions(y)
    e.rebuild()
    _, _, mx = e.neighbor_count(in_rc=False)
    assert mx > cap, "the configuration did not overflow the first capacity"
    e.compute(stats=True)
    check_against_oracle(e, y, 1e-5 if prec == "mixed" else 1e-11)


def test_run_segments_reports_overflow_and_recovers(torch_cuda):
    """Queued segments whose rebuild overflows the capacity: the lists of the
    later segments were cut, so sre_run_segments regrows, rebuilds and
    reports SRE_ERR_LIST_OVERFLOW; the next computation is exact."""
    import torch
    x = lattice_gas()
    y = compressed(x)
    e = make_engine(x)
    pos = torch.from_numpy(x.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    cap = initial_capacity(len(x), e.tuning()["force_lanes"])
    # the positions jump to the clustered configuration behind the engine's
    # back: segment 0 computes with stale lists, its rebuild overflows
    pos.copy_(torch.from_numpy(y))
    torch.cuda.synchronize()
    with pytest.raises(ShortNeighborList) as ei:
        e.run_segments(2, 1, stats=True)
    msg = str(ei.value)
    assert "segment 0" in msg and f"> {cap})" in msg and "regrown" in msg
    _, _, mx = e.neighbor_count(in_rc=False)
    assert mx > cap
    e.compute(stats=True)
    check_against_oracle(e, y, 1e-5)
    # and the queued path runs cleanly from there
    nls, disp = e.run_segments(2, 1, stats=True)
    assert disp == 0.0 and list(nls) == [nls[0]] * 2
    e.compute(stats=True)
    check_against_oracle(e, y, 1e-5)


def test_run_segments_reports_skin_violation(torch_cuda):
    """A segment in which atoms moved more than skin/2 (no overflow) is
    reported as SRE_ERR_SKIN; the lists are current afterwards."""
    import torch
    x = lattice_gas()
    e = make_engine(x)
    pos = torch.from_numpy(x.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    shifted = x.copy()
    shifted[:, 0] = (shifted[:, 0] + 0.4) % L      # 0.4 > skin / 2, uniform
    pos.copy_(torch.from_numpy(shifted))
    torch.cuda.synchronize()
    with pytest.raises(SkinExceeded, match="segment 0"):
        e.run_segments(2, 1, stats=True)
    e.compute(stats=True)
    check_against_oracle(e, shifted, 1e-5)


def test_force_range_flag_is_reported_once_and_cleared(torch_cuda):
    """After a force-range report the forces are invalid (NotReady), and once
    good positions are uploaded and the lists rebuilt the engine computes
    and reads normally again."""
    x = lattice_gas()
    bad = x.copy()
    bad[1, :3] = bad[0, :3] + np.array([0.3, 0.0, 0.0], np.float32)
    e = make_engine(x)
    e.upload_positions(bad)
    e.rebuild()
    e.compute(stats=True)
    with pytest.raises(ForceRange):
        e.stats()
    with pytest.raises(NotReady):
        e.forces()
    e.upload_positions(x)
    e.rebuild()
    e.compute(stats=True)
    check_against_oracle(e, x, 1e-5)


def test_set_system_discards_integrator(torch_cuda):
    """A new system needs a new integrator set-up (its degrees of freedom
    and chain masses belong to the old one)."""
    x = lattice_gas()
    e = make_engine(x)
    e.upload_positions(x)
    e.set_velocities(np.zeros((len(x), 3), np.float32))
    e.set_integrator(0.005)
    e.md_run(2, rebuild_every=1)
    n2 = len(x) - 64
    e.set_system([n2], np.array([[L, L, L]], np.float32))
    e.set_types(np.zeros(n2, np.int32))
    e.upload_positions(x[:n2])
    with pytest.raises(NotReady, match="integrator"):
        e.md_run(1, rebuild_every=1)
