"""GPU tests of the device MD loop (sre_md_run and friends), through the C ABI.

- one NVE step reproduces the FP32 replay of the oracle bit for bit, given the
  forces the device used, and those forces match the FP64 oracle;
- one NHC + COM-removal step matches the FP64 oracle step (FP32 round-off);
- trajectories do not depend on how the run is cut into calls, on the rebuild
  policy (fixed interval, per-step check, adaptive stretches with rollback)
  or on graph vs eager execution (bit-exact, NVE);
- invariants over longer runs (NVE energy, momentum) and a 400-step NHC run
  that follows the oracle's extended energy and temperature.
"""
import numpy as np
import pytest

from oracle import md as M
from oracle import oracle as O
from sre import Engine, configs
from sre.engine import SkinExceeded

pytestmark = pytest.mark.gpu

RC = 2.5
SHIFT = -4.0 * (RC ** -12 - RC ** -6)
DT = 0.004


@pytest.fixture(scope="module")
def fcc():
    return configs.plain_fcc()


def maxwell(n, T, seed, mass=1.0):
    v = np.random.default_rng(seed).normal(0.0, np.sqrt(T / mass), (n, 3))
    return v.astype(np.float32)


def md_engine(torch, w, v, mass=1.0, skin=0.4, thermostat=None, ref_T=0.0, tau_T=0.0,
              com=True, precision="mixed", dt=DT):
    e = Engine(n_types=1, skin=skin, precision=precision)
    e.set_system(w.n_atoms, w.boxes)
    e.set_types(w.types)
    e.set_pair_params(w.eps, w.sigma, np.full((1, 1), SHIFT, np.float32))
    e.set_rc_table(w.rc_tables[0])
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.set_masses([mass])
    e.set_velocities(v)
    e.set_integrator(dt, thermostat, ref_T, tau_T, remove_com=com)
    return e, pos


def oracle_force_fn(box, n):
    b = np.asarray(box, np.float32)
    rc = np.full(n, RC, np.float32)
    one = np.ones((1, 1), np.float32)

    def fn(x):
        x32 = np.asarray(x, np.float32)
        off, j = O.build_pairs(x32, b, rc, 0.0)
        F, E, _, _ = O.forces(x32, b, np.zeros(n, np.int32), one, one, rc, off, j,
                              shift=np.full((1, 1), SHIFT, np.float32))
        return F, E.sum()
    return fn


def pos_of(pos):
    return pos.cpu().numpy()[:, :3].copy()


def test_nve_step_bit_exact(torch_cuda, fcc):
    n = fcc.n
    mass = 1.7
    v0 = maxwell(n, 1.0, 5, mass)
    e, pos = md_engine(torch_cuda, fcc, v0, mass=mass, com=False)
    e.md_run(0)
    x0 = pos_of(pos)
    F0 = e.forces()
    assert np.array_equal(e.velocities(), v0)
    e.md_run(1)
    x1, v1, F1 = pos_of(pos), e.velocities(), e.forces()
    xr, vr = M.replay_nve_step_f32(x0, v0, F0.astype(np.float32), F1.astype(np.float32),
                                   np.full(n, mass, np.float32), DT, fcc.boxes[0])
    assert np.array_equal(x1, xr)
    assert np.array_equal(v1, vr)
    # the forces the device used are the FP64 oracle's at the new positions
    Fo, _ = oracle_force_fn(fcc.boxes[0], n)(x1)
    rms = np.sqrt((Fo ** 2).sum(1).mean())
    assert np.sqrt(((F1 - Fo) ** 2).sum(1)).max() / rms < 1e-5
    st = e.md_state()
    K = 0.5 * float(np.float32(mass)) * (v1.astype(np.float64) ** 2).sum()
    assert abs(st["kinetic"] - K) < 1e-10 * K


@pytest.mark.parametrize("steps", [1, 3])
def test_nhc_com_steps_match_oracle(torch_cuda, fcc, steps):
    n = fcc.n
    v0 = maxwell(n, 1.0, 6) + np.float32(0.05)       # nonzero momentum to remove
    e, pos = md_engine(torch_cuda, fcc, v0, thermostat="nhc2", ref_T=1.5, tau_T=0.2, com=True)
    x0 = pos_of(pos)
    e.md_run(steps, rebuild_every=10)
    x1, v1 = pos_of(pos), e.velocities()
    nhc = M.NHC2(n, 1.5, 0.2)
    fn = oracle_force_fn(fcc.boxes[0], n)
    xo, vo, _, rec = M.md_run(x0, v0, np.ones(n), lambda x: fn(x)[0], DT, steps, nhc=nhc,
                              com=True, box=fcc.boxes[0])
    d = x1 - xo
    d -= fcc.boxes[0] * np.round(d / fcc.boxes[0])
    assert np.abs(d).max() < 2e-5
    assert np.abs(v1 - vo).max() < 1e-4 * np.abs(vo).max()
    st = e.md_state()
    assert abs(st["kinetic"] - rec[-1][0]) < 1e-5 * rec[-1][0]
    # FP32 velocities feed K, so the chain follows the FP64 oracle to ~1e-7
    for k in ("xi1", "xi2"):
        assert abs(st[k] - getattr(nhc, k)) < 1e-6 * max(abs(getattr(nhc, k)), 1.0)
    for k in ("vxi1", "vxi2"):
        assert abs(st[k] - getattr(nhc, k)) < 1e-5 * max(abs(getattr(nhc, k)), 1e-3)
    assert abs(st["nhc_energy"] - nhc.hamiltonian()) < 1e-6 * abs(nhc.hamiltonian())


def _trajectory(torch, w, v0, chunks, rebuild_every, com, precision="mixed", skin=0.4,
                engine_out=None):
    e, pos = md_engine(torch, w, v0, com=com, precision=precision, skin=skin)
    nr = 0
    for c in chunks:
        nr += e.md_run(c, rebuild_every=rebuild_every)
    if engine_out is not None:
        engine_out.append(e)
    return pos_of(pos), e.velocities(), nr


@pytest.mark.parametrize("com", [False, True])
def test_trajectory_independent_of_segmentation(torch_cuda, fcc, com):
    """Forces are exact sums of fixed-point pair terms, so the list layout
    (rebuild times) does not change them; NVE trajectories are bit-identical
    whether run as one call, several calls, other rebuild intervals or the
    eager per-step skin check."""
    v0 = maxwell(fcc.n, 1.0, 7)
    xa, va, na = _trajectory(torch_cuda, fcc, v0, [24], 6, com)
    xb, vb, nb = _trajectory(torch_cuda, fcc, v0, [5, 11, 8], 4, com)
    xc, vc, nc = _trajectory(torch_cuda, fcc, v0, [24], 0, com)
    xd, vd, nd = _trajectory(torch_cuda, fcc, v0, [9, 15], "adaptive", com)
    assert na >= 4 and nb >= 6
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(xa, xc) and np.array_equal(va, vc)
    assert np.array_equal(xa, xd) and np.array_equal(va, vd)


def test_adaptive_rollback_keeps_trajectory(torch_cuda, fcc):
    """With a thin skin the first adaptive stretches overrun skin/2 and are
    rolled back; the kept trajectory is still the eager rule's, bit for bit."""
    v0 = maxwell(fcc.n, 2.0, 14)
    eng = []
    xa, va, na = _trajectory(torch_cuda, fcc, v0, [40], "adaptive", True, skin=0.1,
                             engine_out=eng)
    xe, ve, ne = _trajectory(torch_cuda, fcc, v0, [40], 0, True, skin=0.1)
    st = eng[0].md_state()
    assert st["rollbacks"] >= 1
    assert 1 <= st["adapt_steps"] < 8
    assert ne >= 5
    assert np.array_equal(xa, xe) and np.array_equal(va, ve)


def test_fp64_form_trajectory_close(torch_cuda, fcc):
    v0 = maxwell(fcc.n, 1.0, 8)
    xm, vm, _ = _trajectory(torch_cuda, fcc, v0, [10], 5, True)
    xd, vd, _ = _trajectory(torch_cuda, fcc, v0, [10], 5, True, precision="fp64")
    d = xm - xd
    d -= fcc.boxes[0] * np.round(d / fcc.boxes[0])
    assert np.abs(d).max() < 1e-5
    assert np.abs(vm - vd).max() < 1e-4


def test_nve_energy_conserved(torch_cuda, fcc):
    n = fcc.n
    v0 = maxwell(n, 1.0, 9)
    e, _ = md_engine(torch_cuda, fcc, v0, com=True)
    E = []
    for _ in range(10):
        e.md_run(50, rebuild_every=5, stats=True)
        E.append(e.stats()[0, 0] + e.md_state()["kinetic"])
    E = np.array(E) / n
    # per-atom total energy: fluctuation and drift at dt = 0.004 (O(dt^2))
    assert np.abs(E - E[0]).max() < 2e-3
    # momentum after COM removal stays at FP32 round-off of the kicks
    v = e.velocities().astype(np.float64)
    assert np.abs(v.sum(0)).max() / n < 1e-6


def test_momentum_conserved_without_removal(torch_cuda, fcc):
    """Pair forces are exactly antisymmetric in fixed point: sum F = 0, so
    the momentum stays what it was up to FP32 round-off of the velocities."""
    n = fcc.n
    v0 = maxwell(n, 1.0, 10) + np.float32(0.1)
    P0 = v0.astype(np.float64).sum(0)
    e, _ = md_engine(torch_cuda, fcc, v0, com=False)
    e.md_run(100, rebuild_every=5)
    P1 = e.velocities().astype(np.float64).sum(0)
    assert np.abs(P1 - P0).max() / n < 1e-5
    assert np.abs(e.forces().sum(0)).max() < 1e-6 * n


def test_nhc_run_follows_oracle(torch_cuda, fcc):
    """400 NHC steps (heating from T ~ 0.65 towards 1.5): the extended energy
    U + K + H_NHC and the temperature follow the FP64 oracle's run.  (The
    extended energy itself moves by O(dt^2) per unit of heat pumped in,
    identically in both; the oracle pins that order separately.)"""
    n = fcc.n
    v0 = maxwell(n, 1.0, 11)
    e, pos = md_engine(torch_cuda, fcc, v0, thermostat="nhc2", ref_T=1.5, tau_T=0.2, com=True)
    x0 = pos_of(pos)
    H, T = [], []
    for _ in range(8):
        e.md_run(50, rebuild_every=5, stats=True)
        st = e.md_state()
        H.append(e.stats()[0, 0] + st["kinetic"] + st["nhc_energy"])
        T.append(2.0 * st["kinetic"] / (3.0 * n))
    fn = oracle_force_fn(fcc.boxes[0], n)
    nhc = M.NHC2(n, 1.5, 0.2)
    _, _, _, rec = M.md_run(x0, v0, np.ones(n), lambda x: fn(x)[0], DT, 400, nhc=nhc, com=True,
                            box=fcc.boxes[0], energy_fn=lambda x: fn(x)[1])
    Ho = np.array([k + u + h for k, u, h in rec])[49::50]
    To = np.array([2.0 * k / (3.0 * n) for k, _, _ in rec])[49::50]
    assert np.abs((np.array(H) - Ho) / n).max() < 1e-3
    assert np.abs(np.array(T) - To).max() < 2e-3
    assert T[-1] > T[0] + 0.2                  # the thermostat heats


def test_segment_skin_violation_reported(torch_cuda, fcc):
    v0 = maxwell(fcc.n, 4.0, 12)
    e, _ = md_engine(torch_cuda, fcc, v0, skin=0.05, com=True)
    with pytest.raises(SkinExceeded):
        e.md_run(64, rebuild_every=64)


def test_state_checkpoint_restart(torch_cuda, fcc):
    """Restoring positions, velocities and the chain continues the trajectory."""
    n = fcc.n
    v0 = maxwell(n, 1.0, 13)
    kw = dict(thermostat="nhc2", ref_T=1.2, tau_T=0.5, com=True)
    e, pos = md_engine(torch_cuda, fcc, v0, **kw)
    e.md_run(20, rebuild_every=5)
    x_mid, v_mid, s_mid = pos.clone(), e.velocities(), e.md_state()
    e.md_run(20, rebuild_every=5)
    x_end, v_end = pos_of(pos), e.velocities()
    e2, pos2 = md_engine(torch_cuda, fcc, v_mid, **kw)
    pos2.copy_(x_mid)
    torch_cuda.cuda.synchronize()
    e2.set_md_state(xi1=s_mid["xi1"], xi2=s_mid["xi2"], vxi1=s_mid["vxi1"],
                    vxi2=s_mid["vxi2"])
    e2.rebuild()
    e2.md_run(20, rebuild_every=5)
    assert np.array_equal(pos_of(pos2), x_end)
    assert np.array_equal(e2.velocities(), v_end)
