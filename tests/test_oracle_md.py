"""Pins of the MD-loop oracle (oracle/md.py) by closed forms and invariants.

A plausible mistake in the restatement (a dropped or doubled half kick, a
chain term with the wrong weight, an unsymmetric operator order, a COM
removal with the wrong sign) fails at least one of these.
"""
import numpy as np
import pytest

from oracle import md as M


def _lj_fcc(ncell=3, rho=0.8442, seed=0):
    a = (4.0 / rho) ** (1.0 / 3.0)
    L = ncell * a
    base = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    g = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)])
    x = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    x += np.random.default_rng(seed).uniform(-0.05, 0.05, x.shape)
    return x % L, L


def _lj(L, rc=2.5):
    """All-pairs shifted LJ in FP64 (test force field, minimum image)."""
    shift = -4.0 * (rc ** -12 - rc ** -6)

    def fe(x):
        d = x[None, :, :] - x[:, None, :]
        d -= L * np.round(d / L)
        r2 = (d * d).sum(-1)
        np.fill_diagonal(r2, np.inf)
        inside = r2 < rc * rc
        ir2 = np.where(inside, 1.0 / r2, 0.0)
        ir6 = ir2 ** 3
        fs = np.where(inside, -(48 * ir6 * ir6 - 24 * ir6) * ir2, 0.0)
        F = (fs[:, :, None] * d).sum(1)
        U = 0.5 * np.where(inside, 4 * (ir6 * ir6 - ir6) + shift, 0.0).sum()
        return F, U
    return fe


@pytest.fixture(scope="module")
def lj108():
    x0, L = _lj_fcc()
    n = len(x0)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, (n, 3))
    v0 -= v0.mean(0)
    return x0, v0, np.ones(n), L, _lj(L)


def test_vv_harmonic_shadow_energy():
    """Velocity Verlet on x'' = -w^2 x conserves 1/2 v^2 + 1/2 w^2 (1 - w^2 dt^2/4) x^2
    exactly (the true energy only to O(dt^2))."""
    w2, dt = 1.3, 0.2
    x = np.array([[1.0, -0.4, 0.25]])
    v = np.array([[0.3, 0.7, -0.1]])
    m = np.array([1.0])
    x1, v1, _, _ = M.md_run(x, v, m, lambda y: -w2 * y, dt, 2000, com=False)
    c = 1.0 - w2 * dt * dt / 4.0
    H0 = 0.5 * (v ** 2).sum() + 0.5 * w2 * c * (x ** 2).sum()
    H1 = 0.5 * (v1 ** 2).sum() + 0.5 * w2 * c * (x1 ** 2).sum()
    assert abs(H1 - H0) < 1e-12 * H0
    E1 = 0.5 * (v1 ** 2).sum() + 0.5 * w2 * (x1 ** 2).sum()
    assert abs(E1 - H0) > 1e-4   # the plain energy is not the invariant


def test_remove_com():
    rng = np.random.default_rng(2)
    v = rng.normal(size=(50, 3)) + np.array([0.3, -0.2, 0.5])
    m = rng.uniform(0.5, 2.0, 50)
    K0 = M.kinetic(v, m)
    P0 = (m[:, None] * v).sum(0)
    M.remove_com(v, m)
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
    assert abs(M.kinetic(v, m) - (K0 - 0.5 * (P0 @ P0) / m.sum())) < 1e-10


def _ext_energy_dev(lj108, dt, t_end=0.8):
    x0, v0, m, L, fe = lj108
    n = len(x0)
    nhc = M.NHC2(n, 1.5, 0.2)
    H0 = M.kinetic(v0, m) + fe(x0)[1] + nhc.hamiltonian()
    _, _, _, rec = M.md_run(x0, v0, m, lambda y: fe(y)[0], dt, int(round(t_end / dt)), nhc=nhc,
                            com=True, box=(L, L, L), energy_fn=lambda y: fe(y)[1])
    H = np.array([k + u + h for k, u, h in rec])
    K = np.array([k for k, _, _ in rec])
    return np.abs(H - H0).max(), K


def test_nhc_extended_energy_second_order(lj108):
    """K + U + H_NHC is the conserved quantity of the chain dynamics; the
    symmetric splitting keeps it to O(dt^2) while the thermostat moves K by
    far more."""
    d1, K = _ext_energy_dev(lj108, 0.004)
    d2, _ = _ext_energy_dev(lj108, 0.002)
    assert K.max() - K.min() > 100.0          # the thermostat does work
    assert d1 < 1.0
    assert 3.0 < d1 / d2 < 5.0                # second order


def test_nhc_time_reversible(lj108):
    """(x, v, vxi) -> (x, -v, -vxi) retraces the trajectory (symmetric splitting).
    Kept short: the chain compresses phase space, so the reversed run amplifies
    round-off exponentially over longer stretches."""
    x0, v0, m, L, fe = lj108
    n = len(x0)
    dt, steps = 0.004, 20
    nhc = M.NHC2(n, 1.5, 0.2)
    x1, v1, _, _ = M.md_run(x0, v0, m, lambda y: fe(y)[0], dt, steps, nhc=nhc, com=True,
                            box=(L, L, L))
    xi = (nhc.xi1, nhc.xi2)
    nhc.vxi1, nhc.vxi2 = -nhc.vxi1, -nhc.vxi2
    x2, v2, _, _ = M.md_run(x1, -v1, m, lambda y: fe(y)[0], dt, steps, nhc=nhc, com=True,
                            box=(L, L, L))
    d = x2 - x0
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-8
    assert np.abs(v2 + v0).max() < 1e-7
    assert abs(nhc.xi1 - 1.0) < 1e-10 and abs(nhc.xi2 - 1.0) < 1e-10
    assert abs(xi[0] - 1.0) > 1e-4            # the chain did move


def test_replay_nve_step_f32_matches_fp64():
    """The FP32 single-step replay agrees with the FP64 step to FP32 round-off."""
    rng = np.random.default_rng(3)
    n, L, dt = 64, np.float32(7.5), 0.005
    x = rng.uniform(0, L, (n, 3)).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    F0 = rng.normal(size=(n, 3)).astype(np.float32) * 10
    F1 = rng.normal(size=(n, 3)).astype(np.float32) * 10
    m = rng.uniform(0.5, 2, n).astype(np.float32)
    xs, vs = M.replay_nve_step_f32(x, v, F0, F1, m, dt, (L, L, L))
    x64, v64 = M.vv_part1(x.astype(np.float64), v.astype(np.float64), F0.astype(np.float64),
                          m.astype(np.float64), dt)
    x64 = M.wrap(x64, (L, L, L))
    v64 = M.vv_part2(v64, F1.astype(np.float64), m.astype(np.float64), dt)
    assert np.abs(xs - x64).max() < 1e-5
    assert np.abs(vs - v64).max() < 1e-5
