"""MD loop restatement in numpy (FP64) -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of the device integrator; the product never
imports this module.  Restates, in the reference's order
(shortRange.inhomo/template/tools/simulation.hetero.cutoff/app/lj.nhc-xtc.cu
:158-196, called H/app/lj.nhc-xtc.cu below, H = that template directory):

    remove COM velocity   TranslationalFreedomRemover::remove
                          (H/src/Integrator_interface.cu:40-68)
    NHC half step         NoseHoover_Chains2::operator_L(dt/2)
                          (H/src/Thermostat.cu:23-53 reinit, 56-61 xi,
                           138-174 Cv, 209-243 G1/vxi1/G2, 255-276 operator_L)
    VV part 1             v += F dt/(2m); x += v dt
                          (H/src/Integrator.cu:273-324)
    force at the new x
    VV part 2             v += F dt/(2m)  (H/src/Integrator.cu:327-347)
    NHC half step

All arithmetic in FP64 (the reference runs it in FP32 ScalorType; the device
integrator keeps FP32 velocities and FP64 chain/kinetic sums), so device and
oracle agree to FP32 round-off per step.

Checks that pin this restatement (tests/test_oracle_md.py): the velocity
Verlet map on a harmonic oscillator conserves its closed-form shadow energy
1/2 v^2 + 1/2 w^2 (1 - w^2 dt^2 / 4) x^2 exactly; the full NVE/NHC step is
time-reversible (v, vxi -> -v, -vxi retraces the trajectory); the NHC extended
energy K + U + H_NHC is conserved to second order in dt; COM removal leaves
zero momentum and lowers K by |P|^2 / 2M.
"""
import numpy as np


def kinetic(v, m):
    """sum m v^2 / 2 (calKinetic, H/src/Thermostat.cu:177-207)."""
    return 0.5 * float(np.sum(m[:, None] * v * v))


def remove_com(v, m):
    """v -= sum(m v) / sum(m)  (H/src/Integrator_interface.cu:40-68)."""
    p = (m[:, None] * v).sum(0)
    v -= p / m.sum()
    return v


class NHC2:
    """Two-link Nose-Hoover chain (H/src/Thermostat.cu, NoseHoover_Chains2)."""

    def __init__(self, n_atoms, ref_T, tau_T):
        # reinit (H/src/Thermostat.cu:23-53)
        self.LL = 3.0 * n_atoms
        self.Q1 = tau_T * tau_T * ref_T / (4 * np.pi * np.pi) * self.LL
        self.Q2 = self.Q1 / self.LL
        self.xi1 = self.xi2 = 1.0
        self.vxi1 = self.vxi2 = 0.0
        self.refT = ref_T

    def _G1(self, h, K):
        self.vxi1 += (2.0 * K - self.LL * self.refT) / self.Q1 * h

    def _G2(self, h):
        self.vxi2 += (self.Q1 * self.vxi1 * self.vxi1 - self.refT) / self.Q2 * h

    def _vxi1(self, h):
        self.vxi1 *= np.exp(-self.vxi2 * h)

    def operator_L(self, h, v, m):
        """One chain step of length h (H/src/Thermostat.cu:255-276); scales v."""
        self._G2(0.5 * h)
        self._vxi1(0.25 * h)
        self._G1(0.5 * h, kinetic(v, m))
        self._vxi1(0.25 * h)
        v *= np.exp(-self.vxi1 * h)                 # Cv
        self.xi1 += self.vxi1 * h                   # xi
        self.xi2 += self.vxi2 * h
        self._vxi1(0.25 * h)
        self._G1(0.5 * h, kinetic(v, m))
        self._vxi1(0.25 * h)
        self._G2(0.5 * h)
        return v

    def hamiltonian(self):
        """HamiltonianContribution (H/src/Thermostat.cu:246-253)."""
        return (0.5 * self.Q1 * self.vxi1 ** 2 + 0.5 * self.Q2 * self.vxi2 ** 2
                + self.LL * self.refT * self.xi1 + self.refT * self.xi2)


def vv_part1(x, v, F, m, dt):
    """H/src/Integrator.cu:273-324: v += F dt/(2m); x += v dt."""
    v += F * (0.5 * dt / m)[:, None]
    x += v * dt
    return x, v


def vv_part2(v, F, m, dt):
    """H/src/Integrator.cu:327-347: v += F dt/(2m)."""
    v += F * (0.5 * dt / m)[:, None]
    return v


def wrap(x, box):
    """Back into [0, L) per dimension (normalizeSystem, H/src/MDSystem_interface.cu:19-36)."""
    if box is None:
        return x
    b = np.asarray(box, np.float64)
    return x - np.floor(x / b) * b


def md_run(x, v, m, force_fn, dt, n_steps, nhc=None, com=True, box=None, energy_fn=None):
    """n_steps of the reference loop from (x, v); force_fn(x) -> F (N, 3).

    Returns (x, v, F, record) where record lists per step
    (K, U or None, H_NHC) at the end of the step."""
    x = np.array(x, np.float64)
    v = np.array(v, np.float64)
    m = np.asarray(m, np.float64)
    F = force_fn(x)
    rec = []
    for _ in range(n_steps):
        if com:
            remove_com(v, m)
        if nhc is not None:
            nhc.operator_L(0.5 * dt, v, m)
        x, v = vv_part1(x, v, F, m, dt)
        x = wrap(x, box)
        F = force_fn(x)
        v = vv_part2(v, F, m, dt)
        if nhc is not None:
            nhc.operator_L(0.5 * dt, v, m)
        rec.append((kinetic(v, m), energy_fn(x) if energy_fn else None,
                    nhc.hamiltonian() if nhc is not None else 0.0))
    return x, v, F, rec


def replay_nve_step_f32(x, v, F0, F1, m, dt, box):
    """One NVE step in the device's FP32 operation order, given the forces the
    device used (F0 at x, F1 at the new x, both FP32): v_h = v + c F0,
    x' = x + v_h dt wrapped by one add/subtract of L, v' = v_h + c F1 with
    c = (dt/2) / m.  The device must reproduce it bit for bit."""
    f32 = np.float32
    x = np.asarray(x, f32)
    v = np.asarray(v, f32)
    c = (f32(0.5) * f32(dt) / np.asarray(m, f32))[:, None].astype(f32)
    vh = (v + (c * np.asarray(F0, f32)).astype(f32)).astype(f32)
    xn = (x + (vh * f32(dt)).astype(f32)).astype(f32)
    L = np.asarray(box, f32)[None, :]
    xn = np.where(xn < 0, (xn + L).astype(f32), np.where(xn >= L, (xn - L).astype(f32), xn))
    vn = (vh + (c * np.asarray(F1, f32)).astype(f32)).astype(f32)
    return xn.astype(f32), vn
