"""Differentiation at the binary-slab configuration (N = 64000, A/B 80:20,
256 rc(z) bins): gradients of the energy E and of the force-matching loss
L = sum_i |F_i - F_target,i|^2 with respect to the 3 epsilons, the 3 sigmas
and the 256 rc bins, in both precisions.  F_target is the FP64 oracle's force
at the workload's positions with every pair parameter scaled by
1 + 0.02 N(0,1) (seeded, configs.perturbed_params).

Two checks:
  * the engine's adjoint (sre_adjoint) against the FP64 oracle's analytic
    pair gradient plus its tail gradient (same dL/dF vector);
  * the autograd bridge (lj_energy_forces, with the rc tensor passed so the
    rc path runs: tables set, lists rebuilt) against central finite
    differences with h = 1e-4 of E and L evaluated through the engine.

The cutoff mask is held fixed in the gradient's definition (d/d rc acts
through the tail correction), so the rc finite differences are compared on
the bins whose +-h change moves no pair across a cutoff (the in-rc pair
count is unchanged); every other bin is checked against the oracle by the
first test."""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine, configs

pytestmark = pytest.mark.gpu

H = 1e-4                      # finite-difference step of the configuration
PAIRS = [(0, 0), (0, 1), (1, 1)]


@pytest.fixture(scope="module")
def c3():
    """The workload, the oracle's lists, and F_target."""
    w = configs.binary_slab()
    box = w.boxes[0]
    rcz = w.rc_tables[0]
    rc, _ = O.assign_rc(w.xyz, box, rcz)
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    et, st = configs.perturbed_params(w)
    tail_t = O.tail_atoms(w.xyz, box, w.types, et, st, rcz.ravel(), w.rho[0])
    Ft, _, _, _ = O.forces(w.xyz, box, w.types, et, st, rc, off, j, tail=tail_t)
    return dict(w=w, rc=rc, off=off, j=j, Ft=Ft)


def oracle_gradient(c, dLdE, dLdF):
    w = c["w"]
    box = w.boxes[0]
    pe, ps = O.grad_pairs(w.xyz, box, w.types, w.eps, w.sigma, c["rc"], c["off"], c["j"],
                          dLdE, dLdF)
    te, ts, trc = O.grad_tail(w.xyz, box, w.types, w.eps, w.sigma,
                              w.rc_tables[0].ravel(), w.rho[0], dLdE, dLdF, h=H)
    return pe + te, ps + ts, trc


def engine_for(w, prec):
    e = Engine(n_types=2, skin=w.skin, precision=prec)
    w.setup(e, tail=True)
    e.rebuild()
    return e


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_adjoint_matches_oracle_at_binary_slab_64k(torch_cuda, c3, prec):
    import torch
    w = c3["w"]
    e = engine_for(w, prec)
    e.compute(stats=True, tail=True)
    F = e.forces()
    n_rc = w.rc_tables[0].size
    assert n_rc == 256
    # tolerances: the mixed adjoint evaluates pair terms in FP32 (relative
    # ~1e-6 per term), the FP64 form agrees to round-off
    tol = 2e-5 if prec == "mixed" else 1e-8
    for name, dLdE, dLdF in (("E", 1.0, None), ("L", 0.0, 2.0 * (F - c3["Ft"]))):
        ge, gs, grc = e.adjoint(dLdE=[dLdE],
                                dLdF=None if dLdF is None else torch.from_numpy(dLdF).cuda(),
                                tail=True, n_rc=n_rc)
        re, rs, rrc = oracle_gradient(c3, dLdE, dLdF)
        for a, b in PAIRS:
            assert abs(ge[a, b] - re[a, b]) <= tol * np.abs(re).max(), (name, "eps", a, b)
            assert abs(gs[a, b] - rs[a, b]) <= tol * np.abs(rs).max(), (name, "sigma", a, b)
        np.testing.assert_allclose(grc, rrc, rtol=1e-6, atol=1e-6 * np.abs(rrc).max(),
                                   err_msg=name)
        print(f"{prec} d{name}: eps {[ge[p] for p in PAIRS]}, sigma {[gs[p] for p in PAIRS]}")
    e.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_autograd_matches_finite_differences_at_binary_slab_64k(torch_cuda, c3, prec):
    """Central differences, h = 1e-4, through the same engine calls the
    bridge makes.  Parameters and rc values travel as FP32, so each
    difference quotient divides by the step FP32 actually represents.
    Tolerances: sigma's third derivative (sigma^12) gives a truncation error
    of ~1e-4 of the derivative at h = 1e-4 (the FP64 oracle's own central
    differences show 4e-5 to 9e-5 on this workload); the mixed form's
    force constants are FP32, so a parameter step of 2h is represented to
    ~2^-24 / 2h = 3e-4 of itself, which bounds its difference quotients."""
    import torch
    from sre.autograd import lj_energy_forces, pair_matrix
    w = c3["w"]
    Ft = torch.from_numpy(c3["Ft"]).cuda()
    e = engine_for(w, prec)
    pe = torch.tensor([w.eps[a, b] for a, b in PAIRS], dtype=torch.float64,
                      requires_grad=True)
    ps = torch.tensor([w.sigma[a, b] for a, b in PAIRS], dtype=torch.float64,
                      requires_grad=True)
    rc0 = w.rc_tables[0].ravel().astype(np.float64)
    rct = torch.tensor(rc0, requires_grad=True)

    def losses(pe_, ps_, rc_):
        E, F = lj_energy_forces(e, pair_matrix(pe_, 2), pair_matrix(ps_, 2), rc_, True)
        return E.sum(), ((F - Ft) ** 2).sum()

    E, L = losses(pe, ps, rct)
    grads = {}
    for name, val in (("E", E), ("L", L)):
        g = torch.autograd.grad(val, (pe, ps, rct), retain_graph=True)
        grads[name] = [t.numpy().copy() for t in g]

    def values(pe_, ps_, rc_):
        with torch.no_grad():
            E2, L2 = losses(pe_, ps_, rc_)
        return {"E": float(E2), "L": float(L2)}

    tol_par = {"eps": 3e-4 if prec == "fp64" else 2e-3,
               "sigma": 3e-4 if prec == "fp64" else 2e-3}
    for k in range(3):
        for which in ("eps", "sigma"):
            out = []
            for sgn in (1, -1):
                a_e, a_s = pe.detach().clone(), ps.detach().clone()
                p = a_e if which == "eps" else a_s
                p[k] = float(np.float32(p[k].item() + sgn * H))
                out.append((values(a_e, a_s, rct.detach()), p[k].item()))
            step = out[0][1] - out[1][1]
            for name in ("E", "L"):
                fd = (out[0][0][name] - out[1][0][name]) / step
                g = grads[name][0 if which == "eps" else 1][k]
                assert abs(fd - g) <= tol_par[which] * abs(g), (prec, name, which, k, fd, g)

    # rc bins: the in-rc pair count tells whether a +-h step moved a pair
    # across a cutoff (it can only grow with rc)
    _, nin0, _ = e.neighbor_count()
    # rounding that a mixed-form difference quotient cannot avoid, per bin:
    # the tail force enters each atom's fixed-point force rounded to 2^-32
    # (<= 2^-33 per evaluation), and a rebuilt list that gains or loses a
    # candidate regroups FP32 per-lane energy sums (<= 2^-22 of |E_i|)
    noise = {"E": np.zeros(rc0.size), "L": np.zeros(rc0.size)}
    if prec == "mixed":
        z = np.mod(w.xyz[:, 2].astype(np.float64), w.boxes[0][2])
        bins = np.minimum((z * rc0.size / w.boxes[0][2]).astype(int), rc0.size - 1)
        Fcur = e.forces()
        Ei = np.abs(e.atom_stats()[:, 0])
        np.add.at(noise["L"], bins, 2.0 * np.abs(Fcur[:, 2] - c3["Ft"][:, 2]) * 2.0 ** -32)
        np.add.at(noise["E"], bins, Ei * 2.0 ** -22)
    checked = 0
    for b in range(rc0.size):
        out = []
        for sgn in (1, -1):
            r = rct.detach().clone()
            r[b] = float(np.float32(rc0[b] + sgn * H))
            v = values(pe.detach(), ps.detach(), r)
            out.append((v, r[b].item(), e.neighbor_count()[1]))
        if out[0][2] != nin0 or out[1][2] != nin0:
            continue
        step = out[0][1] - out[1][1]
        for name in ("E", "L"):
            fd = (out[0][0][name] - out[1][0][name]) / step
            g = grads[name][2][b]
            scale = np.abs(grads[name][2]).max()
            tol = 1e-4 * abs(g) + 1e-6 * scale + noise[name][b] / step
            assert abs(fd - g) <= tol, (prec, name, b, fd, g)
        checked += 1
    print(f"{prec}: rc bins checked by finite differences: {checked} of {rc0.size}")
    assert checked >= 48      # 57 such bins on this workload (oracle distances)
    e.close()
