"""Pins the FP64 oracle (CPU) against closed forms, brute force, finite
differences, invariants and the reference's own fixtures.

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign
or index, a transposed operand, a wrong factor of 2) fails it.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ONE = np.ones((1, 1), np.float32)


def lj_closed(r, eps=1.0, sig=1.0):
    """u(r) and the radial force -du/dr of 4 eps [(s/r)^12 - (s/r)^6]."""
    s6 = (sig / r) ** 6
    u = 4 * eps * (s6 * s6 - s6)
    f = 24 * eps * (2 * s6 * s6 - s6) / r
    return u, f


def two_atoms(r, box=20.0):
    x = np.array([[5.0, 5.0, 5.0, 0], [5.0 + r, 5.0, 5.0, 0]], np.float32)
    return x, np.array([box] * 3, np.float32)


@pytest.mark.parametrize("r,eps,sig", [(0.95, 1.0, 1.0), (1.12, 1.0, 1.0), (1.6, 1.0, 1.0),
                                       (2.4, 1.0, 1.0), (0.9, 1.5, 0.8), (1.3, 0.5, 0.88)])
def test_two_body_closed_form(r, eps, sig):
    x, box = two_atoms(r)
    r = float(x[1, 0] - x[0, 0])          # the FP32 separation actually stored
    rc = np.full(2, 3.0, np.float32)
    off, j = O.build_pairs(x, box, rc, 0.3, method="brute")
    assert list(off) == [0, 1, 2] and list(j) == [1, 0]
    e = np.full((1, 1), eps, np.float32)
    s = np.full((1, 1), sig, np.float32)
    F, E, W, nin = O.forces(x, box, None, e, s, rc, off, j)
    u, f = lj_closed(r, float(e[0, 0]), float(s[0, 0]))
    assert nin == 2
    # force on atom 0 along +x is -f (repulsive f > 0 pushes atom 0 to -x)
    np.testing.assert_allclose(F[0], [-f, 0, 0], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(F[1], [f, 0, 0], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(E, [u / 2, u / 2], rtol=1e-13)
    # reference virial convention: W_xx = 1/2 F_i,x (x_j - x_i) per atom
    np.testing.assert_allclose(W[0, 0], 0.5 * (-f) * r, rtol=1e-13)
    np.testing.assert_allclose(W[:, 1:], 0.0, atol=1e-14)


def test_shift_and_cutoff_mask():
    x, box = two_atoms(2.0)
    sh = np.full((1, 1), 0.25, np.float32)
    # inside rc: energy includes the shift, outside: the pair is listed but masked
    off, j = O.build_pairs(x, box, np.full(2, 2.5, np.float32), 0.3)
    _, E, _, nin = O.forces(x, box, None, ONE, ONE, np.full(2, 2.5, np.float32), off, j, shift=sh)
    u, _ = lj_closed(float(x[1, 0] - x[0, 0]))
    np.testing.assert_allclose(E, (u + 0.25) / 2, rtol=1e-12)
    rc = np.array([1.8, 2.5], np.float32)     # asymmetric cutoffs: only atom 1 sees atom 0
    off, j = O.build_pairs(x, box, rc, 0.3)
    assert list(off) == [0, 1, 2]             # both lists hold the pair (rc + skin)
    F, E, _, nin = O.forces(x, box, None, ONE, ONE, rc, off, j)
    assert nin == 1 and E[0] == 0.0 and np.all(F[0] == 0.0) and F[1, 0] != 0.0


def test_minimum_image_across_boundary():
    box = np.array([10.0, 10.0, 10.0], np.float32)
    x = np.array([[0.2, 5, 5, 0], [9.3, 5, 5, 0]], np.float32)
    rc = np.full(2, 2.5, np.float32)
    off, j = O.build_pairs(x, box, rc, 0.3)
    assert off[-1] == 2
    F, _, _, _ = O.forces(x, box, None, ONE, ONE, rc, off, j)
    r = 0.2 + 10.0 - 9.3
    _, f = lj_closed(np.float64(np.float32(0.2)) + 10.0 - np.float64(np.float32(9.3)))
    # atom 0 is pushed towards +x (its neighbour sits at -0.9 through the boundary)
    assert F[0, 0] > 0 and abs(F[0, 0] - f) < 1e-9 * f
    assert r > 0


def brute_pairs_f64(xyz, box, rc, skin):
    """Independent check: explicit periodic images in FP64 (no minimum image)."""
    n = len(xyz)
    out = []
    for i in range(n):
        d = xyz[:, None, :3].astype(np.float64) - xyz[i, :3].astype(np.float64)
        sh = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)])
        dd = d + sh[None] * box.astype(np.float64)
        r2 = (dd ** 2).sum(-1).min(1)
        rl = float(rc[i]) + skin
        out.append(np.nonzero((r2 < rl * rl) & (np.arange(n) != i))[0])
    return out


def test_pairs_cells_equal_brute_and_explicit_images():
    from sre import configs
    w = configs.slab(seed=5, n=3000, box=(14.0, 14.0, 40.0), nz=32)
    box = w.boxes[0]
    rc, _ = O.assign_rc(w.xyz, box, w.rc_tables[0])
    o1, j1 = O.build_pairs(w.xyz, box, rc, w.skin, method="cells")
    o2, j2 = O.build_pairs(w.xyz, box, rc, w.skin, method="brute")
    assert np.array_equal(o1, o2) and np.array_equal(j1, j2)
    ref = brute_pairs_f64(w.xyz[:400], box, rc[:400], w.skin)  # exact maths, FP64
    sub_o, sub_j = O.build_pairs(w.xyz[:400], box, rc[:400], w.skin, method="brute")
    mism = sum(len(set(ref[i]) ^ set(sub_j[sub_o[i]:sub_o[i + 1]])) for i in range(400))
    assert mism == 0        # no pair sits within FP32 rounding of rc + skin here


def test_lj27_fixture_against_explicit_images():
    """The reference's 29-atom real-error example (lj.27.gro, 6^3 box)."""
    d = np.load(os.path.join(GOLD, "lj27.npz"))
    xyz, box = d["xyz"], d["box"]
    n = len(xyz)
    rc = np.full(n, 2.9, np.float32)
    off, j = O.build_pairs(xyz, box, rc, 0.05)
    F, E, W, nin = O.forces(xyz, box, None, ONE, ONE, rc, off, j)
    # explicit-image FP64 sum, each image counted once (rc < L/2)
    Fr = np.zeros((n, 3))
    Er = np.zeros(n)
    sh = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]) * 6.0
    X = xyz.astype(np.float64)
    for i in range(n):
        for k in range(n):
            if k == i:
                continue
            for s in sh:
                dv = X[k] + s - X[i]
                r = np.sqrt(dv @ dv)
                if r < 2.9:
                    u, f = lj_closed(r)
                    Fr[i] -= f * dv / r
                    Er[i] += u / 2
    np.testing.assert_allclose(F, Fr, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(E, Er, rtol=1e-12, atol=1e-12)


def _fcc_small():
    from sre import configs
    return configs.plain_fcc(seed=9, ncell=5)


def test_force_is_minus_gradient_of_energy():
    w = _fcc_small()
    box = w.boxes[0]
    rc = np.full(w.n, 2.5, np.float32)
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    F, _, _, _ = O.forces(w.xyz, box, None, ONE, ONE, rc, off, j)
    h = 2.0 ** -12                         # exact offset for |x| < 16
    rng = np.random.default_rng(0)
    for i in rng.choice(w.n, 6, replace=False):
        for a in range(3):
            xp, xm = w.xyz.copy(), w.xyz.copy()
            xp[i, a] += h
            xm[i, a] -= h
            Ep = O.forces(xp, box, None, ONE, ONE, rc, off, j)[1].sum()
            Em = O.forces(xm, box, None, ONE, ONE, rc, off, j)[1].sum()
            fd = -(Ep - Em) / (2 * h)
            assert abs(fd - F[i, a]) <= 2e-3 * max(1.0, abs(F[i, a])), (i, a, fd, F[i, a])


def test_virial_is_box_scaling_derivative():
    """W_aa (reference convention) = dE/d lambda_a under x_a -> (1+lambda) x_a."""
    w = _fcc_small()
    box = w.boxes[0].astype(np.float64)
    rc = np.full(w.n, 2.5, np.float32)
    off, j = O.build_pairs(w.xyz, w.boxes[0], rc, w.skin)
    _, E0, W, _ = O.forces(w.xyz, w.boxes[0], None, ONE, ONE, rc, off, j)
    lam = 1e-3
    for a in range(3):
        vals = []
        for sgn in (1, -1):
            x = w.xyz.astype(np.float64).copy()
            b = box.copy()
            x[:, a] *= 1 + sgn * lam
            b[a] *= 1 + sgn * lam
            vals.append(O.forces(x.astype(np.float32), b.astype(np.float32), None, ONE, ONE,
                                 rc, off, j)[1].sum())
        fd = (vals[0] - vals[1]) / (2 * lam)
        assert abs(fd - W[:, a].sum()) <= 1e-2 * abs(W[:, a].sum()), (a, fd, W[:, a].sum())


def test_translation_and_permutation_invariance():
    from sre import configs
    w = configs.slab(seed=6, n=4000, box=(14.0, 14.0, 50.0), nz=40)
    box = w.boxes[0]
    # positions on a 2^-10 grid, so that the translation below is exact in FP32
    w.xyz[:, :3] = np.mod(np.round(w.xyz[:, :3] * 1024.0) / 1024.0, box)
    rc = np.full(w.n, 2.8, np.float32)   # uniform rc: invariant under translation
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    F, E, _, _ = O.forces(w.xyz, box, None, ONE, ONE, rc, off, j)
    # sum of forces vanishes for a uniform cutoff (Newton's third law holds)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).sum()
    perm = np.random.default_rng(1).permutation(w.n)
    xp = w.xyz[perm]
    o2, j2 = O.build_pairs(xp, box, rc, w.skin)
    F2, E2, _, _ = O.forces(xp, box, None, ONE, ONE, rc, o2, j2)
    np.testing.assert_allclose(F2, F[perm], rtol=1e-11, atol=1e-10)
    # translation by a multiple of the cell used for positions keeps FP32 values exact
    xt = w.xyz.copy()
    xt[:, 2] = np.mod(xt[:, 2] + np.float32(8.0), np.float32(50.0))
    o3, j3 = O.build_pairs(xt, box, rc, w.skin)
    F3, E3, _, _ = O.forces(xt, box, None, ONE, ONE, rc, o3, j3)
    np.testing.assert_allclose(F3, F, rtol=1e-9, atol=1e-9)


def test_rc_assignment_bins():
    box = np.array([10.0, 10.0, 20.0], np.float32)
    table = np.arange(1, 9, dtype=np.float32).reshape(1, 1, 8) * 0.5 + 2.0
    z = np.array([0.0, 2.49, 2.5, 19.99, -0.5, 20.0, 20.5], np.float32)
    x = np.stack([np.full_like(z, 1.0), np.full_like(z, 1.0), z, np.zeros_like(z)], 1)
    rc, bz = O.assign_rc(x, box, table)
    # bin width 2.5; -0.5 wraps to 19.5 (bin 7); 20.0 wraps to 0 (bin 0)
    assert list(bz) == [0, 0, 1, 7, 7, 0, 0]
    np.testing.assert_array_equal(rc, table.ravel()[bz])


def test_pair_count_matches_published_statistics():
    """PRE Table I: 2.9e6 (rc 3.5) and 7.8e6 (rc 5.0) list entries per step for the
    16000-atom slab (statistical: a trajectory average; here one stored frame)."""
    d = np.load(os.path.join(GOLD, "conf16000.npz"))
    xyz, box = d["xyz"], d["box"]
    for rc_v, pub in [(3.5, 2.9e6), (5.0, 7.8e6)]:
        rc = np.full(len(xyz), rc_v, np.float32)
        off, _ = O.build_pairs(xyz, box, rc, 0.49)
        assert abs(off[-1] - pub) / pub < 0.10, (rc_v, off[-1])


# ---------------------------------------------------------------- tail term --

def test_tail_homogeneous_closed_form():
    rho0, nz, Lz = 0.7, 50, 40.0
    rho = np.full((1, nz), rho0)
    for R in (2.5, 3.7, 5.0):
        for z in (0.3, 17.1, 39.9):
            F, E = O.tail_point(z, R, 0, ONE, ONE, rho, Lz)
            classic = 8.0 / 3.0 * np.pi * rho0 * ((1.0 / 3.0) * R ** -9 - R ** -3)
            assert abs(F) < 1e-12
            assert abs(E - classic) < 1e-12 * abs(classic)


def _slab_rho(nz, Lz, z0, z1, rl, rv):
    zc = (np.arange(nz) + 0.5) * Lz / nz
    return np.where((zc > z0) & (zc < z1), rl, rv)[None, :]


def test_tail_slab_against_quadrature():
    """Mean error force and tail energy of a planar slab against direct 2-D
    numerical integration of the pair force/energy beyond rc (scipy)."""
    from scipy import integrate
    nz, Lz = 200, 100.0
    rho = _slab_rho(nz, Lz, 30.0, 70.0, 0.7, 0.02)
    R = 3.0

    def rho_at(zz):
        k = int(np.floor(np.mod(zz, Lz) / (Lz / nz)))
        return rho[0, min(k, nz - 1)]

    def u(r):
        return 4 * (r ** -12 - r ** -6)

    def du(r):
        return -24 * (2 * r ** -13 - r ** -7)

    for z in (28.0, 31.5, 50.0, 69.0):
        F, E = O.tail_point(z, R, 0, ONE, ONE, rho, Lz)
        # integrate s over |s| <= 250 (the force integrand decays as s^-5, the
        # energy one as s^-4) and rho_perp; breakpoints at the density jumps
        faces = [f + k * Lz - z for f in (30.0, 70.0) for k in range(-3, 4)]
        edges = sorted({-250.0, 250.0, *[e for e in faces if -250.0 < e < 250.0]})
        fq = eq = 0.0
        for a, b in zip(edges[:-1], edges[1:]):
            def fz(s):
                lo = np.sqrt(max(R * R - s * s, 0.0))
                val, _ = integrate.quad(lambda p: 2 * np.pi * p * du(np.hypot(p, s)) * s
                                        / np.hypot(p, s), lo, np.inf, epsabs=1e-13)
                return rho_at(z + s) * val

            def ez(s):
                lo = np.sqrt(max(R * R - s * s, 0.0))
                val, _ = integrate.quad(lambda p: 2 * np.pi * p * u(np.hypot(p, s)), lo, np.inf,
                                        epsabs=1e-13)
                return rho_at(z + s) * val
            pts = [x for x in (-R, R) if a < x < b]
            fq += integrate.quad(fz, a, b, points=pts or None, limit=200, epsabs=1e-12)[0]
            eq += integrate.quad(ez, a, b, points=pts or None, limit=200, epsabs=1e-12)[0]
        eq *= 0.5
        # force on the atom = sum of -grad u at r_i - r_j: f_z = -u'(r) (z_i - z_j)/r
        #                   = u'(r) s / r with s = z_j - z_i (as integrated above)
        assert abs(F - fq) < 2e-6 * max(1e-3, abs(fq)), (z, F, fq)
        assert abs(E - eq) < 2e-6 * abs(eq), (z, E, eq)


def test_tail_force_points_into_the_liquid():
    nz, Lz = 100, 100.0
    rho = _slab_rho(nz, Lz, 30.0, 70.0, 0.7, 0.02)
    F_left, _ = O.tail_point(29.0, 3.0, 0, ONE, ONE, rho, Lz)    # just outside the left face
    F_right, _ = O.tail_point(71.0, 3.0, 0, ONE, ONE, rho, Lz)
    F_mid, _ = O.tail_point(50.0, 3.0, 0, ONE, ONE, rho, Lz)
    assert F_left > 0 and F_right < 0 and abs(F_mid) < 1e-6
    assert abs(F_left + F_right) < 1e-9


# ---------------------------------------------------------------- gradients --

def test_pair_gradient_against_central_differences():
    from sre import configs
    w = configs.slab(seed=8, n=3000, box=(14.0, 14.0, 40.0), nz=32, binary=True)
    box = w.boxes[0]
    rc, _ = O.assign_rc(w.xyz, box, w.rc_tables[0])
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    g = np.random.default_rng(3).normal(size=(w.n, 3))
    dLdE = 0.7
    ge, gs = O.grad_pairs(w.xyz, box, w.types, w.eps, w.sigma, rc, off, j, dLdE, g)

    def L(eps, sig):
        F, E, _, _ = O.forces(w.xyz, box, w.types, eps, sig, rc, off, j)
        return dLdE * E.sum() + (g * F).sum()

    for a, b in [(0, 0), (0, 1), (1, 1)]:
        for which, grad in (("eps", ge), ("sig", gs)):
            base = (w.eps if which == "eps" else w.sigma).astype(np.float64)
            h = 2.0 ** -12
            vals = []
            for sgn in (1, -1):
                p = base.copy()
                p[a, b] += sgn * h
                p[b, a] = p[a, b]
                args = (p.astype(np.float32), w.sigma) if which == "eps" else (w.eps, p.astype(np.float32))
                vals.append(L(*args))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grad[a, b]) <= 1e-4 * max(1.0, abs(grad[a, b])), (which, a, b, fd, grad[a, b])


def test_tail_gradient_rc_bins_against_differences():
    nz, Lz = 16, 40.0
    box = np.array([12.0, 12.0, Lz], np.float32)
    rng = np.random.default_rng(4)
    n = 400
    xyz = np.zeros((n, 4), np.float32)
    xyz[:, :3] = rng.uniform(0, 1, (n, 3)) * box
    rho = _slab_rho(nz, Lz, 12.0, 28.0, 0.7, 0.05)
    rcz = np.linspace(2.5, 4.0, nz).astype(np.float32)
    g = rng.normal(size=(n, 3))
    dLdE = 0.3
    ge, gs, grc = O.grad_tail(xyz, box, None, ONE, ONE, rcz, rho, dLdE, g, h=1e-4)
    t0 = O.tail_atoms(xyz, box, None, ONE, ONE, rcz, rho)
    for b in (0, 5, 11):
        vals = []
        for sgn in (1, -1):
            r = rcz.astype(np.float64).copy()
            h = 2.0 ** -10
            r[b] += sgn * h
            t = O.tail_atoms(xyz, box, None, ONE, ONE, r.astype(np.float32), rho)
            vals.append(dLdE * t[:, 1].sum() + (g[:, 2] * t[:, 0]).sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grc[b]) <= 1e-5 * max(1.0, abs(fd)), (b, fd, grc[b])
    # d/d eps of a term linear in eps equals the term itself at eps = 1
    tot = dLdE * t0[:, 1].sum() + (g[:, 2] * t0[:, 0]).sum()
    assert abs(ge[0, 0] - tot) <= 1e-10 * max(1.0, abs(tot))
