"""Adaptive-cutoff producer restated in numpy/scipy (FP64) -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of the device rc-table producer; the product
never imports this module.  Restates (H = shortRange.inhomo/template/tools/
simulation.hetero.cutoff):

  DensityProfile_PiecewiseConst  grid (H/src/DensityProfile.cu:27-57: n = int(L/refh),
                                 made odd), deposit (:69-90), calculate (:92-99)
  AdaptRCut::reinit              candidate list rc = rcmin, rcmin+step, ... <= rcmax
                                 (H/src/AdaptRCut.cu:96-163)
  AdaptRCut::makeS1k             homogeneity kernel (:265-312):
                                 S1(k) = (2/k) P1 int_rc^30 x^-13 sin(2 pi k x) dx,
                                 S1(0) = 4 pi P1 / (11 rc^11),  P1 = 16*36 eps sigma^12
  AdaptRCut::makeS2k             mean-error-force kernel (:314-416):
                                 S2_a(k) = i k_a G(k),
                                 G(k) = 2 pi (P2/k) int_rc^30 x^-5 sin(2 pi k x) dx
                                        - 2 pi rc^2 (4 eps sigma^6/rc^6)
                                          (2 cos(phi)/phi - 2 sin(phi)/phi^2) / k,
                                 phi = 2 pi k rc, P2 = 8 eps sigma^6, S2(0) = 0
  AdaptRCut::calError            rho_k = FFT(rho) V/n; e1 = IFFT(S1 rho_k)/V,
                                 e2_a = IFFT(S2_a rho_k)/V (FFTW backward, unnormalised);
                                 error = sqrt(e1 + |e2|^2)  (:182-236)
  AdaptRCut::calRCutOnePoint     end checks then bisection over candidates (:418-460)
  AdaptRCut::refineRCut          index = max over the 3x3x3 periodic neighbourhood (:470-508)

The integrals use scipy's QAWO (quad with weight='sin'), not the reference's
adaptive Gauss rule; both converge to 1e-12.

Checks that pin it (tests/test_oracle_arc.py): a uniform density gives the
closed-form homogeneous error sqrt(4 pi P1 rho / (11 rc^11)) everywhere, and
the reference's own liquid-region output matches that value; S1 and G equal
the 3-D Fourier transforms of |f|^2 and of the dispersion force, evaluated
independently with spherical Bessel functions; a sinusoidal density gives
e1 = rho0 S1(0) + a S1(k0) cos(2 pi k0 x) exactly; for a planar profile the
mean error force equals the closed-form planar tail force of oracle.oracle
(independent formulation); the refine equals scipy.ndimage's wrapped maximum
filter; the bisection equals a brute-force search when the error decreases
with rc.  Pressure correction: the trace of the kernels equals the transform of
the isotropic virial -24 r^-6, and for rho(z) the FFT route equals the direct
planar double integral with closed-form transverse kernels (both components).
"""
import numpy as np
from scipy import integrate

P1 = 16.0 * 36.0       # eps = sigma = 1 (makeS1k(1., 1.))
P2 = 8.0
X_MAX = 30.0


def grid_dims(box, refh):
    """DensityProfile_PiecewiseConst::reinit: n = unsigned(L/refh), odd."""
    dims = []
    for L in box:
        n = int(float(np.float32(L)) / refh)
        if n % 2 == 0:
            n += 1
        dims.append(n)
    return tuple(dims)


def rc_list(rcmin, rcmax, rcstep):
    """for (tmprc = rcmin; tmprc <= rcmax; tmprc += rcstep) -- accumulated."""
    out, t = [], float(rcmin)
    while t <= rcmax:
        out.append(t)
        t += rcstep
    return np.array(out)


def deposit(counts, xyz, box, dims):
    """Adds one sample (H/src/DensityProfile.cu:69-90); bin index clamped."""
    b = np.asarray([float(np.float32(v)) for v in box])
    h = b / np.asarray(dims)
    idx = []
    for a in range(3):
        t = np.asarray(xyz[:, a], np.float64)
        t = np.where(t >= b[a], t - b[a], np.where(t < 0, t + b[a], t))
        idx.append(np.minimum((t / h[a]).astype(np.int64), dims[a] - 1))
    np.add.at(counts, (idx[0], idx[1], idx[2]), 1.0)
    return counts


def density(counts, ndata, box, dims):
    """calculate(): counts / (dV * ndata)."""
    b = np.asarray([float(np.float32(v)) for v in box])
    dv = np.prod(b / np.asarray(dims))
    return counts / (dv * ndata)


def k_grid(box, dims):
    """k vectors at FFT index positions (signed frequency / L)."""
    b = np.asarray([float(np.float32(v)) for v in box])
    ks = [np.fft.fftfreq(n, d=1.0 / n) / b[a] for a, n in enumerate(dims)]
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    return kx, ky, kz


def _int_sin(power, k, a, b):
    """int_a^b x^-power sin(2 pi k x) dx (QAWO)."""
    return integrate.quad(lambda x: x ** (-power), a, b, weight="sin", wvar=2 * np.pi * k,
                          epsabs=1e-17, epsrel=1e-13, limit=400)[0]


def kernels(box, dims, rcs):
    """S1[c] (real) and G[c] on the k grid (makeS1k / makeS2k)."""
    kx, ky, kz = k_grid(box, dims)
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    S1 = np.empty((len(rcs),) + k.shape)
    G = np.empty((len(rcs),) + k.shape)
    flat = k.ravel()
    uk, inv = np.unique(flat, return_inverse=True)
    for c, rc in enumerate(rcs):
        s1 = np.empty(uk.size)
        g = np.empty(uk.size)
        for q, kk in enumerate(uk):
            if kk == 0.0:
                s1[q] = 4.0 * np.pi * P1 / 11.0 * rc ** -11
                g[q] = 0.0
                continue
            s1[q] = 2.0 / kk * P1 * _int_sin(13, kk, rc, X_MAX)
            phi = 2.0 * np.pi * kk * rc
            size = (-2.0 * np.pi * rc * rc * (4.0 / rc ** 6)
                    * (2.0 * np.cos(phi) / phi - 2.0 * np.sin(phi) / (phi * phi))) / kk
            g[q] = 2.0 * np.pi * P2 / kk * _int_sin(5, kk, rc, X_MAX) + size
        S1[c] = s1[inv].reshape(k.shape)
        G[c] = g[inv].reshape(k.shape)
    return S1, G, (kx, ky, kz)


def cal_error(rho, box, S1, G, kvec, parts=False):
    """error[c] = sqrt(e1 + |e2|^2) on the grid (calError)."""
    b = np.asarray([float(np.float32(v)) for v in box])
    V = float(np.prod(b))
    n = rho.size
    rhok = np.fft.fftn(rho) * (V / n)
    kx, ky, kz = kvec
    err = np.empty(S1.shape)
    e1s, e2s = [], []
    for c in range(S1.shape[0]):
        back = lambda a: np.fft.ifftn(a).real * n / V       # FFTW backward, / V
        e1 = back(S1[c] * rhok)
        e2 = [back(1j * kk * G[c] * rhok) for kk in (kx, ky, kz)]
        err[c] = np.sqrt(e1 + e2[0] ** 2 + e2[1] ** 2 + e2[2] ** 2)
        if parts:
            e1s.append(e1)
            e2s.append(np.stack(e2))
    if parts:
        return err, np.array(e1s), np.array(e2s)
    return err


def cal_rcut(err, prec):
    """calRCutOnePoint for every cell: index of the chosen candidate."""
    nrc = err.shape[0]
    flat = err.reshape(nrc, -1)
    out = np.empty(flat.shape[1], np.int64)
    for i in range(flat.shape[1]):
        a, b = 0, nrc - 1
        if flat[a, i] - prec <= 0:
            out[i] = a
        elif flat[b, i] - prec >= 0:
            out[i] = b
        else:
            while b - a > 1:
                c = (a + b) // 2
                if flat[c, i] > prec:
                    a = c
                else:
                    b = c
            out[i] = b
    return out.reshape(err.shape[1:])


def refine(idx):
    """refineRCut: max over the periodic 3x3x3 neighbourhood."""
    out = np.zeros_like(idx)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                out = np.maximum(out, np.roll(idx, (dx, dy, dz), axis=(0, 1, 2)))
    return out


def _int_bessel(nu, k, a, b):
    """int_a^b -24 x^-4.5 J_nu(2 pi k x) dx (FF_I1 / FF_I5, exact Bessel)."""
    from scipy import special
    return integrate.quad(lambda x: -24.0 * x ** -4.5 * special.jv(nu, 2 * np.pi * k * x),
                          a, b, epsabs=1e-17, epsrel=1e-13, limit=800)[0]


def pressure_kernels(box, dims, rcs):
    """Kxx, Kyy, Kzz[c] on the k grid (PressureCorrection::makeKernel,
    H/src/PressureCorrection.cu:262-355): with i1, i5 = int_rc^30 -24 x^-4.5
    J_{1/2 or 5/2}(2 pi k x) dx and the direction of k,
    K_aa = (pi / sqrt(k)) (2/3 i1 + (2/3 - 2 khat_a^2) i5);  K(0) = 0."""
    kx, ky, kz = k_grid(box, dims)
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    K = np.zeros((3, len(rcs)) + k.shape)
    uk, inv = np.unique(k.ravel(), return_inverse=True)
    for c, rc in enumerate(rcs):
        i1 = np.zeros(uk.size)
        i5 = np.zeros(uk.size)
        for q, kk in enumerate(uk):
            if kk == 0.0:
                continue
            i1[q] = _int_bessel(0.5, kk, rc, X_MAX)
            i5[q] = _int_bessel(2.5, kk, rc, X_MAX)
        I1 = i1[inv].reshape(k.shape)
        I5 = i5[inv].reshape(k.shape)
        with np.errstate(divide="ignore", invalid="ignore"):
            scale = np.where(k > 0, np.pi / np.sqrt(np.where(k > 0, k, 1.0)), 0.0)
            for a, ka in enumerate((kx, ky, kz)):
                kh2 = np.where(k > 0, (ka / np.where(k > 0, k, 1.0)) ** 2, 0.0)
                K[a, c] = scale * (2.0 / 3.0 * I1 + (2.0 / 3.0 - 2.0 * kh2) * I5)
    return K


def pressure_correction(rho, box, K, idx):
    """PressureCorrection::correction (H/src/PressureCorrection.cu:176-211):
    p_aa = sum_cells 1/2 conv_aa[idx](cell) rho(cell) dV / V^2,
    conv = IFFT(K rho_k) (unnormalised), rho_k = FFT(rho) V / n."""
    b = np.asarray([float(np.float32(v)) for v in box])
    V = float(np.prod(b))
    n = rho.size
    dv = np.prod(b / np.asarray(rho.shape))
    rhok = np.fft.fftn(rho) * (V / n)
    out = np.zeros(3)
    for a in range(3):
        conv = np.array([np.fft.ifftn(K[a, c] * rhok).real * n for c in range(K.shape[1])])
        sel = np.take_along_axis(conv, idx[None], 0)[0]
        out[a] = (0.5 * sel * rho / V * dv).sum() / V
    return out


def homogeneous_error(rho, rc):
    """Closed form for a uniform density (S1(0) term alone)."""
    return np.sqrt(4.0 * np.pi * P1 / 11.0 * rho * rc ** -11.0)
