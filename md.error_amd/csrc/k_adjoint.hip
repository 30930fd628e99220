// k_adjoint.hip -- reverse-mode derivative of
//     L = sum_r dLdE_r E_r + sum_i dLdF_i . F_i
// with respect to the LJ parameters eps_ab, sigma_ab and the rc lattice.
// The reference has no derivative code; this is new.
//
// Pair terms are recomputed from the neighbor list (nothing stored by the
// forward pass).  For a pair (i, j) inside rc_i with s6 = (sig/r)^6,
// s12 = s6^2, d = x_j - x_i, g = dLdF_i:
//     E_i/2 part:  d/deps = 2 (s12 - s6),         d/dsig = 12 eps (2 s12 - s6) / sig
//     F_i part:    d/deps = -24 (2 s12 - s6) (g.d) / r^2
//                  d/dsig = -144 eps (4 s12 - s6) (g.d) / (sig r^2)
// The cutoff mask is held fixed, so the pair part has no rc derivative.  The
// tail correction (bin mode) is a function of (eps, sig, rc_bin); its
// derivatives are taken analytically from the closed forms of k_tail.hip,
// written as  eps * (s6 X6(R) + s12 X12(R))  per species pair.
//
// Accumulation is deterministic: per-pair values go into per-wave int64
// fixed-point LDS sums (scale 2^24), per-atom tail weights into int64 global
// sums (scale 2^32), and all FP64 reductions run in a fixed order.
#include "sre_internal.h"

namespace sre {

constexpr double kGradScale = 16777216.0;  // 2^24
constexpr double kGradInv = 1.0 / 16777216.0;
constexpr int kAdjBlock = 256;

__device__ __forceinline__ long long grad_fix(double v) {
  const double t = __builtin_fma(v, kGradScale, kMagic);
  return (long long)(__double_as_longlong(t) - __double_as_longlong(kMagic));
}

template <int T, bool FP64>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint_pairs(AdjointArgs a) {
  extern __shared__ unsigned long long acc[];  // [waves_per_block][2*npt]
  const int npt = a.n_types * a.n_types;
  const int nv = 2 * npt;
  const int wib = threadIdx.x / kWave;
  unsigned long long *wacc = acc + (size_t)wib * nv;
  for (int k = threadIdx.x % kWave; k < nv; k += kWave) wacc[k] = 0ull;
  __syncthreads();
  constexpr int A = kWave / T;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + lane / T;
  if (s < a.n) {
    const uint4 pi = a.pos4[s];
    const int rep = (int)(pi.w >> 8);
    const RepInfo &R = a.reps[rep];
    const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
    const double qdx = R.qd[0], qdy = R.qd[1], qdz = R.qd[2];
    const float rc = a.rcut[s];
    const float rc2 = rc * rc;
    const int nn = a.nnb[s];
    const int ti = (int)(pi.w & 0xffu);
    const double wE = 0.5 * a.dLdE[rep];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (a.dLdF) {
      const size_t i3 = (size_t)a.perm[s] * 3;
      gx = a.dLdF[i3];
      gy = a.dLdF[i3 + 1];
      gz = a.dLdF[i3 + 2];
    }
    const int *lp = a.list + (size_t)wave * (size_t)a.cap * A + lane;
    const int iters = (nn + T - 1) / T;
    for (int m = 0; m < iters; ++m) {
      const int k = m * T + l;
      if (k >= nn) break;
      const int j = lp[(size_t)m * kWave];
      const uint4 pj = a.pos4[j];
      const float dxf = disp(pj.x, pi.x, qx), dyf = disp(pj.y, pi.y, qy),
                  dzf = disp(pj.z, pi.z, qz);
      if (!(dist2(dxf, dyf, dzf) < rc2)) continue;
      const int p = ti * a.n_types + (int)(pj.w & 0xffu);
      const PairParam &P = a.params[p];
      double de, ds;
      if (FP64) {
        const double dx = (double)(int32_t)(pj.x - pi.x) * qdx;
        const double dy = (double)(int32_t)(pj.y - pi.y) * qdy;
        const double dz = (double)(int32_t)(pj.z - pi.z) * qdz;
        const double ri2 = 1.0 / (dx * dx + dy * dy + dz * dz);
        const double sg = (double)P.sig, ep = (double)P.eps;
        const double sg2 = sg * sg * ri2;
        const double s6 = sg2 * sg2 * sg2, s12 = s6 * s6;
        const double gd = gx * dx + gy * dy + gz * dz;
        de = wE * 4.0 * (s12 - s6) - 24.0 * (2.0 * s12 - s6) * ri2 * gd;
        ds = (wE * 24.0 * ep * (2.0 * s12 - s6) -
              144.0 * ep * (4.0 * s12 - s6) * ri2 * gd) / sg;
      } else {
        const float ri2 = __builtin_amdgcn_rcpf(dist2(dxf, dyf, dzf));
        const float sg2 = P.sig * P.sig * ri2;
        const float s6 = sg2 * sg2 * sg2, s12 = s6 * s6;
        const float gd = (float)gx * dxf + (float)gy * dyf + (float)gz * dzf;
        de = wE * (double)(4.f * (s12 - s6)) +
             (double)(-24.f * (2.f * s12 - s6) * ri2 * gd);
        ds = (wE * (double)(24.f * P.eps * (2.f * s12 - s6)) +
              (double)(-144.f * P.eps * (4.f * s12 - s6) * ri2 * gd)) /
             (double)P.sig;
      }
      atomicAdd(&wacc[p], (unsigned long long)grad_fix(de));
      atomicAdd(&wacc[npt + p], (unsigned long long)grad_fix(ds));
    }
    // tail weights of this atom (one lane per atom)
    if (a.use_tail && l == 0) {
      const int te = a.tail_idx[s];
      if (te >= 0) {
        atomicAdd((unsigned long long *)&a.tail_w[2 * (size_t)te],
                  (unsigned long long)to_fix(gz));
        atomicAdd((unsigned long long *)&a.tail_w[2 * (size_t)te + 1],
                  (unsigned long long)to_fix(a.dLdE[rep]));
      }
    }
  }
  __syncthreads();
  // every launched wave writes its slot (zeros beyond the atoms); value-major
  // so that the reduction below reads each value's row contiguously
  const size_t n_waves = (size_t)gridDim.x * (blockDim.x / kWave);
  for (int k = threadIdx.x % kWave; k < nv; k += kWave)
    a.wave_part[(size_t)k * n_waves + wave] = (double)(long long)wacc[k] * kGradInv;
}

// Fixed-order sum over waves: one block per output value, coalesced reads
// of the value's row.
__global__ __launch_bounds__(256) void k_adjoint_reduce(int n_waves, int n_vals,
                                                        const double *wave_part,
                                                        double *out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  double s = 0.0;
  const double *row = wave_part + (size_t)v * n_waves;
  for (int w = threadIdx.x; w < n_waves; w += blockDim.x) s += row[w];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0];
}

// --------------------------------------------------------------- tail part --

// Geometric integrals of one species profile: F = sum over bins/images of
// rho*(G(s1)-G(s0)) split into the sigma^6 and sigma^12 parts, likewise the
// energy integral of P, and their derivatives with respect to R.
constexpr int kTailAdjBlock = 128;  // threads per tail bin in the adjoint

struct TailX {
  double F6, F12, E6, E12, dF6, dF12, dE6, dE12;
};

__device__ inline void g_parts(double s, double R, double *g6, double *g12,
                               double *dg6, double *dg12) {
  const double a = fabs(s);
  const double R2 = R * R, R4 = R2 * R2, R6 = R4 * R2;
  if (a <= R) {
    const double s2 = s * s;
    *g6 = -2.0 * s2 / R6;
    *g12 = 2.0 * s2 / (R6 * R6);
    *dg6 = 12.0 * s2 / (R6 * R);
    *dg12 = -24.0 * s2 / (R6 * R6 * R);
  } else {
    const double a2 = a * a, a4 = a2 * a2;
    *g6 = 1.0 / a4 - 3.0 / R4;
    *g12 = -0.4 / (a4 * a4 * a2) + 2.4 / (R4 * R6);
    *dg6 = 12.0 / (R4 * R);
    *dg12 = -24.0 / (R6 * R4 * R);
  }
}

// V(a) with P(s) = sign(s) * (-2 pi) * V(|s|)
__device__ inline void v_parts(double a, double R, double *v6, double *v12,
                               double *dv6, double *dv12) {
  const double R2 = R * R, R3 = R2 * R, R4 = R2 * R2, R9 = R4 * R4 * R;
  if (a <= R) {
    *v6 = a / R4;
    *v12 = -0.4 * a / (R9 * R);
    *dv6 = -4.0 * a / (R4 * R);
    *dv12 = 4.0 * a / (R9 * R2);
  } else {
    const double a3 = a * a * a, a9 = a3 * a3 * a3;
    *v6 = (4.0 / 3.0) / R3 - 1.0 / (3.0 * a3);
    *v12 = -(4.0 / 9.0) / R9 + (0.4 / 9.0) / a9;
    *dv6 = -4.0 / R4;
    *dv12 = 4.0 / (R9 * R);
  }
}

// Sums of the planar tail integrals and their rc derivatives over the
// (image, source bin) pairs idx = t0, t0 + stride, ... (acc[0..7] in TailX
// order) and of the density over bins t0, t0 + stride, ... (acc[8]); a block
// splits the work over its threads and tail_x_finish adds the mean-density
// remainder outside the image window.
__device__ void tail_x_part(double z, double R, const double *rb, int nz,
                            double Lz, int t0, int stride, double acc[9]) {
  const int M = max(1, (int)ceil(30.0 / Lz));
  const double w = Lz / nz;
  const double two_pi = 6.283185307179586;
  for (int c = 0; c < 9; ++c) acc[c] = 0.0;
  const int total = (2 * M + 1) * nz;
  for (int idx = t0; idx < total; idx += stride) {
    const int img = idx / nz - M, k = idx - (idx / nz) * nz;
    const double r = rb[k];
    if (r == 0.0) continue;
    const double s0 = k * w + img * Lz - z, s1 = s0 + w;
    double a6, a12, da6, da12, b6, b12, db6, db12;
    g_parts(s1, R, &a6, &a12, &da6, &da12);
    g_parts(s0, R, &b6, &b12, &db6, &db12);
    acc[0] += r * (a6 - b6);
    acc[1] += r * (a12 - b12);
    acc[4] += r * (da6 - db6);
    acc[5] += r * (da12 - db12);
    double v6, v12, dv6, dv12, u6, u12, du6, du12;
    v_parts(fabs(s1), R, &v6, &v12, &dv6, &dv12);
    v_parts(fabs(s0), R, &u6, &u12, &du6, &du12);
    const double g1 = s1 < 0.0 ? two_pi : -two_pi;
    const double g0 = s0 < 0.0 ? two_pi : -two_pi;
    acc[2] += r * (g1 * v6 - g0 * u6);
    acc[3] += r * (g1 * v12 - g0 * u12);
    acc[6] += r * (g1 * dv6 - g0 * du6);
    acc[7] += r * (g1 * dv12 - g0 * du12);
  }
  for (int k = t0; k < nz; k += stride) acc[8] += rb[k];
}

__device__ TailX tail_x_finish(double z, double R, int nz, double Lz,
                               const double acc[9]) {
  const int M = max(1, (int)ceil(30.0 / Lz));
  const double two_pi = 6.283185307179586;
  TailX x{acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], acc[6], acc[7]};
  const double rbar = acc[8] / nz;
  const double sL = -M * Lz - z, sU = (M + 1) * Lz - z;
  double a6, a12, da6, da12, b6, b12, db6, db12;
  g_parts(sL, R, &a6, &a12, &da6, &da12);
  g_parts(sU, R, &b6, &b12, &db6, &db12);
  x.F6 += rbar * (a6 - b6);
  x.F12 += rbar * (a12 - b12);
  x.dF6 += rbar * (da6 - db6);
  x.dF12 += rbar * (da12 - db12);
  const double R2 = R * R, R3 = R2 * R, R4 = R2 * R2, R9 = R4 * R4 * R;
  const double vi6 = (4.0 / 3.0) / R3, vi12 = -(4.0 / 9.0) / R9;
  const double dvi6 = -4.0 / R4, dvi12 = 4.0 / (R9 * R);
  double v6, v12, dv6, dv12, u6, u12, du6, du12;
  v_parts(fabs(sU), R, &v6, &v12, &dv6, &dv12);
  v_parts(fabs(sL), R, &u6, &u12, &du6, &du12);
  x.E6 += rbar * (-two_pi) * (2.0 * vi6 - v6 - u6);
  x.E12 += rbar * (-two_pi) * (2.0 * vi12 - v12 - u12);
  x.dE6 += rbar * (-two_pi) * (2.0 * dvi6 - dv6 - du6);
  x.dE12 += rbar * (-two_pi) * (2.0 * dvi12 - dv12 - du12);
  return x;
}

// One block per tail bin (over all replicas), its threads splitting the
// planar integrals (fixed-order tree reduction): combines the per-entry
// weights W_F = sum g_z and W_E = sum dLdE with the derivatives of the table
// entries.  Writes d/drc of the bin to g_rc and the eps/sigma parts to
// contrib[bin][2*npt].
__global__ __launch_bounds__(kTailAdjBlock) void k_adjoint_tail(
    int n_bins, const int *entry_rep, const RepInfo *reps, int n_types,
    const PairParamD *params, const float *rc_table, const double *rho,
    const int64_t *wts, double *contrib, double *g_rc) {
  __shared__ double red[9][kTailAdjBlock];
  const int bg = blockIdx.x;
  const int t = threadIdx.x;
  const int npt = n_types * n_types;
  const int r = entry_rep[bg];
  const RepInfo &R = reps[r];
  const int bin = bg - R.tail_off;
  const int nz = R.rc_n[2];
  const double Lz = (double)R.L[2];
  const double z = (bin + 0.5) * Lz / nz;
  const double rc = (double)rc_table[R.rc_off + bin];
  double *c = contrib + (size_t)bg * 2 * npt;
  if (t == 0)
    for (int k = 0; k < 2 * npt; ++k) c[k] = 0.0;
  double grc = 0.0;
  for (int ta = 0; ta < n_types; ++ta) {
    const int te = bg * n_types + ta;
    const double WF = (double)wts[2 * (size_t)te] * kFixInv;
    const double WE = (double)wts[2 * (size_t)te + 1] * kFixInv;
    if (WF == 0.0 && WE == 0.0) continue;  // block-uniform
    for (int tb = 0; tb < n_types; ++tb) {
      double acc[9];
      tail_x_part(z, rc, rho + R.rho_off + (size_t)tb * R.rho_nz, R.rho_nz, Lz,
                  t, kTailAdjBlock, acc);
      for (int k = 0; k < 9; ++k) red[k][t] = acc[k];
      __syncthreads();
      for (int o = kTailAdjBlock / 2; o > 0; o >>= 1) {
        if (t < o)
          for (int k = 0; k < 9; ++k) red[k][t] += red[k][t + o];
        __syncthreads();
      }
      if (t == 0) {
        for (int k = 0; k < 9; ++k) acc[k] = red[k][0];
        const TailX x = tail_x_finish(z, rc, R.rho_nz, Lz, acc);
        const PairParamD &P = params[ta * n_types + tb];
        const double sg = P.sig, ep = P.eps;
        const double s2 = sg * sg, s6 = s2 * s2 * s2, s12 = s6 * s6;
        // T_F = -2 pi eps (s6 F6 + s12 F12),  T_E = eps/2 (s6 E6 + s12 E12)
        const double cF = -6.283185307179586, cE = 0.5;
        const double tF = s6 * x.F6 + s12 * x.F12, tE = s6 * x.E6 + s12 * x.E12;
        const double ds6 = 6.0 * s6 / sg, ds12 = 12.0 * s12 / sg;
        const int p = ta * n_types + tb;
        c[p] += WF * cF * tF + WE * cE * tE;
        c[npt + p] += ep * (WF * cF * (ds6 * x.F6 + ds12 * x.F12) +
                            WE * cE * (ds6 * x.E6 + ds12 * x.E12));
        grc += ep * (WF * cF * (s6 * x.dF6 + s12 * x.dF12) +
                     WE * cE * (s6 * x.dE6 + s12 * x.dE12));
      }
      __syncthreads();  // red[] reused
    }
  }
  if (t == 0) g_rc[R.rc_off + bin] = grc;
}

__global__ __launch_bounds__(256) void k_adjoint_tail_reduce(
    int n_bins, int n_vals, const double *contrib, double *out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < n_bins; b += blockDim.x)
    s += contrib[(size_t)b * n_vals + v];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0];
}

template <int T>
static void launch_pairs_t(const AdjointArgs &a, bool fp64, hipStream_t st) {
  const int A = kWave / T;
  const int waves = (a.n + A - 1) / A;
  const int grid = (waves * kWave + kAdjBlock - 1) / kAdjBlock;
  const size_t shm = (size_t)(kAdjBlock / kWave) * 2 * a.n_types * a.n_types *
                     sizeof(unsigned long long);
  if (fp64)
    hipLaunchKernelGGL((k_adjoint_pairs<T, true>), dim3(grid), dim3(kAdjBlock),
                       shm, st, a);
  else
    hipLaunchKernelGGL((k_adjoint_pairs<T, false>), dim3(grid),
                       dim3(kAdjBlock), shm, st, a);
}

int adjoint_grid_waves(int n, int lanes) {
  const int A = kWave / lanes;
  const int waves = (n + A - 1) / A;
  return ((waves * kWave + kAdjBlock - 1) / kAdjBlock) * (kAdjBlock / kWave);
}

void launch_adjoint_pairs(const AdjointArgs &a, int precision, int lanes,
                          double *out, hipStream_t st) {
  const int npt = a.n_types * a.n_types;
  switch (lanes) {
    case 1: launch_pairs_t<1>(a, precision != 0, st); break;
    case 2: launch_pairs_t<2>(a, precision != 0, st); break;
    case 4: launch_pairs_t<4>(a, precision != 0, st); break;
    case 8: launch_pairs_t<8>(a, precision != 0, st); break;
    default: launch_pairs_t<16>(a, precision != 0, st); break;
  }
  hipLaunchKernelGGL(k_adjoint_reduce, dim3(2 * npt), dim3(256), 0, st,
                     adjoint_grid_waves(a.n, lanes), 2 * npt, a.wave_part, out);
}

void launch_adjoint_tail(int n_bins, const int *entry_rep, const RepInfo *reps,
                         int n_types, const PairParamD *params,
                         const float *rc_table, const double *rho,
                         const int64_t *tail_w, double *contrib, double *g_rc,
                         double *out, hipStream_t st) {
  if (n_bins <= 0) return;
  const int npt = n_types * n_types;
  hipLaunchKernelGGL(k_adjoint_tail, dim3(n_bins), dim3(kTailAdjBlock), 0, st,
                     n_bins, entry_rep, reps, n_types, params, rc_table, rho,
                     tail_w, contrib, g_rc);
  hipLaunchKernelGGL(k_adjoint_tail_reduce, dim3(2 * npt), dim3(256), 0, st,
                     n_bins, 2 * npt, contrib, out);
}

}  // namespace sre
