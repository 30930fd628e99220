// k_tail.hip -- FP64 inhomogeneous tail correction for a planar density
// profile rho_t(z).
//
// The force-correction engine adds the mean error force of the pairs beyond
// the cutoff,  <dF>(r) = int f^c(r - r') rho(r') dr'  (PRE:256-262, 564-567),
// per particle by bin (FC/src/InteractionEngine_interface.cu:1118,
// FC/src/AssignForceCorr.cu:95-128), computing it with FFTs on a 3-D grid
// (FC/src/ForceCorr.cu:131-162).  For rho = rho(z) the integral over the
// plane has a closed form.  With s = z' - z and u~ = 4 (sig^12 r^-12 - sig^6 r^-6):
//
//   F_z(z) = -2 pi eps  int ds rho(z+s) s u~(max(R, |s|))
//   E(z)   =  1/2  eps  int ds rho(z+s) Phi(s),  Phi(s) = -2 pi H(max(R, |s|))
//
// where R = rc of the atom, H(t) = sig^6 t^-4 - 0.4 sig^12 t^-10 (H' = t u~).
// The s-integral over a piecewise-constant rho uses the antiderivatives
//   G(s) = u~(R) s^2/2                          |s| <= R
//        = H(|s|) - H(R) + u~(R) R^2/2           |s| >  R
//   P(s) = sign(s) * (-2 pi) * { H(R)|s|                      |s| <= R
//                               { H(R) R + K(|s|) - K(R)       |s| >  R
//   K(t) = -sig^6/(3 t^3) + 0.4 sig^12/(9 t^9)   (K' = H).
// Bins are summed over the periodic images n = -M..M, M = max(1, ceil(30/Lz)),
// and beyond that window the mean density is integrated analytically.
#include "sre_internal.h"

namespace sre {

struct TailLJ {
  double s6, s12;
};

__device__ __forceinline__ double tH(double t, const TailLJ &p) {
  const double t2 = t * t, t4 = t2 * t2, t10 = t4 * t4 * t2;
  return p.s6 / t4 - 0.4 * p.s12 / t10;
}
__device__ __forceinline__ double tK(double t, const TailLJ &p) {
  const double t3 = t * t * t, t9 = t3 * t3 * t3;
  return -p.s6 / (3.0 * t3) + 0.4 * p.s12 / (9.0 * t9);
}
__device__ __forceinline__ double tU(double r, const TailLJ &p) {
  const double r2 = r * r, r6 = r2 * r2 * r2;
  return 4.0 * (p.s12 / (r6 * r6) - p.s6 / r6);
}
__device__ __forceinline__ double tG(double s, double R, double uR, double HR,
                                     const TailLJ &p) {
  const double a = fabs(s);
  if (a <= R) return 0.5 * uR * s * s;
  return tH(a, p) - HR + 0.5 * uR * R * R;
}
__device__ __forceinline__ double tP(double s, double R, double HR, double KR,
                                     const TailLJ &p) {
  const double a = fabs(s);
  const double v = (a <= R) ? HR * a : HR * R + tK(a, p) - KR;
  const double two_pi = 6.283185307179586;
  return (s < 0.0 ? two_pi : -two_pi) * v;
}

// Tail force (z) and energy of one atom of type ta at height z with cutoff R.
// Work items of the (image, bin) double loop are split over nthr threads and
// the partial sums are returned for a fixed-order reduction by the caller.
__device__ void tail_partial(double z, double R, int ta, int n_types,
                             const PairParamD *params, const double *rho,
                             int nz, double Lz, int tid, int nthr, double *pf,
                             double *pe) {
  const int M = max(1, (int)ceil(30.0 / Lz));
  const double w = Lz / nz;
  double F = 0.0, E = 0.0;
  for (int tb = 0; tb < n_types; ++tb) {
    const PairParamD &P = params[ta * n_types + tb];
    const double s2 = P.sig * P.sig;
    TailLJ p;
    p.s6 = s2 * s2 * s2;
    p.s12 = p.s6 * p.s6;
    const double uR = tU(R, p), HR = tH(R, p), KR = tK(R, p);
    const double *rb = rho + (size_t)tb * nz;
    double fb = 0.0, eb = 0.0;
    const int items = (2 * M + 1) * nz;
    for (int it = tid; it < items; it += nthr) {
      const int img = it / nz - M, k = it % nz;
      const double r = rb[k];
      if (r == 0.0) continue;
      const double s0 = k * w + img * Lz - z, s1 = s0 + w;
      fb += r * (tG(s1, R, uR, HR, p) - tG(s0, R, uR, HR, p));
      eb += r * (tP(s1, R, HR, KR, p) - tP(s0, R, HR, KR, p));
    }
    if (tid == 0) {
      // beyond the image window: mean density, G even, P odd
      double rbar = 0.0;
      for (int k = 0; k < nz; ++k) rbar += rb[k];
      rbar /= nz;
      const double sL = -M * Lz - z, sU = (M + 1) * Lz - z;
      const double Pinf = -6.283185307179586 * (HR * R - KR);
      fb += rbar * (tG(sL, R, uR, HR, p) - tG(sU, R, uR, HR, p));
      eb += rbar * (2.0 * Pinf - tP(sU, R, HR, KR, p) + tP(sL, R, HR, KR, p));
    }
    F += -6.283185307179586 * P.eps * fb;
    E += 0.5 * P.eps * eb;
  }
  *pf = F;
  *pe = E;
}

// one block per (replica bin, type): value at the bin centre with the bin's rc
__global__ __launch_bounds__(256) void k_tail_table(
    const RepInfo *reps, const int *entry_rep, int n_types,
    const PairParamD *params, const float *rc_table, const double *rho,
    double2 *tail) {
  __shared__ double sf[256], se[256];
  const int e = blockIdx.x;           // (tail_off + bin) * n_types + type
  const int ta = e % n_types;
  const int bin_g = e / n_types;
  const int r = entry_rep[bin_g];
  const RepInfo &R = reps[r];
  const int bin = bin_g - R.tail_off;
  const int nz = R.rc_n[2];
  const double Lz = (double)R.L[2];
  const double z = (bin + 0.5) * Lz / nz;
  const double rc = (double)rc_table[R.rc_off + bin];
  double f, en;
  tail_partial(z, rc, ta, n_types, params, rho + R.rho_off, R.rho_nz, Lz,
               threadIdx.x, blockDim.x, &f, &en);
  sf[threadIdx.x] = f;
  se[threadIdx.x] = en;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sf[threadIdx.x] += sf[threadIdx.x + o];
      se[threadIdx.x] += se[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) tail[e] = make_double2(sf[0], se[0]);
}

// entry_rep maps each tail bin (over all replicas) to its replica.
void launch_tail_table(int n_entries, const RepInfo *reps,
                        const int *entry_rep, int n_types,
                        const PairParamD *params, const float *rc_table,
                        const double *rho, double2 *tail, hipStream_t st) {
  if (n_entries <= 0) return;
  hipLaunchKernelGGL(k_tail_table, dim3(n_entries), dim3(256), 0, st, reps,
                     entry_rep, n_types, params, rc_table, rho, tail);
}

// per atom at its own height (one wave per atom)
__global__ __launch_bounds__(256) void k_tail_exact(
    int n, int n_types, const uint4 *pos4, const float *rcut,
    const RepInfo *reps, const PairParamD *params, const double *rho,
    double2 *out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int s = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (s >= n) return;
  const uint4 pi = pos4[s];
  const RepInfo &R = reps[pi.w >> 8];
  double f = 0.0, en = 0.0;
  if (R.rho_off >= 0) {
    const double z = (double)pi.z * R.qd[2];
    tail_partial(z, (double)rcut[s], (int)(pi.w & 0xffu), n_types, params,
                 rho + R.rho_off, R.rho_nz, (double)R.L[2], lane, kWave, &f,
                 &en);
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    f += __shfl_xor(f, o, kWave);
    en += __shfl_xor(en, o, kWave);
  }
  if (lane == 0) out[s] = make_double2(f, en);
}

void launch_tail_exact(int n, int n_types, const uint4 *pos4,
                        const float *rcut, const RepInfo *reps,
                        const PairParamD *params, const double *rho,
                        double2 *out, hipStream_t st) {
  if (n <= 0) return;
  const int block = 256;
  const int grid = (int)(((size_t)n * kWave + block - 1) / block);
  hipLaunchKernelGGL(k_tail_exact, dim3(grid), dim3(block), 0, st, n, n_types,
                     pos4, rcut, reps, params, rho, out);
}

}  // namespace sre
