// k_md.hip -- the MD step either side of the force path: velocity Verlet
// (H/src/Integrator.cu:273-347, velocityVerlet_part1/part2), the two-link
// Nose-Hoover chain thermostat (H/src/Thermostat.cu:255-276, NoseHoover_Chains2::
// operator_L) and the centre-of-mass velocity removal
// (H/src/Integrator_interface.cu:40-68, TranslationalFreedomRemover::remove),
// in the order of the reference driver (H/app/lj.nhc-xtc.cu:158-196):
//   remove COM -> NHC(dt/2) -> kick(dt/2) + drift(dt) -> force -> kick(dt/2)
//   -> NHC(dt/2).
// GPU-side formulation: one reduction of (kinetic energy, momentum) per step,
// taken in the kick kernel; the chain variables live on the device and one
// single-block kernel turns the reduced sums into the velocity factors of the
// next step (v <- a2 (a1 v - v_com)), using that scaling all velocities by a
// multiplies K by a^2 and that removing the COM velocity lowers K by
// |P|^2 / 2M.  The drift is fused into the position refresh.  No host
// synchronisation: the whole step is graph-capturable.
#include "sre_internal.h"

namespace sre {

constexpr int kMDBlock = 256;

// one NHC half step of length h on the chain variables, given the current
// kinetic energy K; returns the velocity scale factor (operator_L order)
__device__ double nhc_operator(MDState &m, double h, double K) {
  const double T = m.refT;
  m.vxi2 += (m.Q1 * m.vxi1 * m.vxi1 - T) / m.Q2 * (0.5 * h);   // G2
  m.vxi1 *= exp(-m.vxi2 * 0.25 * h);                             // vxi1
  m.vxi1 += (2.0 * K - m.LL * T) / m.Q1 * (0.5 * h);             // G1
  m.vxi1 *= exp(-m.vxi2 * 0.25 * h);                             // vxi1
  const double a = exp(-m.vxi1 * h);                             // Cv
  K *= a * a;
  m.xi1 += m.vxi1 * h;                                           // xi
  m.xi2 += m.vxi2 * h;
  m.vxi1 *= exp(-m.vxi2 * 0.25 * h);                             // vxi1
  m.vxi1 += (2.0 * K - m.LL * T) / m.Q1 * (0.5 * h);             // G1
  m.vxi1 *= exp(-m.vxi2 * 0.25 * h);                             // vxi1
  m.vxi2 += (m.Q1 * m.vxi1 * m.vxi1 - T) / m.Q2 * (0.5 * h);     // G2
  return a;
}

// Kick v += F dt/(2m) (caller order; forces read from the sorted slot) and
// save the force in caller order for the next step's first kick (a rebuild
// in between reorders the sorted arrays).  Per-block partial sums of
// (K = sum m v^2 / 2, P = sum m v) in a fixed order -> part[block][4].
__global__ __launch_bounds__(kMDBlock) void k_md_kick(
    int n, const uint32_t *__restrict__ inv_perm, const int *__restrict__ types,
    const int64_t *__restrict__ fx, const int64_t *__restrict__ fy,
    const int64_t *__restrict__ fz, const double *__restrict__ fxd,
    const double *__restrict__ fyd, const double *__restrict__ fzd,
    const float *__restrict__ mass, float hdt, float *__restrict__ vel,
    float *__restrict__ fcall, double *__restrict__ part) {
  __shared__ double red[kMDBlock / kWave][4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double k = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  if (i < n) {
    const int s = (int)inv_perm[i];
    float Fx, Fy, Fz;
    if (fx) {
      Fx = (float)((double)fx[s] * kFixInv);
      Fy = (float)((double)fy[s] * kFixInv);
      Fz = (float)((double)fz[s] * kFixInv);
    } else {
      Fx = (float)fxd[s];
      Fy = (float)fyd[s];
      Fz = (float)fzd[s];
    }
    const float m = mass[types[i]];
    const float c = hdt / m;
    float *v = vel + (size_t)i * 3;
    const float vx = v[0] + c * Fx, vy = v[1] + c * Fy, vz = v[2] + c * Fz;
    v[0] = vx;
    v[1] = vy;
    v[2] = vz;
    float *f = fcall + (size_t)i * 3;
    f[0] = Fx;
    f[1] = Fy;
    f[2] = Fz;
    k = 0.5 * (double)m * ((double)vx * vx + (double)vy * vy + (double)vz * vz);
    px = (double)m * vx;
    py = (double)m * vy;
    pz = (double)m * vz;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    k += __shfl_xor(k, o, kWave);
    px += __shfl_xor(px, o, kWave);
    py += __shfl_xor(py, o, kWave);
    pz += __shfl_xor(pz, o, kWave);
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[w][0] = k; red[w][1] = px; red[w][2] = py; red[w][3] = pz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kMDBlock / kWave; ++q) {
      k += red[q][0]; px += red[q][1]; py += red[q][2]; pz += red[q][3];
    }
    double *d = part + (size_t)blockIdx.x * 4;
    d[0] = k; d[1] = px; d[2] = py; d[3] = pz;
  }
}

// fixed-order sum of the kick partials by one block
__device__ void reduce_parts(int nparts, const double *part, double out[4]) {
  __shared__ double red[kMDBlock / kWave][4];
  double a[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
    for (int c = 0; c < 4; ++c) a[c] += part[(size_t)b * 4 + c];
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) a[c] += __shfl_xor(a[c], o, kWave);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0)
    for (int c = 0; c < 4; ++c) red[w][c] = a[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 1; q < kMDBlock / kWave; ++q)
      for (int c = 0; c < 4; ++c) a[c] += red[q][c];
  for (int c = 0; c < 4; ++c) out[c] = a[c];  // valid on thread 0
}

// Start of a step: the pending end-of-step NHC half step of the previous
// step, the COM velocity removal and this step's first NHC half step, all
// from the kinetic energy and momentum after the previous kick.
__global__ __launch_bounds__(kMDBlock) void k_md_factors(int nparts,
                                                         const double *part,
                                                         MDState *st) {
  double s[4];
  reduce_parts(nparts, part, s);
  if (threadIdx.x != 0) return;
  MDState m = *st;
  const double h = 0.5 * m.dt;
  double K = s[0], P[3] = {s[1], s[2], s[3]};
  double a1 = 1.0, a2 = 1.0;
  if (m.thermostat && m.have_end) {
    a1 = nhc_operator(m, h, K);
    K *= a1 * a1;
    for (int c = 0; c < 3; ++c) P[c] *= a1;
  }
  for (int c = 0; c < 3; ++c) m.vcom[c] = 0.0;
  if (m.remove_com) {
    for (int c = 0; c < 3; ++c) m.vcom[c] = P[c] / m.total_mass;
    K -= 0.5 * (P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) / m.total_mass;
  }
  if (m.thermostat) {
    a2 = nhc_operator(m, h, K);
    K *= a2 * a2;
  }
  m.a1 = a1;
  m.a2 = a2;
  m.have_end = 1;
  m.K = K;
  *st = m;
}

// End of a run: the pending end-of-step NHC half step; v <- a1 v follows
// (k_md_scale).  K is the kinetic energy after it.
__global__ __launch_bounds__(kMDBlock) void k_md_finish(int nparts,
                                                        const double *part,
                                                        MDState *st) {
  double s[4];
  reduce_parts(nparts, part, s);
  if (threadIdx.x != 0) return;
  MDState m = *st;
  double K = s[0], a1 = 1.0;
  if (m.thermostat && m.have_end) {
    a1 = nhc_operator(m, 0.5 * m.dt, K);
    K *= a1 * a1;
  }
  m.a1 = a1;
  m.a2 = 1.0;
  for (int c = 0; c < 3; ++c) m.vcom[c] = 0.0;
  m.have_end = 0;
  m.K = K;
  *st = m;
}

__global__ __launch_bounds__(kMDBlock) void k_md_scale(int n, const MDState *st,
                                                       float *vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  vel[i] = (float)(st->a1 * (double)vel[i]);
}

// Copies up to kCopySegs buffers (whole 4-byte words) in one launch: the
// snapshot / restore of an adaptive MD stretch.
__global__ __launch_bounds__(kMDBlock) void k_md_copy(CopySegs c) {
  size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < c.count; ++k) {
    if (w < c.words[k]) {
      c.dst[k][w] = c.src[k][w];
      return;
    }
    w -= c.words[k];
  }
}

void launch_md_copy(const CopySegs &c, hipStream_t st) {
  size_t total = 0;
  for (int k = 0; k < c.count; ++k) total += c.words[k];
  if (total == 0) return;
  hipLaunchKernelGGL(k_md_copy, dim3((unsigned)((total + kMDBlock - 1) / kMDBlock)),
                     dim3(kMDBlock), 0, st, c);
}

int md_part_count(int n) { return (n + kMDBlock - 1) / kMDBlock; }

void launch_md_kick(int n, const uint32_t *inv_perm, const int *types,
                    const int64_t *fx, const int64_t *fy, const int64_t *fz,
                    const double *fxd, const double *fyd, const double *fzd,
                    const float *mass, float hdt, float *vel, float *fcall,
                    double *part, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_md_kick, dim3(md_part_count(n)), dim3(kMDBlock), 0, st,
                     n, inv_perm, types, fx, fy, fz, fxd, fyd, fzd, mass, hdt,
                     vel, fcall, part);
}

void launch_md_factors(int n, const double *part, MDState *state,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_md_factors, dim3(1), dim3(kMDBlock), 0, st,
                     md_part_count(n), part, state);
}

void launch_md_finish(int n, const double *part, MDState *state, float *vel,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_md_finish, dim3(1), dim3(kMDBlock), 0, st,
                     md_part_count(n), part, state);
  hipLaunchKernelGGL(k_md_scale, dim3((3 * n + kMDBlock - 1) / kMDBlock),
                     dim3(kMDBlock), 0, st, n, state, vel);
}

}  // namespace sre
