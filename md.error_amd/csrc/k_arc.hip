// k_arc.hip -- the adaptive-cutoff producer on the GPU: density deposit on a
// grid, the error kernels of every rc candidate, the FFT-route error estimate,
// the per-cell choice of rc and its 3x3x3 widening.  Restates (H =
// shortRange.inhomo/template/tools/simulation.hetero.cutoff):
//   DensityProfile_PiecewiseConst::{reinit, deposite, calculate}
//       (H/src/DensityProfile.cu:27-99)
//   AdaptRCut::{reinit, makeS1k, makeS2k, calError, calRCut, refineRCut}
//       (H/src/AdaptRCut.cu:96-163, 265-416, 182-236, 418-468, 470-508)
//   PressureCorrection::{makeKernel, correction}
//       (H/src/PressureCorrection.cu:262-355, 176-211)
//
// GPU formulation:
//  * kernel tables: one thread per grid cell integrates x^-13 sin(2 pi k x)
//    and x^-5 sin(2 pi k x) from 30 down to rc_min in panels (10-point
//    Gauss-Legendre, panel width <= 1/8 and <= 1/(8k)), recording the running
//    integral at every candidate, so all candidates cost one pass;
//  * transforms: the grid edges are odd (the reference makes them so), so a
//    3-D DFT is three passes of direct per-axis DFTs (one thread per output
//    element, twiddles from an exact table), O(nele * (nx + ny + nz));
//  * the four inverse transforms per candidate are packed into two complex
//    ones (every output is real: Z1 = (S1 - kx G) rho_k -> e1 + i e2x,
//    Z2 = (i ky - kz) G rho_k -> e2y + i e2z), all candidates in one batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sre_host.h"
#include "sre_internal.h"

namespace sre {

constexpr int kArcBlock = 256;
constexpr int kGL = 10;
constexpr double kArcXMax = 30.0;             // integration limit of makeS1k/S2k
constexpr double kArcP1 = 16.0 * 36.0;         // eps = sigma = 1
constexpr double kArcP2 = 8.0;

struct ArcQuad {
  double x[kGL], w[kGL];   // Gauss-Legendre on [-1, 1]
};

struct ArcGrid {
  int n[3];
  int nele;
  double L[3];
};

__device__ inline void arc_cell(const ArcGrid &g, int cell, int i[3]) {
  i[2] = cell % g.n[2];
  const int t = cell / g.n[2];
  i[1] = t % g.n[1];
  i[0] = t / g.n[1];
}

// signed frequency / L at FFT index positions (odd n: 0..(n-1)/2, then negative)
__device__ inline void arc_k(const ArcGrid &g, const int i[3], double k[3]) {
  for (int a = 0; a < 3; ++a) {
    const int f = i[a] <= (g.n[a] - 1) / 2 ? i[a] : i[a] - g.n[a];
    k[a] = (double)f / g.L[a];
  }
}

// sqrt(2/(pi z)) sin z and sqrt(2z/pi) j2(z): J_{1/2}(z), J_{5/2}(z)
// (half-integer Bessel functions in closed form; a series for j2 at small z)
__device__ inline void bessel_half(double z, double sz, double cz, double *j1,
                                   double *j5) {
  const double r = sqrt(2.0 / (M_PI * z));
  *j1 = r * sz;
  double j2;
  if (z < 0.5) {
    const double z2 = z * z;
    j2 = z2 / 15.0 *
         (1.0 - z2 / 14.0 * (1.0 - z2 / 36.0 * (1.0 - z2 / 66.0 * (1.0 - z2 / 104.0))));
  } else {
    const double iz = 1.0 / z;
    j2 = (3.0 * iz * iz * iz - iz) * sz - 3.0 * cz * iz * iz;
  }
  *j5 = j2 * z * r;   // sqrt(2z/pi) j2 = z sqrt(2/(pi z)) j2
}

// makeS1k / makeS2k for every candidate, layout [c][cell]; with I1 != NULL
// also the integrals i1, i5 = int_rc^30 -24 x^-4.5 J_{1/2, 5/2}(2 pi k x) dx
// of PressureCorrection::makeKernel
__global__ __launch_bounds__(kArcBlock) void k_arc_kernels(
    ArcGrid g, int nrc, const double *__restrict__ rcs, ArcQuad q,
    double *__restrict__ S1, double *__restrict__ G, double *__restrict__ I1,
    double *__restrict__ I5) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.nele) return;
  int i[3];
  double kv[3];
  arc_cell(g, cell, i);
  arc_k(g, i, kv);
  const double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  if (k == 0.0) {
    for (int c = 0; c < nrc; ++c) {
      const double rc = rcs[c];
      const double r2 = rc * rc, r4 = r2 * r2, r8 = r4 * r4;
      S1[(size_t)c * g.nele + cell] = 4.0 * M_PI * kArcP1 / 11.0 / (r8 * r2 * rc);
      G[(size_t)c * g.nele + cell] = 0.0;
      I1[(size_t)c * g.nele + cell] = 0.0;
      I5[(size_t)c * g.nele + cell] = 0.0;
    }
    return;
  }
  const double hmax = fmin(0.125, 0.125 / k);
  double acc13 = 0.0, acc5 = 0.0, accb1 = 0.0, accb5 = 0.0, upper = kArcXMax;
  for (int c = nrc - 1; c >= 0; --c) {
    const double lo = rcs[c];
    const int m = max(1, (int)ceil((upper - lo) / hmax));
    const double hp = (upper - lo) / m;
    for (int p = 0; p < m; ++p) {
      const double mid = lo + (p + 0.5) * hp;
      double s13 = 0.0, s5 = 0.0, b1 = 0.0, b5 = 0.0;
#pragma unroll
      for (int j = 0; j < kGL; ++j) {
        const double x = mid + 0.5 * hp * q.x[j];
        double s, cs;
        sincospi(2.0 * k * x, &s, &cs);
        const double inv = 1.0 / x;
        const double i2 = inv * inv, i4 = i2 * i2;
        const double i5 = i4 * inv;
        s13 += q.w[j] * (i5 * i4 * i4) * s;
        s5 += q.w[j] * i5 * s;
        double j1, j5;
        bessel_half(2.0 * M_PI * k * x, s, cs, &j1, &j5);
        const double f = -24.0 * i4 * sqrt(inv);
        b1 += q.w[j] * f * j1;
        b5 += q.w[j] * f * j5;
      }
      acc13 += 0.5 * hp * s13;
      acc5 += 0.5 * hp * s5;
      accb1 += 0.5 * hp * b1;
      accb5 += 0.5 * hp * b5;
    }
    upper = lo;
    const double rc = lo;
    const double phi = 2.0 * M_PI * k * rc;
    const double rc6 = rc * rc * rc * rc * rc * rc;
    const double size = -2.0 * M_PI * rc * rc * (4.0 / rc6) *
                        (2.0 * cos(phi) / phi - 2.0 * sin(phi) / (phi * phi)) / k;
    S1[(size_t)c * g.nele + cell] = 2.0 / k * kArcP1 * acc13;
    G[(size_t)c * g.nele + cell] = 2.0 * M_PI * kArcP2 / k * acc5 + size;
    I1[(size_t)c * g.nele + cell] = accb1;
    I5[(size_t)c * g.nele + cell] = accb5;
  }
}

// K_aa(k) = (pi / sqrt(k)) (2/3 i1 + (2/3 - 2 khat_a^2) i5), K(0) = 0
__device__ inline void pressure_kernel(const double kv[3], double i1, double i5,
                                       double K[3]) {
  const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  if (k2 == 0.0) {
    K[0] = K[1] = K[2] = 0.0;
    return;
  }
  const double k = sqrt(k2);
  const double sc = M_PI / sqrt(k);
  for (int a = 0; a < 3; ++a)
    K[a] = sc * (2.0 / 3.0 * i1 + (2.0 / 3.0 - 2.0 * kv[a] * kv[a] / k2) * i5);
}

// Z[2c] = (Kxx + i Kyy) rho_k, Z[2c+1] = Kzz rho_k   (rho_k scaled)
__global__ __launch_bounds__(kArcBlock) void k_arc_pack_pressure(
    ArcGrid g, int nrc, double scale, const double *__restrict__ I1,
    const double *__restrict__ I5, const double2 *__restrict__ rhok,
    double2 *__restrict__ Z) {
  const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (size_t)nrc * g.nele) return;
  const int c = (int)(gi / g.nele);
  const int cell = (int)(gi % g.nele);
  int i[3];
  double kv[3], K[3];
  arc_cell(g, cell, i);
  arc_k(g, i, kv);
  pressure_kernel(kv, I1[gi], I5[gi], K);
  const double2 r = make_double2(rhok[cell].x * scale, rhok[cell].y * scale);
  Z[(size_t)(2 * c) * g.nele + cell] =
      make_double2(K[0] * r.x - K[1] * r.y, K[0] * r.y + K[1] * r.x);
  Z[(size_t)(2 * c + 1) * g.nele + cell] = make_double2(K[2] * r.x, K[2] * r.y);
}

// per-block partial sums of 1/2 conv_aa[idx](cell) rho(cell), fixed order
__global__ __launch_bounds__(kArcBlock) void k_arc_pressure_sum(
    ArcGrid g, const int *__restrict__ idx, const double *__restrict__ rho,
    const double2 *__restrict__ Z, double *__restrict__ part) {
  __shared__ double red[kArcBlock / kWave][3];
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (cell < g.nele) {
    const int c = idx[cell];
    const double2 z1 = Z[(size_t)(2 * c) * g.nele + cell];
    const double2 z2 = Z[(size_t)(2 * c + 1) * g.nele + cell];
    const double w = 0.5 * rho[cell];
    v[0] = w * z1.x;
    v[1] = w * z1.y;
    v[2] = w * z2.x;
  }
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v[a] += __shfl_xor(v[a], o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0)
    for (int a = 0; a < 3; ++a) red[threadIdx.x / kWave][a] = v[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kArcBlock / kWave; ++q)
      for (int a = 0; a < 3; ++a) v[a] += red[q][a];
    for (int a = 0; a < 3; ++a) part[(size_t)blockIdx.x * 3 + a] = v[a];
  }
}

// twiddles tw[m] = exp(sign 2 pi i m / n)
__global__ void k_arc_twiddle(int n, int sign, double2 *tw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[m] = make_double2(c, sign * s);
}

// one axis of a batched 3-D DFT, out-of-place
__global__ __launch_bounds__(kArcBlock) void k_arc_dft(
    ArcGrid g, int batch, int axis, const double2 *__restrict__ tw,
    const double2 *__restrict__ in, double2 *__restrict__ out) {
  const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (size_t)batch * g.nele) return;
  const int cell = (int)(gi % g.nele);
  int i[3];
  arc_cell(g, cell, i);
  const int n = g.n[axis];
  const int stride = axis == 0 ? g.n[1] * g.n[2] : (axis == 1 ? g.n[2] : 1);
  const int kk = i[axis];
  const double2 *line = in + (gi - (size_t)kk * stride);
  double re = 0.0, im = 0.0;
  int m = 0;
  for (int j = 0; j < n; ++j) {
    const double2 v = line[(size_t)j * stride];
    const double2 w = tw[m];
    re += v.x * w.x - v.y * w.y;
    im += v.x * w.y + v.y * w.x;
    m += kk;
    if (m >= n) m -= n;
  }
  out[gi] = make_double2(re, im);
}

__global__ void k_arc_real_to_complex(int nele, const double *rho, double2 *z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nele) z[i] = make_double2(rho[i], 0.0);
}

// Z[2c] = (S1 - kx G) rho_k, Z[2c+1] = (i ky - kz) G rho_k   (rho_k scaled)
__global__ __launch_bounds__(kArcBlock) void k_arc_pack(
    ArcGrid g, int nrc, double scale, const double *__restrict__ S1,
    const double *__restrict__ G, const double2 *__restrict__ rhok,
    double2 *__restrict__ Z) {
  const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (size_t)nrc * g.nele) return;
  const int c = (int)(gi / g.nele);
  const int cell = (int)(gi % g.nele);
  int i[3];
  double kv[3];
  arc_cell(g, cell, i);
  arc_k(g, i, kv);
  const double2 r = make_double2(rhok[cell].x * scale, rhok[cell].y * scale);
  const double s1 = S1[gi], gg = G[gi];
  const double a = s1 - kv[0] * gg;
  Z[(size_t)(2 * c) * g.nele + cell] = make_double2(a * r.x, a * r.y);
  // (-kz G + i ky G) * r
  const double br = -kv[2] * gg, bi = kv[1] * gg;
  Z[(size_t)(2 * c + 1) * g.nele + cell] =
      make_double2(br * r.x - bi * r.y, br * r.y + bi * r.x);
}

// error[c][cell] = sqrt(e1 + e2x^2 + e2y^2 + e2z^2), every part divided by V
__global__ __launch_bounds__(kArcBlock) void k_arc_combine(
    ArcGrid g, int nrc, double inv_vol, const double2 *__restrict__ Z,
    double *__restrict__ err, double *__restrict__ parts) {
  const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= (size_t)nrc * g.nele) return;
  const int c = (int)(gi / g.nele);
  const int cell = (int)(gi % g.nele);
  const double2 z1 = Z[(size_t)(2 * c) * g.nele + cell];
  const double2 z2 = Z[(size_t)(2 * c + 1) * g.nele + cell];
  const double e1 = z1.x * inv_vol, ex = z1.y * inv_vol;
  const double ey = z2.x * inv_vol, ez = z2.y * inv_vol;
  err[gi] = sqrt(e1 + ex * ex + ey * ey + ez * ez);
  if (parts) {
    double *p = parts + gi * 4;
    p[0] = e1;
    p[1] = ex;
    p[2] = ey;
    p[3] = ez;
  }
}

// calRCutOnePoint: end checks, then bisection over the candidates
__global__ __launch_bounds__(kArcBlock) void k_arc_bisect(
    int nele, int nrc, double prec, const double *__restrict__ err,
    int *__restrict__ idx) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nele) return;
  int a = 0, b = nrc - 1;
  const double ea = err[cell], eb = err[(size_t)b * nele + cell];
  int r;
  if (ea - prec <= 0.0) {
    r = a;
  } else if (eb - prec >= 0.0) {
    r = b;
  } else {
    while (b - a > 1) {
      const int c = (a + b) / 2;
      if (err[(size_t)c * nele + cell] > prec) a = c;
      else b = c;
    }
    r = b;
  }
  idx[cell] = r;
}

// refineRCut: max of the candidate index over the periodic 3x3x3 neighbourhood
__global__ __launch_bounds__(kArcBlock) void k_arc_refine(
    ArcGrid g, const int *__restrict__ in, int *__restrict__ out) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.nele) return;
  int i[3];
  arc_cell(g, cell, i);
  int best = 0;
  for (int dx = -1; dx <= 1; ++dx) {
    int jx = i[0] + dx;
    jx = jx < 0 ? jx + g.n[0] : (jx >= g.n[0] ? jx - g.n[0] : jx);
    for (int dy = -1; dy <= 1; ++dy) {
      int jy = i[1] + dy;
      jy = jy < 0 ? jy + g.n[1] : (jy >= g.n[1] ? jy - g.n[1] : jy);
      for (int dz = -1; dz <= 1; ++dz) {
        int jz = i[2] + dz;
        jz = jz < 0 ? jz + g.n[2] : (jz >= g.n[2] ? jz - g.n[2] : jz);
        best = max(best, in[jz + g.n[2] * (jy + g.n[1] * jx)]);
      }
    }
  }
  out[cell] = best;
}

__global__ __launch_bounds__(kArcBlock) void k_arc_result(
    int nele, const int *__restrict__ idx, const double *__restrict__ rcs,
    const double *__restrict__ err, float *__restrict__ rc,
    double *__restrict__ res_err) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nele) return;
  const int c = idx[cell];
  rc[cell] = (float)rcs[c];
  res_err[cell] = err[(size_t)c * nele + cell];
}

// deposite: one count per atom in its cell (one-image wrap, index clamped)
__global__ __launch_bounds__(kArcBlock) void k_arc_deposit(
    ArcGrid g, int n, const float *__restrict__ pos, int stride,
    unsigned *__restrict__ counts) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  int ix[3];
  for (int d = 0; d < 3; ++d) {
    double t = (double)pos[(size_t)a * stride + d];
    if (t >= g.L[d]) t -= g.L[d];
    else if (t < 0.0) t += g.L[d];
    const double h = g.L[d] / g.n[d];
    int v = (int)(t / h);
    ix[d] = min(max(v, 0), g.n[d] - 1);
  }
  atomicAdd(&counts[ix[2] + g.n[2] * (ix[1] + g.n[1] * ix[0])], 1u);
}

// the mean error force (e2x, e2y, e2z) of each cell at its chosen candidate
__global__ __launch_bounds__(kArcBlock) void k_arc_force_grid(
    int nele, const int *__restrict__ idx, const double *__restrict__ parts,
    float *__restrict__ out) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nele) return;
  const double *p = parts + ((size_t)idx[cell] * nele + cell) * 4;
  out[3 * (size_t)cell] = (float)p[1];
  out[3 * (size_t)cell + 1] = (float)p[2];
  out[3 * (size_t)cell + 2] = (float)p[3];
}

__global__ void k_arc_density(int nele, double norm, const unsigned *counts,
                              double *rho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nele) rho[i] = (double)counts[i] / norm;
}

static unsigned arc_blocks(size_t n) {
  return (unsigned)((n + kArcBlock - 1) / kArcBlock);
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
static ArcQuad gauss_legendre() {
  ArcQuad q{};
  const int n = kGL;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double x = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0);
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    q.x[i] = -x;
    q.w[i] = w;
    q.x[n - 1 - i] = x;
    q.w[n - 1 - i] = w;
  }
  return q;
}

}  // namespace sre

using namespace sre;

struct sre_arc {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t engine_done = nullptr;         // engine work before a deposit from its
  hipEvent_t deposit_done = nullptr;        // positions, and that deposit before later
                                            // engine work
  ArcGrid g{};
  std::vector<double> rcs;
  double *d_rcs = nullptr;
  double *d_S1 = nullptr, *d_G = nullptr;   // [nrc][nele]
  double *d_I1 = nullptr, *d_I5 = nullptr;  // pressure-kernel integrals [nrc][nele]
  double *d_psum = nullptr;                 // pressure partial sums, 3 per block
  double2 *d_tw[3][2] = {};                 // per axis: forward, backward
  unsigned *d_counts = nullptr;
  int ndata = 0;
  double *d_rho = nullptr;
  bool rho_valid = false;
  double2 *d_rhok = nullptr, *d_zr = nullptr, *d_za = nullptr, *d_zb = nullptr;
  double *d_err = nullptr;                  // [nrc][nele]
  double *d_parts = nullptr;                // [nrc][nele][4], on demand
  int *d_idx = nullptr, *d_idx2 = nullptr;
  float *d_rc = nullptr;
  double *d_res_err = nullptr;
  float *d_fgrid = nullptr;                 // [nele][3], on demand
  bool parts_valid = false;
  bool err_valid = false, rc_valid = false;
  ~sre_arc();
};

sre_arc::~sre_arc() {
  dfree(d_rcs); dfree(d_S1); dfree(d_G); dfree(d_I1); dfree(d_I5); dfree(d_psum);
  for (auto &a : d_tw) { dfree(a[0]); dfree(a[1]); }
  dfree(d_counts); dfree(d_rho); dfree(d_rhok); dfree(d_zr); dfree(d_za);
  dfree(d_zb); dfree(d_err); dfree(d_parts); dfree(d_idx); dfree(d_idx2);
  dfree(d_rc); dfree(d_res_err); dfree(d_fgrid);
  if (engine_done) (void)hipEventDestroy(engine_done);
  if (deposit_done) (void)hipEventDestroy(deposit_done);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// batched 3-D DFT of za (batch grids) in place (zb scratch of the same size)
static sre_status arc_dft3(sre_arc *a, int batch, int dir, double2 *za, double2 *zb) {
  const ArcGrid &g = a->g;
  const size_t n = (size_t)batch * g.nele;
  hipLaunchKernelGGL(k_arc_dft, dim3(arc_blocks(n)), dim3(kArcBlock), 0, a->stream,
                     g, batch, 0, a->d_tw[0][dir], za, zb);
  hipLaunchKernelGGL(k_arc_dft, dim3(arc_blocks(n)), dim3(kArcBlock), 0, a->stream,
                     g, batch, 1, a->d_tw[1][dir], zb, za);
  hipLaunchKernelGGL(k_arc_dft, dim3(arc_blocks(n)), dim3(kArcBlock), 0, a->stream,
                     g, batch, 2, a->d_tw[2][dir], za, zb);
  HIP_TRY(hipMemcpyAsync(za, zb, sizeof(double2) * n, hipMemcpyDeviceToDevice,
                         a->stream));
  HIP_TRY(hipGetLastError());
  return SRE_OK;
}

extern "C" sre_status sre_arc_create(sre_arc **out, const double box[3],
                                     double refh, int device, void *stream) {
  if (!out || !box || !(refh > 0.0))
    return fail(SRE_ERR_INVALID, "sre_arc_create: bad arguments");
  *out = nullptr;
  auto *a = new sre_arc();
  // DensityProfile_PiecewiseConst::reinit: n = unsigned(L / refh), made odd
  size_t nele = 1;
  for (int d = 0; d < 3; ++d) {
    const double L = (double)(float)box[d];
    if (!(L > 0.0)) {
      delete a;
      return fail(SRE_ERR_INVALID, "sre_arc_create: box edge %d is %g", d, L);
    }
    int n = (int)(L / refh);
    if (n < 1) n = 1;
    if (n % 2 == 0) ++n;
    a->g.n[d] = n;
    a->g.L[d] = L;
    nele *= (size_t)n;
  }
  if (nele > (size_t)1 << 26) {
    delete a;
    return fail(SRE_ERR_INVALID, "sre_arc_create: grid of %zu cells too large", nele);
  }
  a->g.nele = (int)nele;
  sre_status s = [&]() -> sre_status {
    if (device >= 0) {
      a->device = device;
      HIP_TRY(hipSetDevice(device));
    } else {
      HIP_TRY(hipGetDevice(&a->device));
    }
    if (stream) {
      a->stream = (hipStream_t)stream;
    } else {
      HIP_TRY(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
      a->own_stream = true;
    }
    HIP_TRY(hipEventCreateWithFlags(&a->engine_done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&a->deposit_done, hipEventDisableTiming));
    HIP_TRY(dalloc(&a->d_counts, nele));
    HIP_TRY(dalloc(&a->d_rho, nele));
    HIP_TRY(dalloc(&a->d_rhok, nele));
    HIP_TRY(dalloc(&a->d_zr, nele));
    HIP_TRY(dalloc(&a->d_idx, nele));
    HIP_TRY(dalloc(&a->d_idx2, nele));
    HIP_TRY(dalloc(&a->d_rc, nele));
    HIP_TRY(dalloc(&a->d_res_err, nele));
    HIP_TRY(dalloc(&a->d_psum, (size_t)3 * arc_blocks(nele)));
    for (int d = 0; d < 3; ++d)
      for (int dir = 0; dir < 2; ++dir) {
        HIP_TRY(dalloc(&a->d_tw[d][dir], a->g.n[d]));
        hipLaunchKernelGGL(k_arc_twiddle, dim3(arc_blocks(a->g.n[d])), dim3(kArcBlock),
                           0, a->stream, a->g.n[d], dir == 0 ? -1 : 1, a->d_tw[d][dir]);
      }
    HIP_TRY(hipMemsetAsync(a->d_counts, 0, sizeof(unsigned) * nele, a->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(a->stream));
    return SRE_OK;
  }();
  if (s != SRE_OK) {
    delete a;
    return s;
  }
  *out = a;
  return SRE_OK;
}

// AdaptRCut::reinit: candidate list rc = rcmin, rcmin + rcstep, ... <= rcmax
// (accumulated as the reference does) and the kernel tables of every candidate
extern "C" sre_status sre_arc_set_candidates(sre_arc *a, double rcmin, double rcmax,
                                             double rcstep) {
  if (!a || !(rcmin > 0.0) || !(rcstep > 0.0) || !(rcmax >= rcmin) ||
      !(rcmax < kArcXMax))
    return fail(SRE_ERR_INVALID, "sre_arc_set_candidates: bad arguments");
  std::vector<double> rcs;
  for (double t = rcmin; t <= rcmax; t += rcstep) rcs.push_back(t);
  if (rcs.size() > 4096)
    return fail(SRE_ERR_INVALID, "sre_arc_set_candidates: %zu candidates", rcs.size());
  HIP_TRY(hipSetDevice(a->device));
  HIP_TRY(hipStreamSynchronize(a->stream));
  dfree(a->d_rcs); dfree(a->d_S1); dfree(a->d_G); dfree(a->d_za); dfree(a->d_zb);
  dfree(a->d_I1); dfree(a->d_I5);
  dfree(a->d_err); dfree(a->d_parts);
  a->rcs = rcs;
  a->err_valid = a->rc_valid = false;
  const int nrc = (int)rcs.size();
  const size_t tab = (size_t)nrc * a->g.nele;
  HIP_TRY(dalloc(&a->d_rcs, nrc));
  HIP_TRY(dalloc(&a->d_S1, tab));
  HIP_TRY(dalloc(&a->d_G, tab));
  HIP_TRY(dalloc(&a->d_I1, tab));
  HIP_TRY(dalloc(&a->d_I5, tab));
  HIP_TRY(dalloc(&a->d_za, 2 * tab));
  HIP_TRY(dalloc(&a->d_zb, 2 * tab));
  HIP_TRY(dalloc(&a->d_err, tab));
  HIP_TRY(hipMemcpyAsync(a->d_rcs, rcs.data(), sizeof(double) * nrc,
                         hipMemcpyHostToDevice, a->stream));
  hipLaunchKernelGGL(k_arc_kernels, dim3(arc_blocks(a->g.nele)), dim3(kArcBlock), 0,
                     a->stream, a->g, nrc, a->d_rcs, gauss_legendre(), a->d_S1, a->d_G,
                     a->d_I1, a->d_I5);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

extern "C" void sre_arc_destroy(sre_arc *a) {
  if (!a) return;
  (void)hipSetDevice(a->device);
  (void)hipStreamSynchronize(a->stream);
  delete a;
}

extern "C" sre_status sre_arc_dims(sre_arc *a, int *dims, int *n_rc) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (dims)
    for (int d = 0; d < 3; ++d) dims[d] = a->g.n[d];
  if (n_rc) *n_rc = (int)a->rcs.size();
  return SRE_OK;
}

extern "C" sre_status sre_arc_rc_list(sre_arc *a, double *rcs) {
  if (!a || !rcs) return fail(SRE_ERR_INVALID, "null argument");
  std::copy(a->rcs.begin(), a->rcs.end(), rcs);
  return SRE_OK;
}

extern "C" sre_status sre_arc_clear(sre_arc *a) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  HIP_TRY(hipSetDevice(a->device));
  HIP_TRY(hipMemsetAsync(a->d_counts, 0, sizeof(unsigned) * a->g.nele, a->stream));
  a->ndata = 0;
  a->rho_valid = false;
  return SRE_OK;
}

extern "C" sre_status sre_arc_deposit(sre_arc *a, const float *d_pos, int n,
                                      int stride) {
  if (!a || (n > 0 && !d_pos) || n < 0 || stride < 3)
    return fail(SRE_ERR_INVALID, "sre_arc_deposit: bad arguments");
  HIP_TRY(hipSetDevice(a->device));
  if (n > 0)
    hipLaunchKernelGGL(k_arc_deposit, dim3(arc_blocks(n)), dim3(kArcBlock), 0,
                       a->stream, a->g, n, d_pos, stride, a->d_counts);
  HIP_TRY(hipGetLastError());
  ++a->ndata;
  a->rho_valid = false;
  return SRE_OK;
}

extern "C" sre_status sre_arc_deposit_engine(sre_arc *a, sre_engine *e, int replica) {
  if (!a || !e) return fail(SRE_ERR_INVALID, "null argument");
  const float *pos;
  int n, stride;
  float box[3];
  sre_status s = engine_replica_positions(e, replica, &pos, &n, &stride, box);
  if (s != SRE_OK) return s;
  for (int d = 0; d < 3; ++d)
    if ((double)box[d] != a->g.L[d])
      return fail(SRE_ERR_INVALID, "replica %d box differs from the density grid's",
                  replica);
  hipStream_t es = engine_stream(e);
  if (es != a->stream) {
    HIP_TRY(hipSetDevice(a->device));
    HIP_TRY(hipEventRecord(a->engine_done, es));
    HIP_TRY(hipStreamWaitEvent(a->stream, a->engine_done, 0));
  }
  s = sre_arc_deposit(a, pos, n, stride);
  if (s != SRE_OK || es == a->stream) return s;
  HIP_TRY(hipEventRecord(a->deposit_done, a->stream));
  HIP_TRY(hipStreamWaitEvent(es, a->deposit_done, 0));
  return SRE_OK;
}

// DensityProfile_PiecewiseConst::calculate
static sre_status arc_calculate(sre_arc *a) {
  if (a->rho_valid) return SRE_OK;
  if (a->ndata == 0) return fail(SRE_ERR_NOT_READY, "no density samples deposited");
  const ArcGrid &g = a->g;
  const double dv = (g.L[0] / g.n[0]) * (g.L[1] / g.n[1]) * (g.L[2] / g.n[2]);
  hipLaunchKernelGGL(k_arc_density, dim3(arc_blocks(g.nele)), dim3(kArcBlock), 0,
                     a->stream, g.nele, dv * a->ndata, a->d_counts, a->d_rho);
  HIP_TRY(hipGetLastError());
  a->rho_valid = true;
  return SRE_OK;
}

extern "C" sre_status sre_arc_set_density(sre_arc *a, const double *rho) {
  if (!a || !rho) return fail(SRE_ERR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(a->device));
  HIP_TRY(hipMemcpyAsync(a->d_rho, rho, sizeof(double) * a->g.nele,
                         hipMemcpyHostToDevice, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  a->rho_valid = true;
  a->ndata = 0;
  return SRE_OK;
}

extern "C" sre_status sre_arc_density(sre_arc *a, double *rho) {
  if (!a || !rho) return fail(SRE_ERR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(a->device));
  sre_status s = arc_calculate(a);
  if (s != SRE_OK) return s;
  HIP_TRY(hipMemcpyAsync(rho, a->d_rho, sizeof(double) * a->g.nele,
                         hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

extern "C" sre_status sre_arc_cal_error(sre_arc *a, int keep_parts) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (a->rcs.empty()) return fail(SRE_ERR_NOT_READY, "sre_arc_set_candidates first");
  HIP_TRY(hipSetDevice(a->device));
  sre_status s = arc_calculate(a);
  if (s != SRE_OK) return s;
  const ArcGrid &g = a->g;
  const int nrc = (int)a->rcs.size();
  const size_t tab = (size_t)nrc * g.nele;
  const double vol = g.L[0] * g.L[1] * g.L[2];
  // rho_k = FFT(rho) (forward: exp(-2 pi i ...)), scaled by V / nele in pack
  hipLaunchKernelGGL(k_arc_real_to_complex, dim3(arc_blocks(g.nele)), dim3(kArcBlock),
                     0, a->stream, g.nele, a->d_rho, a->d_rhok);
  s = arc_dft3(a, 1, 0, a->d_rhok, a->d_zr);
  if (s != SRE_OK) return s;
  hipLaunchKernelGGL(k_arc_pack, dim3(arc_blocks(tab)), dim3(kArcBlock), 0, a->stream,
                     g, nrc, vol / g.nele, a->d_S1, a->d_G, a->d_rhok, a->d_za);
  s = arc_dft3(a, 2 * nrc, 1, a->d_za, a->d_zb);
  if (s != SRE_OK) return s;
  if (keep_parts && !a->d_parts) HIP_TRY(dalloc(&a->d_parts, 4 * tab));
  hipLaunchKernelGGL(k_arc_combine, dim3(arc_blocks(tab)), dim3(kArcBlock), 0,
                     a->stream, g, nrc, 1.0 / vol, a->d_za, a->d_err,
                     keep_parts ? a->d_parts : nullptr);
  HIP_TRY(hipGetLastError());
  a->err_valid = true;
  a->parts_valid = keep_parts != 0;
  a->rc_valid = false;
  return SRE_OK;
}

extern "C" sre_status sre_arc_cal_rcut(sre_arc *a, double prec) {
  if (!a || !(prec > 0.0)) return fail(SRE_ERR_INVALID, "sre_arc_cal_rcut: bad arguments");
  if (!a->err_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_error first");
  HIP_TRY(hipSetDevice(a->device));
  hipLaunchKernelGGL(k_arc_bisect, dim3(arc_blocks(a->g.nele)), dim3(kArcBlock), 0,
                     a->stream, a->g.nele, (int)a->rcs.size(), prec, a->d_err,
                     a->d_idx);
  HIP_TRY(hipGetLastError());
  a->rc_valid = true;
  return SRE_OK;
}

extern "C" sre_status sre_arc_refine(sre_arc *a) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (!a->rc_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_rcut first");
  HIP_TRY(hipSetDevice(a->device));
  hipLaunchKernelGGL(k_arc_refine, dim3(arc_blocks(a->g.nele)), dim3(kArcBlock), 0,
                     a->stream, a->g, a->d_idx, a->d_idx2);
  HIP_TRY(hipGetLastError());
  std::swap(a->d_idx, a->d_idx2);
  return SRE_OK;
}

// load_rc: a saved candidate index per cell (the errors are not restored)
extern "C" sre_status sre_arc_set_index(sre_arc *a, const int *idx) {
  if (!a || !idx) return fail(SRE_ERR_INVALID, "null argument");
  if (a->rcs.empty()) return fail(SRE_ERR_NOT_READY, "sre_arc_set_candidates first");
  const int nrc = (int)a->rcs.size();
  for (int i = 0; i < a->g.nele; ++i)
    if (idx[i] < 0 || idx[i] >= nrc)
      return fail(SRE_ERR_INVALID, "index %d of cell %d outside [0, %d)", idx[i], i, nrc);
  HIP_TRY(hipSetDevice(a->device));
  HIP_TRY(hipMemcpyAsync(a->d_idx, idx, sizeof(int) * a->g.nele, hipMemcpyHostToDevice,
                         a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  if (!a->err_valid)   // no error estimate behind this index: report zeros
    HIP_TRY(hipMemsetAsync(a->d_err, 0, sizeof(double) * a->g.nele * nrc, a->stream));
  a->err_valid = true;
  a->rc_valid = true;
  return SRE_OK;
}

extern "C" sre_status sre_arc_result(sre_arc *a, float *rc, double *err, int *idx) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (!a->rc_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_rcut first");
  HIP_TRY(hipSetDevice(a->device));
  const int nele = a->g.nele;
  hipLaunchKernelGGL(k_arc_result, dim3(arc_blocks(nele)), dim3(kArcBlock), 0,
                     a->stream, nele, a->d_idx, a->d_rcs, a->d_err, a->d_rc,
                     a->d_res_err);
  HIP_TRY(hipGetLastError());
  if (rc)
    HIP_TRY(hipMemcpyAsync(rc, a->d_rc, sizeof(float) * nele, hipMemcpyDeviceToHost,
                           a->stream));
  if (err)
    HIP_TRY(hipMemcpyAsync(err, a->d_res_err, sizeof(double) * nele,
                           hipMemcpyDeviceToHost, a->stream));
  if (idx)
    HIP_TRY(hipMemcpyAsync(idx, a->d_idx, sizeof(int) * nele, hipMemcpyDeviceToHost,
                           a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

extern "C" sre_status sre_arc_errors(sre_arc *a, double *err, double *parts) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (!a->err_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_error first");
  if (parts && !a->d_parts)
    return fail(SRE_ERR_NOT_READY, "parts need sre_arc_cal_error(keep_parts = 1)");
  HIP_TRY(hipSetDevice(a->device));
  const size_t tab = a->rcs.size() * (size_t)a->g.nele;
  if (err)
    HIP_TRY(hipMemcpyAsync(err, a->d_err, sizeof(double) * tab, hipMemcpyDeviceToHost,
                           a->stream));
  if (parts)
    HIP_TRY(hipMemcpyAsync(parts, a->d_parts, sizeof(double) * 4 * tab,
                           hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

extern "C" sre_status sre_arc_kernels(sre_arc *a, double *S1, double *G) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (a->rcs.empty()) return fail(SRE_ERR_NOT_READY, "sre_arc_set_candidates first");
  HIP_TRY(hipSetDevice(a->device));
  const size_t tab = a->rcs.size() * (size_t)a->g.nele;
  if (S1)
    HIP_TRY(hipMemcpyAsync(S1, a->d_S1, sizeof(double) * tab, hipMemcpyDeviceToHost,
                           a->stream));
  if (G)
    HIP_TRY(hipMemcpyAsync(G, a->d_G, sizeof(double) * tab, hipMemcpyDeviceToHost,
                           a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

// PressureCorrection::correction: p_aa = sum_cells 1/2 conv_aa[idx](cell)
// rho(cell) dV / V^2 with the current density and candidate per cell
extern "C" sre_status sre_arc_pressure(sre_arc *a, double *p) {
  if (!a || !p) return fail(SRE_ERR_INVALID, "null argument");
  if (!a->rc_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_rcut first");
  HIP_TRY(hipSetDevice(a->device));
  sre_status s = arc_calculate(a);
  if (s != SRE_OK) return s;
  const ArcGrid &g = a->g;
  const int nrc = (int)a->rcs.size();
  const size_t tab = (size_t)nrc * g.nele;
  const double vol = g.L[0] * g.L[1] * g.L[2];
  hipLaunchKernelGGL(k_arc_real_to_complex, dim3(arc_blocks(g.nele)), dim3(kArcBlock),
                     0, a->stream, g.nele, a->d_rho, a->d_rhok);
  s = arc_dft3(a, 1, 0, a->d_rhok, a->d_zr);
  if (s != SRE_OK) return s;
  hipLaunchKernelGGL(k_arc_pack_pressure, dim3(arc_blocks(tab)), dim3(kArcBlock), 0,
                     a->stream, g, nrc, vol / g.nele, a->d_I1, a->d_I5, a->d_rhok,
                     a->d_za);
  s = arc_dft3(a, 2 * nrc, 1, a->d_za, a->d_zb);
  if (s != SRE_OK) return s;
  const unsigned nb = arc_blocks(g.nele);
  hipLaunchKernelGGL(k_arc_pressure_sum, dim3(nb), dim3(kArcBlock), 0, a->stream, g,
                     a->d_idx, a->d_rho, a->d_za, a->d_psum);
  HIP_TRY(hipGetLastError());
  std::vector<double> part((size_t)3 * nb);
  HIP_TRY(hipMemcpyAsync(part.data(), a->d_psum, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  const double dv = (g.L[0] / g.n[0]) * (g.L[1] / g.n[1]) * (g.L[2] / g.n[2]);
  for (int d = 0; d < 3; ++d) {
    double acc = 0.0;
    for (unsigned b = 0; b < nb; ++b) acc += part[(size_t)b * 3 + d];
    p[d] = acc / vol * dv / vol;
  }
  return SRE_OK;
}

extern "C" sre_status sre_arc_pressure_integrals(sre_arc *a, double *I1, double *I5) {
  if (!a) return fail(SRE_ERR_INVALID, "null arc");
  if (a->rcs.empty()) return fail(SRE_ERR_NOT_READY, "sre_arc_set_candidates first");
  HIP_TRY(hipSetDevice(a->device));
  const size_t tab = a->rcs.size() * (size_t)a->g.nele;
  if (I1)
    HIP_TRY(hipMemcpyAsync(I1, a->d_I1, sizeof(double) * tab, hipMemcpyDeviceToHost,
                           a->stream));
  if (I5)
    HIP_TRY(hipMemcpyAsync(I5, a->d_I5, sizeof(double) * tab, hipMemcpyDeviceToHost,
                           a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

// ForceCorr::getForceCorr{X,Y,Z}: the mean error force at each cell's chosen
// candidate, f[3 * cell + a] (FP32, as AssignForceCorr::getForceCorr copies it)
extern "C" sre_status sre_arc_force_grid(sre_arc *a, float *f) {
  if (!a || !f) return fail(SRE_ERR_INVALID, "null argument");
  if (!a->rc_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_rcut first");
  if (!a->parts_valid)
    return fail(SRE_ERR_NOT_READY, "the force grid needs sre_arc_cal_error(a, 1)");
  HIP_TRY(hipSetDevice(a->device));
  const int nele = a->g.nele;
  if (!a->d_fgrid) HIP_TRY(dalloc(&a->d_fgrid, (size_t)3 * nele));
  hipLaunchKernelGGL(k_arc_force_grid, dim3(arc_blocks(nele)), dim3(kArcBlock), 0,
                     a->stream, nele, a->d_idx, a->d_parts, a->d_fgrid);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(f, a->d_fgrid, sizeof(float) * 3 * nele, hipMemcpyDeviceToHost,
                         a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return SRE_OK;
}

// AssignForceCorr::getForceCorr: installs the force grid into an engine replica
extern "C" sre_status sre_arc_apply_force_correction(sre_arc *a, sre_engine *e,
                                                     int replica) {
  if (!a || !e) return fail(SRE_ERR_INVALID, "null argument");
  std::vector<float> f((size_t)3 * a->g.nele);
  sre_status s = sre_arc_force_grid(a, f.data());
  if (s != SRE_OK) return s;
  return sre_set_force_grid(e, replica, a->g.n[0], a->g.n[1], a->g.n[2], f.data());
}

extern "C" sre_status sre_arc_apply(sre_arc *a, sre_engine *e, int replica) {
  if (!a || !e) return fail(SRE_ERR_INVALID, "null argument");
  if (!a->rc_valid) return fail(SRE_ERR_NOT_READY, "sre_arc_cal_rcut first");
  std::vector<float> rc((size_t)a->g.nele);
  sre_status s = sre_arc_result(a, rc.data(), nullptr, nullptr);
  if (s != SRE_OK) return s;
  return sre_set_rc_table(e, replica, a->g.n[0], a->g.n[1], a->g.n[2], rc.data());
}
