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
// Accumulation is exact: per-pair terms and per-atom tail weights are summed
// as integers (scale 2^32) -- in registers, LDS histograms and integer
// reductions -- so the gradients are the same bits for any launch shape;
// the remaining FP64 reductions (tail part) run in a fixed order.
#include "sre_internal.h"

namespace sre {

// integer scale of the pair terms: 2^32, so that rounding each FP32 term down
// (the mixed form's exact sums) biases a total over ~10^8 pairs by < 0.03
constexpr double kGradScale = 4294967296.0;  // 2^32
constexpr double kGradInv = 1.0 / 4294967296.0;
constexpr int kAdjBlock = 256;

__device__ __forceinline__ long long grad_fix(double v) {
  const double t = __builtin_fma(v, kGradScale, kMagic);
  return (long long)(__double_as_longlong(t) - __double_as_longlong(kMagic));
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Pair part.  For a pair (i, j) inside rc_i (s6 = (sig/r)^6, s12 = s6^2,
// ri2 = 1/r^2, g = dLdF_i, d = x_j - x_i, wE = dLdE_r / 2) four terms are
// summed per ordered type pair (a, b):
//     t1 = wE (s12 - s6)              t2 = (2 s12 - s6) ri2 (g.d)
//     t3 = wE (2 s12 - s6)            t4 = (4 s12 - s6) ri2 (g.d)
// and the host forms  dL/deps_ab = 4 T1 - 24 T2  and
// dL/dsig_ab = eps_ab / sig_ab (24 T3 - 144 T4)  from the totals.  Every term
// is taken to an integer at scale 2^32 (the mixed form evaluates the terms
// in FP32 and sums them exactly in FP64 accumulators under round toward
// zero, the FP64 form rounds each FP64 term), so the totals are exact sums,
// the same bits whatever the lane count, block size or list order.
// Register sums per species of j (NT = n_types of 1 or 2; the lane's own
// species is fixed); NT = 0: any n_types, per-wave LDS sums through integer
// atomics.  Every launched wave writes its integer partials, value-major:
// part[(k * npt + p) * n_waves + wave], k = term.
constexpr int kAdjTerms = 4;

template <int T, bool FP64, int NT>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint_pairs(AdjointArgs a) {
  extern __shared__ unsigned long long acc[];  // NT == 0: [waves_per_block][4*npt]
  const int npt = a.n_types * a.n_types;
  const int nv = kAdjTerms * npt;
  const int wib = threadIdx.x / kWave;
  unsigned long long *wacc = acc + (size_t)wib * nv;
  if (NT == 0) {
    for (int k = threadIdx.x % kWave; k < nv; k += kWave) wacc[k] = 0ull;
    __syncthreads();
  }
  constexpr int A = kWave / T;
  constexpr int NR = NT > 0 ? NT : 1;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + lane / T;
  // mixed: FP64 accumulators for all j (accA) and for j of species 1 (accB);
  // FP64 form: integer sums per species of j
  double accA[kAdjTerms], accB[kAdjTerms];
  long long si64[NR][kAdjTerms];
#pragma unroll
  for (int k = 0; k < kAdjTerms; ++k) {
    accA[k] = accB[k] = kAccBase;
#pragma unroll
    for (int c = 0; c < NR; ++c) si64[c][k] = 0;
  }
  int ti = 0;
  if (s < a.n) {
    const uint4 pi = a.pos4[s];
    const int rep = (int)(pi.w >> 8);
    const RepInfo &R = a.reps[rep];
    const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
    const double qdx = R.qd[0], qdy = R.qd[1], qdz = R.qd[2];
    const float rc = a.rcut[s];
    const float rc2 = rc * rc;
    const int nn = a.nnb[s];
    ti = (int)(pi.w & 0xffu);
    // wE and g carry the integer scale 2^32
    const double wE = 0.5 * a.dLdE[rep] * kGradScale;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (a.dLdF) {
      const size_t i3 = (size_t)a.perm[s] * 3;
      gx = a.dLdF[i3] * kGradScale;
      gy = a.dLdF[i3 + 1] * kGradScale;
      gz = a.dLdF[i3 + 2] * kGradScale;
    }
    float wEf = (float)wE, gxf = (float)gx, gyf = (float)gy, gzf = (float)gz;
    float sig2[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const float sg = a.params[ti * a.n_types + c].sig;
      sig2[c] = sg * sg;
    }
    const int *lp = a.list + (size_t)wave * (size_t)a.cap * A + lane;
    const int iters = (nn + T - 1) / T;
    if (!FP64) {
      // the lane constants are materialised before the rounding mode changes
      asm volatile("" : "+v"(wEf), "+v"(gxf), "+v"(gyf), "+v"(gzf));
      round_f64_toward_zero();
    }
    for (int m = 0; m < iters; ++m) {
      const int k = m * T + l;
      if (k >= nn) break;
      const int j = lp[(size_t)m * kWave];
      const uint4 pj = a.pos4[j];
      const float dxf = disp(pj.x, pi.x, qx), dyf = disp(pj.y, pi.y, qy),
                  dzf = disp(pj.z, pi.z, qz);
      const float d2f = dist2(dxf, dyf, dzf);
      if (!(d2f < rc2)) continue;
      const int tj = (int)(pj.w & 0xffu);
      if (FP64) {
        const double dx = (double)(int32_t)(pj.x - pi.x) * qdx;
        const double dy = (double)(int32_t)(pj.y - pi.y) * qdy;
        const double dz = (double)(int32_t)(pj.z - pi.z) * qdz;
        const double ri2 = 1.0 / (dx * dx + dy * dy + dz * dz);
        const double sg = (double)a.params[ti * a.n_types + tj].sig;
        const double sg2 = sg * sg * ri2;
        const double s6 = sg2 * sg2 * sg2, s12 = s6 * s6;
        const double rgd = ri2 * (gx * dx + gy * dy + gz * dz);
        const double b = 2.0 * s12 - s6;
        const double t[kAdjTerms] = {wE * (s12 - s6), b * rgd, wE * b, (4.0 * s12 - s6) * rgd};
        if (NT == 0) {
          const int p = ti * a.n_types + tj;
#pragma unroll
          for (int q = 0; q < kAdjTerms; ++q)
            atomicAdd(&wacc[q * npt + p], (unsigned long long)grad_fix(t[q] * kGradInv));
        } else {
#pragma unroll
          for (int q = 0; q < kAdjTerms; ++q) {
            const long long f = grad_fix(t[q] * kGradInv);
#pragma unroll
            for (int c = 0; c < NR; ++c) si64[c][q] += tj == c ? f : 0;
          }
        }
      } else {
        const float ri2 = __builtin_amdgcn_rcpf(d2f);
        const float sgt = NT == 0 ? a.params[ti * a.n_types + tj].sig : 0.f;
        const float sg2 = (NT == 2 ? (tj == 1 ? sig2[NR - 1] : sig2[0])
                                   : (NT == 1 ? sig2[0] : sgt * sgt)) * ri2;
        const float s6 = sg2 * sg2 * sg2, s12 = s6 * s6;
        const float rgd = ri2 * __builtin_fmaf(gzf, dzf, __builtin_fmaf(gyf, dyf, gxf * dxf));
        const float b = __builtin_fmaf(2.f, s12, -s6);
        const float t[kAdjTerms] = {wEf * (s12 - s6), b * rgd, wEf * b,
                                    __builtin_fmaf(4.f, s12, -s6) * rgd};
#pragma unroll
        for (int q = 0; q < kAdjTerms; ++q) {
          acc_floor(accA[q], t[q]);
          if (NT == 2) acc_floor(accB[q], tj == 1 ? t[q] : 0.f);
        }
        if (NT == 0) {
          // general species: the sums leave the registers per pair
          const int p = ti * a.n_types + tj;
#pragma unroll
          for (int q = 0; q < kAdjTerms; ++q) {
            const long long f = __double_as_longlong(accA[q]) - __double_as_longlong(kAccBase);
            accA[q] = kAccBase;
            atomicAdd(&wacc[q * npt + p], (unsigned long long)f);
          }
        }
      }
    }
    if (!FP64) round_f64_nearest();
    if (!FP64 && NT > 0) {
      const long long b0 = __double_as_longlong(kAccBase);
#pragma unroll
      for (int q = 0; q < kAdjTerms; ++q) {
        const long long all = __double_as_longlong(accA[q]) - b0;
        const long long one = __double_as_longlong(accB[q]) - b0;
        si64[0][q] = NT == 2 ? all - one : all;
        if (NT == 2) si64[NR - 1][q] = one;
      }
    }
  }
  const size_t n_waves = (size_t)gridDim.x * (blockDim.x / kWave);
  if (NT == 0) {
    __syncthreads();
    for (int k = threadIdx.x % kWave; k < nv; k += kWave)
      a.wave_part[(size_t)k * n_waves + wave] = (long long)wacc[k];
  } else {
    // per ordered type pair p = (ta, tb): the lanes of atoms of species ta
    // contribute their species-tb sums
#pragma unroll
    for (int ta = 0; ta < NR; ++ta)
#pragma unroll
      for (int tb = 0; tb < NR; ++tb)
#pragma unroll
        for (int q = 0; q < kAdjTerms; ++q) {
          const long long v = wave_sum_i64(ti == ta ? si64[tb][q] : 0);
          if (lane == 0)
            a.wave_part[(size_t)(q * npt + ta * NR + tb) * n_waves + wave] = v;
        }
  }
}

// The mixed form for one or two species in the force kernel's pipelined
// loop (k_force.hip, k_force_mixed): a wave-uniform number of groups of U list
// rows, list rows two groups ahead through a per-wave buffer resource,
// position gathers one group ahead; a lane past its atom's list or outside
// rc_i gets 1/r^2 = 0, which zeroes all four terms.  Same terms, sums and
// output as k_adjoint_pairs.
template <int T, int NT>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint_piped(AdjointArgs a) {
  constexpr int A = kWave / T;
  constexpr int U = 2;
  const int npt = NT * NT;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const size_t n_waves = (size_t)gridDim.x * (blockDim.x / kWave);
  const int s = wave * A + lane / T;
  const bool act = s < a.n;
  const int s0 = wave * A;
  if (s0 >= a.n) {     // a wave beyond the atoms writes zero partials
    if (lane == 0)
      for (int v = 0; v < kAdjTerms * npt; ++v) a.wave_part[(size_t)v * n_waves + wave] = 0;
    return;
  }
  const int si = act ? s : s0;
  const uint4 pi = a.pos4[si];
  const int rep = (int)(pi.w >> 8);
  const RepInfo &R = a.reps[rep];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  const float rc = a.rcut[si];
  const float rc2 = rc * rc;
  const int nn = act ? a.nnb[si] : 0;
  const int lim = nn - l;
  const int ti = (int)(pi.w & 0xffu);
  int groups;
  if (a.wave_len) {
    const int m = min(a.wave_len[__builtin_amdgcn_readfirstlane(wave)].x, a.cap);
    groups = __builtin_amdgcn_readfirstlane((m + T * U - 1) / (T * U));
  } else {
    int g = (nn + T * U - 1) / (T * U);
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) g = max(g, __shfl_xor(g, o, kWave));
    groups = __builtin_amdgcn_readfirstlane(g);
  }
  // wE and g carry the integer scale 2^32
  const double wE = 0.5 * a.dLdE[rep] * kGradScale;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (a.dLdF && act) {
    const size_t i3 = (size_t)a.perm[si] * 3;
    gx = a.dLdF[i3] * kGradScale;
    gy = a.dLdF[i3 + 1] * kGradScale;
    gz = a.dLdF[i3 + 2] * kGradScale;
  }
  float wEf = (float)wE, gxf = (float)gx, gyf = (float)gy, gzf = (float)gz;
  float sig2[NT];
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    const float sg = a.params[ti * a.n_types + c].sig;
    sig2[c] = sg * sg;
  }
  double accA[kAdjTerms], accB[kAdjTerms];
#pragma unroll
  for (int q = 0; q < kAdjTerms; ++q) accA[q] = accB[q] = kAccBase;
  const size_t tile0 = (size_t)__builtin_amdgcn_readfirstlane(wave) * (size_t)a.cap * A;
  const size_t lrem = a.list_len > tile0 ? a.list_len - tile0 : 0;
  const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<int *>(a.list + tile0), (short)0,
      (int)min(lrem * 4, (size_t)0x7fffffff), 0x00020000);
  const int loff = lane * 4;
  auto row = [&](int r) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(lrs, loff, r * (kWave * 4), 0);
  };
  const __amdgpu_buffer_rsrc_t prs = pos_rsrc_idx(a.pos4, a.n);
  auto eval_group = [&](const uint4 *pj, int g) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool v = (g * U + u) * T < lim;
      const float dx = disp(pj[u].x, pi.x, qx);
      const float dy = disp(pj[u].y, pi.y, qy);
      const float dz = disp(pj[u].z, pi.z, qz);
      const float d2 = dist2(dx, dy, dz);
      const bool in = v && (d2 < rc2);
      const float ri2 = in ? __builtin_amdgcn_rcpf(d2) : 0.f;
      const bool one = NT == 2 && (pj[u].w & 0xffu) == 1u;
      const float sg2 = (one ? sig2[NT - 1] : sig2[0]) * ri2;
      const float s6 = sg2 * sg2 * sg2, s12 = s6 * s6;
      const float rgd = ri2 * __builtin_fmaf(gzf, dz, __builtin_fmaf(gyf, dy, gxf * dx));
      const float b = __builtin_fmaf(2.f, s12, -s6);
      const float t[kAdjTerms] = {wEf * (s12 - s6), b * rgd, wEf * b,
                                  __builtin_fmaf(4.f, s12, -s6) * rgd};
#pragma unroll
      for (int q = 0; q < kAdjTerms; ++q) {
        acc_floor(accA[q], t[q]);
        if (NT == 2) acc_floor(accB[q], one ? t[q] : 0.f);
      }
    }
  };
  auto prefetch = [&](uint4 *p, int *jr, int g) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      p[u] = gather_pos_idx<(NT > 1)>(prs, (g * U + u) * T < lim ? jr[u] : si);
    const int gr = max(min(g + 2, groups - 1), 0);   // past the last group: re-read
#pragma unroll
    for (int u = 0; u < U; ++u) jr[u] = row(gr * U + u);
    asm volatile("" ::: "memory");
  };
  int ja[U], jb[U];
  uint4 pa[U], pb[U];
#pragma unroll
  for (int u = 0; u < U; ++u) ja[u] = row(u);
#pragma unroll
  for (int u = 0; u < U; ++u) jb[u] = row(U + u);
  asm volatile("" ::: "memory");
  prefetch(pa, ja, 0);
  // the lane constants are materialised before the rounding mode changes
  asm volatile("" : "+v"(wEf), "+v"(gxf), "+v"(gyf), "+v"(gzf));
  round_f64_toward_zero();
  int g = 0;
  for (; g + 1 < groups; g += 2) {
    prefetch(pb, jb, g + 1);
    eval_group(pa, g);
    prefetch(pa, ja, g + 2);
    eval_group(pb, g + 1);
  }
  if (g < groups) eval_group(pa, g);  // wave-uniform
  round_f64_nearest();
  const long long b0 = __double_as_longlong(kAccBase);
  long long si64[NT][kAdjTerms];
#pragma unroll
  for (int q = 0; q < kAdjTerms; ++q) {
    const long long all = __double_as_longlong(accA[q]) - b0;
    const long long one = __double_as_longlong(accB[q]) - b0;
    si64[0][q] = NT == 2 ? all - one : all;
    if (NT == 2) si64[NT - 1][q] = one;
  }
  const int tis = act ? ti : -1;   // lanes past the atoms add nothing
#pragma unroll
  for (int ta = 0; ta < NT; ++ta)
#pragma unroll
    for (int tb = 0; tb < NT; ++tb)
#pragma unroll
      for (int q = 0; q < kAdjTerms; ++q) {
        const long long v = wave_sum_i64(tis == ta ? si64[tb][q] : 0);
        if (lane == 0) a.wave_part[(size_t)(q * npt + ta * NT + tb) * n_waves + wave] = v;
      }
}

// Integer sum over waves (exact, so in any order): blocks of kAdjChunk waves
// per value, whose sums are added to the value's integer total.
constexpr int kAdjChunk = 4096;
__global__ __launch_bounds__(256) void k_adjoint_reduce(int n_waves, const long long *wave_part,
                                                        unsigned long long *out) {
  __shared__ long long red[256];
  const int v = blockIdx.y;
  const int w0 = blockIdx.x * kAdjChunk, w1 = min(n_waves, w0 + kAdjChunk);
  long long s = 0;
  const long long *row = wave_part + (size_t)v * n_waves;
  for (int w = w0 + threadIdx.x; w < w1; w += blockDim.x) s += row[w];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out[v], (unsigned long long)red[0]);
}

// Tail weights per tail entry (rc bin x species of i): (sum of dLdF_z,
// sum of dLdE) over its atoms, scale 2^32.  Each block makes an integer
// histogram of a contiguous range of sorted atoms in LDS and writes it whole;
// k_tail_w_reduce adds the blocks' histograms.  Integer sums: exact.
constexpr int kTailWBlocks = 256;
constexpr int kTailWThreads = 1024;
constexpr int kTailWMaxEntries = 2048;   // LDS: 2 x 2048 x 8 B = 32 KB

__global__ __launch_bounds__(kTailWThreads) void k_tail_w_hist(
    int n, int n_entries, const int *__restrict__ tail_idx, const uint32_t *__restrict__ perm,
    const uint4 *__restrict__ pos4, const double *__restrict__ dLdF,
    const double *__restrict__ dLdE, long long *__restrict__ part) {
  __shared__ unsigned long long h[2 * kTailWMaxEntries];
  for (int e = threadIdx.x; e < 2 * n_entries; e += blockDim.x) h[e] = 0ull;
  __syncthreads();
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int s0 = blockIdx.x * per, s1 = min(n, s0 + per);
  for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    const int te = tail_idx[s];
    if (te < 0) continue;
    const double gz = dLdF ? dLdF[(size_t)perm[s] * 3 + 2] : 0.0;
    atomicAdd(&h[2 * te], (unsigned long long)to_fix(gz));
    atomicAdd(&h[2 * te + 1], (unsigned long long)to_fix(dLdE[pos4[s].w >> 8]));
  }
  __syncthreads();
  // value-major, so that one wavefront of the reduction reads a value's row
  for (int e = threadIdx.x; e < 2 * n_entries; e += blockDim.x)
    part[(size_t)e * gridDim.x + blockIdx.x] = (long long)h[e];
}

// One wavefront per value: the blocks' integer histograms added (exact).
__global__ __launch_bounds__(256) void k_tail_w_reduce(int n_blocks, int n_vals,
                                                       const long long *__restrict__ part,
                                                       int64_t *__restrict__ tail_w) {
  const int v = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (v >= n_vals) return;
  long long s = 0;
  for (int b = lane; b < n_blocks; b += kWave) s += part[(size_t)v * n_blocks + b];
  s = wave_sum_i64(s);
  if (lane == 0) tail_w[v] = s;
}

// Entries beyond the LDS histogram: direct integer atomics (exact).
__global__ __launch_bounds__(256) void k_tail_w_atomic(
    int n, const int *__restrict__ tail_idx, const uint32_t *__restrict__ perm,
    const uint4 *__restrict__ pos4, const double *__restrict__ dLdF,
    const double *__restrict__ dLdE, int64_t *__restrict__ tail_w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int te = tail_idx[s];
  if (te < 0) return;
  const double gz = dLdF ? dLdF[(size_t)perm[s] * 3 + 2] : 0.0;
  atomicAdd((unsigned long long *)&tail_w[2 * (size_t)te], (unsigned long long)to_fix(gz));
  atomicAdd((unsigned long long *)&tail_w[2 * (size_t)te + 1],
            (unsigned long long)to_fix(dLdE[pos4[s].w >> 8]));
}

size_t tail_w_scratch(int n_entries) {
  return n_entries <= kTailWMaxEntries ? (size_t)kTailWBlocks * 2 * n_entries : 0;
}

void launch_tail_weights(int n, int n_entries, const int *tail_idx, const uint32_t *perm,
                         const uint4 *pos4, const double *dLdF, const double *dLdE,
                         long long *scratch, int64_t *tail_w, hipStream_t st) {
  if (n <= 0 || n_entries <= 0) return;
  if (n_entries <= kTailWMaxEntries) {
    hipLaunchKernelGGL(k_tail_w_hist, dim3(kTailWBlocks), dim3(kTailWThreads), 0, st, n,
                       n_entries, tail_idx, perm, pos4, dLdF, dLdE, scratch);
    hipLaunchKernelGGL(k_tail_w_reduce, dim3((2 * n_entries * kWave + 255) / 256), dim3(256),
                       0, st, kTailWBlocks, 2 * n_entries, scratch, tail_w);
  } else {
    launch_fill_i64(tail_w, 0, (size_t)2 * n_entries, st);
    hipLaunchKernelGGL(k_tail_w_atomic, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       tail_idx, perm, pos4, dLdF, dLdE, tail_w);
  }
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

template <int T, bool FP64>
static void launch_pairs_f(const AdjointArgs &a, int grid, hipStream_t st) {
  if (!FP64 && a.n_types <= 2) {
    if (a.n_types == 1)
      hipLaunchKernelGGL((k_adjoint_piped<T, 1>), dim3(grid), dim3(kAdjBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_adjoint_piped<T, 2>), dim3(grid), dim3(kAdjBlock), 0, st, a);
  } else if (a.n_types == 1) {
    hipLaunchKernelGGL((k_adjoint_pairs<T, FP64, 1>), dim3(grid), dim3(kAdjBlock), 0, st, a);
  } else if (a.n_types == 2) {
    hipLaunchKernelGGL((k_adjoint_pairs<T, FP64, 2>), dim3(grid), dim3(kAdjBlock), 0, st, a);
  } else {
    const size_t shm = (size_t)(kAdjBlock / kWave) * kAdjTerms * a.n_types * a.n_types *
                       sizeof(unsigned long long);
    hipLaunchKernelGGL((k_adjoint_pairs<T, FP64, 0>), dim3(grid), dim3(kAdjBlock), shm,
                       st, a);
  }
}

template <int T>
static void launch_pairs_t(const AdjointArgs &a, bool fp64, hipStream_t st) {
  const int A = kWave / T;
  const int waves = (a.n + A - 1) / A;
  const int grid = (waves * kWave + kAdjBlock - 1) / kAdjBlock;
  if (fp64)
    launch_pairs_f<T, true>(a, grid, st);
  else
    launch_pairs_f<T, false>(a, grid, st);
}

int adjoint_grid_waves(int n, int lanes) {
  const int A = kWave / lanes;
  const int waves = (n + A - 1) / A;
  return ((waves * kWave + kAdjBlock - 1) / kAdjBlock) * (kAdjBlock / kWave);
}

void launch_adjoint_pairs(const AdjointArgs &a, int precision, int lanes,
                          long long *out, hipStream_t st) {
  const int npt = a.n_types * a.n_types;
  launch_fill_i64(reinterpret_cast<int64_t *>(out), 0, (size_t)kAdjTerms * npt, st);
  switch (lanes) {
    case 1: launch_pairs_t<1>(a, precision != 0, st); break;
    case 2: launch_pairs_t<2>(a, precision != 0, st); break;
    case 4: launch_pairs_t<4>(a, precision != 0, st); break;
    case 8: launch_pairs_t<8>(a, precision != 0, st); break;
    default: launch_pairs_t<16>(a, precision != 0, st); break;
  }
  const int nw = adjoint_grid_waves(a.n, lanes);
  hipLaunchKernelGGL(k_adjoint_reduce, dim3((nw + kAdjChunk - 1) / kAdjChunk, kAdjTerms * npt),
                     dim3(256), 0, st, nw, a.wave_part,
                     reinterpret_cast<unsigned long long *>(out));
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
