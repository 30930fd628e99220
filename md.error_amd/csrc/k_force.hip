// k_force.hip -- per-step kernels: position refresh + skin test, the LJ
// force/energy/virial kernels (mixed and FP64), the statistic reduction and
// the force export.
//
// Force semantics (H/src/InteractionEngine_interface.cu:1204-1281 with
// H/include/NonBondedInteraction.h:866-889):
//   for k < Nneighbor[i]: d = minimg(x_j - x_i); if |d|^2 < rc_i^2:
//       F_i += d * (-4 eps (12 s^12 - 6 s^6) / r^2)
//       E_i += 4 eps (s^12 - s^6) + shift ;  W_i,aa += F_a d_a
//   stored per atom: E_i / 2 and W_i / 2.
// The list is full (no Newton's third law), so every atom owns its sums and no
// atomics touch forces.  A group of T lanes shares one atom's list; entry k of
// the atom is read by lane k % T at row k / T of the wave's tile, so each list
// row is one coalesced 256-byte load.
#include "sre_internal.h"

namespace sre {

// ----------------------------------------------------- update + skin test --

// Thread i refreshes atom i (caller order, so every load is coalesced and
// independent): with a perturbation the new position is written back to the
// bound array first; the lattice coordinates go to the atom's sorted slot
// inv_perm[i].  The skin test compares with the rebuild-time lattice position
// backup[i]; each wavefront keeps the largest squared displacement it has
// seen since the last rebuild in its own slot of skin_part (no atomics).
// With drift.vel set the refresh is the first half of a velocity Verlet step
// (H/src/Integrator.cu:273-324, velocityVerlet_part1) after the thermostat
// and centre-of-mass factors of the step: v <- a2 (a1 v - vcom) + F dt/(2m),
// x <- x + v dt, then x is wrapped back into [0, L) by one add or subtract of
// L (the reference wraps at each rebuild, H/src/MDSystem_interface.cu:19-36).
__device__ __forceinline__ Geom geom_of(const RepInfo &R) {
  Geom g;
  for (int d = 0; d < 3; ++d) {
    g.L[d] = R.L[d];
    g.q[d] = R.q[d];
    g.invq[d] = R.invq[d];
    g.qd[d] = R.qd[d];
  }
  return g;
}

// Every load an atom needs is issued before any of them is used (the
// refresh runs one wavefront per SIMD at 64k atoms, so the memory round trip
// is its cost); with one replica the geometry comes in the kernel arguments.
__global__ __launch_bounds__(256) void k_update_positions(
    int n, const uint32_t *__restrict__ inv_perm, float *__restrict__ pos,
    int stride, const RepInfo *__restrict__ reps, Geom g1, int one_rep,
    uint4 *__restrict__ pos4, const uint4 *__restrict__ backup,
    float4 *__restrict__ res4, float *__restrict__ skin_part, PerturbArgs pert,
    DriftArgs drift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = i / kWave;  // slots exist for the waves holding atoms
  const bool wslot = w * kWave < n;
  // issued first: this wave's running maximum, compared at the end
  const float prev = ((threadIdx.x & (kWave - 1)) == 0 && wslot) ? skin_part[w] : 0.f;
  float d2 = 0.f;
  if (i < n) {
    const uint4 b = backup[i];
    const uint32_t s = inv_perm[i];
    float *p = pos + (size_t)i * stride;
    float x, y, z;
    bool wrap = false;
    if (pert.step_ctr) {
      const float *o = pert.anchor ? pert.anchor + (size_t)i * stride : p;
      const float ox = o[0], oy = o[1], oz = o[2];
      const float3 g = gauss3(pert.seed, *pert.step_ctr, (uint64_t)i);
      x = ox + pert.sigma * g.x;
      y = oy + pert.sigma * g.y;
      z = oz + pert.sigma * g.z;
      p[0] = x;
      p[1] = y;
      p[2] = z;
    } else if (drift.vel) {
      float *v = drift.vel + (size_t)i * 3;
      const float *f = drift.fcall + (size_t)i * 3;
      const float v0 = v[0], v1 = v[1], v2 = v[2];
      const float f0 = f[0], f1 = f[1], f2 = f[2];
      const float p0 = p[0], p1 = p[1], p2 = p[2];
      const float m = drift.mass[drift.types[i]];
      const MDState &ms = *drift.state;
      const float a1 = (float)ms.a1, a2 = (float)ms.a2;
      const float c = 0.5f * drift.dt / m;
      const float vx = a2 * (a1 * v0 - (float)ms.vcom[0]) + c * f0;
      const float vy = a2 * (a1 * v1 - (float)ms.vcom[1]) + c * f1;
      const float vz = a2 * (a1 * v2 - (float)ms.vcom[2]) + c * f2;
      v[0] = vx;
      v[1] = vy;
      v[2] = vz;
      x = p0 + vx * drift.dt;
      y = p1 + vy * drift.dt;
      z = p2 + vz * drift.dt;
      wrap = true;
    } else {
      x = p[0];
      y = p[1];
      z = p[2];
    }
    // R: the atom's replica; with one replica its geometry is an argument
    const Geom G = one_rep ? g1 : geom_of(reps[b.w >> 8]);
    if (wrap) {
      // x + v dt wrapped back into [0, L) by one add or subtract of L (the
      // reference wraps at each rebuild, H/src/MDSystem_interface.cu:19-36)
      x = x < 0.f ? x + G.L[0] : (x >= G.L[0] ? x - G.L[0] : x);
      y = y < 0.f ? y + G.L[1] : (y >= G.L[1] ? y - G.L[1] : y);
      z = z < 0.f ? z + G.L[2] : (z >= G.L[2] ? z - G.L[2] : z);
      p[0] = x;
      p[1] = y;
      p[2] = z;
    }
    const uint32_t ux = quantize(x, G.invq[0]);
    const uint32_t uy = quantize(y, G.invq[1]);
    const uint32_t uz = quantize(z, G.invq[2]);
    pos4[s] = make_uint4(ux, uy, uz, b.w);
    if (res4)
      res4[s] = make_float4(quant_residual(x, G.invq[0], G.qd[0]),
                            quant_residual(y, G.invq[1], G.qd[1]),
                            quant_residual(z, G.invq[2], G.qd[2]), 0.f);
    d2 = dist2(disp(ux, b.x, G.q[0]), disp(uy, b.y, G.q[1]), disp(uz, b.z, G.q[2]));
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
    d2 = fmaxf(d2, __shfl_xor(d2, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && wslot && d2 > prev) skin_part[w] = d2;
}

// ---------------------------------------- next-step refresh in a force launch --
//
// The first nx.n_blocks blocks of a force launch that refreshes the next
// step do only that, one atom per thread in cell order (slot s = block *
// blockDim + thread): the arithmetic of k_update_positions with an
// anchored or free perturbation, reading the cell-order copies of the
// rebuild-time lattice positions and of the anchor (coalesced) and writing
// the next coordinate buffer (coalesced) and the caller's positions.  They
// run beside the force wavefronts and fill their idle issue slots; each
// wavefront keeps its running maximum in skin_part[s / 64].
__device__ __forceinline__ void refresh_block(const NextRefresh &nx, const RepInfo *reps,
                                              int n, uint64_t step, int rb) {
  const int s = rb * (int)blockDim.x + (int)threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = s / kWave;
  const bool wslot = w * kWave < n;
  const float prev = (lane == 0 && wslot) ? nx.skin_part[w] : 0.f;
  float d2 = 0.f;
  if (s < n) {
    const uint32_t i = nx.perm[s];
    const uint4 b = nx.backup_s[s];
    float ox, oy, oz;
    if (nx.anchor_s) {
      const float4 o = nx.anchor_s[s];
      ox = o.x;
      oy = o.y;
      oz = o.z;
    } else {
      const float *o = nx.pos + (size_t)i * nx.stride;
      ox = o[0];
      oy = o[1];
      oz = o[2];
    }
    const float3 g = gauss3(nx.seed, step, (uint64_t)i);
    const float x = ox + nx.sigma * g.x;
    const float y = oy + nx.sigma * g.y;
    const float z = oz + nx.sigma * g.z;
    float *p = nx.pos + (size_t)i * nx.stride;
    p[0] = x;
    p[1] = y;
    p[2] = z;
    const Geom G = nx.one_rep ? nx.g1 : geom_of(reps[b.w >> 8]);
    const uint32_t ux = quantize(x, G.invq[0]);
    const uint32_t uy = quantize(y, G.invq[1]);
    const uint32_t uz = quantize(z, G.invq[2]);
    nx.pos4[s] = make_uint4(ux, uy, uz, b.w);
    if (nx.res4)
      nx.res4[s] = make_float4(quant_residual(x, G.invq[0], G.qd[0]),
                               quant_residual(y, G.invq[1], G.qd[1]),
                               quant_residual(z, G.invq[2], G.qd[2]), 0.f);
    d2 = dist2(disp(ux, b.x, G.q[0]), disp(uy, b.y, G.q[1]), disp(uz, b.z, G.q[2]));
  }
  d2 = wave_max_nonneg(d2);
  if (lane == 0 && wslot && d2 > prev) nx.skin_part[w] = d2;
}

// Block index and count of the force part of a launch (the refresh blocks,
// a multiple of 8 so that the XCD map is unchanged, come first).
template <bool NEXT>
__device__ __forceinline__ bool force_block(const ForceArgs &a, int *bid, int *nb) {
  *bid = blockIdx.x;
  *nb = gridDim.x;
  if (a.count.n_blocks > 0) {   // block-uniform
    if (*bid < a.count.n_blocks) {
      const int i = *bid * (int)blockDim.x + (int)threadIdx.x;
      if (i < a.n) {
        // keys are < n_cells by construction; the clamp keeps the counts in bounds
        const uint32_t k = min(cell_key(a.count.pos, a.count.stride, a.count.rep_of,
                                        a.reps, i),
                               (uint32_t)(a.count.n_cells - 1));
        a.count.keys[i] = k;
        atomicAdd(&a.count.count[k], 1u);
      }
      return false;
    }
    *bid -= a.count.n_blocks;
    *nb -= a.count.n_blocks;
  }
  if (NEXT) {
    if (*bid < a.next.n_blocks) {
      refresh_block(a.next, a.reps, a.n, *a.ctr_in + 1, *bid);
      return false;
    }
    *bid -= a.next.n_blocks;
    *nb -= a.next.n_blocks;
  }
  return true;
}

void launch_update_positions(int n, const uint32_t *inv_perm, float *pos,
                             int stride, const RepInfo *reps, const RepInfo *host_rep0,
                             uint4 *pos4, const uint4 *backup, float4 *res4,
                             float *skin_part, const PerturbArgs *pert,
                             const DriftArgs *drift, hipStream_t st) {
  if (n <= 0) return;
  PerturbArgs p{};
  if (pert) p = *pert;
  DriftArgs d{};
  if (drift) d = *drift;
  Geom g{};
  if (host_rep0)
    for (int k = 0; k < 3; ++k) {
      g.L[k] = host_rep0->L[k];
      g.q[k] = host_rep0->q[k];
      g.invq[k] = host_rep0->invq[k];
      g.qd[k] = host_rep0->qd[k];
    }
  hipLaunchKernelGGL(k_update_positions, dim3((n + 255) / 256), dim3(256), 0,
                     st, n, inv_perm, pos, stride, reps, g, host_rep0 ? 1 : 0, pos4,
                     backup, res4, skin_part, p, d);
}

// ------------------------------------------------------- rebuild info --

__global__ __launch_bounds__(1024) void k_rebuild_info(
    int n_build, const int2 *__restrict__ build_part,
    unsigned long long *counters, int n_skin, float *__restrict__ skin_part,
    int clear_skin, unsigned *skin_bits, unsigned long long *log,
    unsigned *log_idx, int log_cap) {
  __shared__ unsigned long long s_tot[1024 / kWave];
  __shared__ int s_max[1024 / kWave];
  __shared__ float s_d2[1024 / kWave];
  int mx = 0;
  unsigned long long tot = 0;
  float d2 = 0.f;
  for (int k = threadIdx.x; k < n_build; k += blockDim.x) {
    const int2 v = build_part[k];
    mx = max(mx, v.x);
    tot += (unsigned long long)(unsigned)v.y;
  }
  for (int k = threadIdx.x; k < n_skin; k += blockDim.x) {
    d2 = fmaxf(d2, skin_part[k]);
    if (clear_skin) skin_part[k] = 0.f;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mx = max(mx, __shfl_xor(mx, o, kWave));
    tot += __shfl_xor(tot, o, kWave);
    d2 = fmaxf(d2, __shfl_xor(d2, o, kWave));
  }
  const int wv = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_max[wv] = mx;
    s_tot[wv] = tot;
    s_d2[wv] = d2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / kWave); ++k) {
      mx = max(mx, s_max[k]);
      tot += s_tot[k];
      d2 = fmaxf(d2, s_d2[k]);
    }
    if (n_build > 0) {
      counters[0] = (unsigned long long)mx;
      counters[1] = tot;
    }
    if (n_skin > 0) *skin_bits = __float_as_uint(d2);
    if (log) {
      // the next entry of the rebuild log (host memory): the host counts
      // the logged launches it queues, so it knows which entry each wrote
      const unsigned k = *log_idx;
      unsigned long long *dst = log + 3 * (size_t)(k % (unsigned)log_cap);
      dst[0] = (unsigned long long)mx;
      dst[1] = tot;
      dst[2] = (unsigned long long)__float_as_uint(d2);
      *log_idx = k + 1;
      __threadfence_system();
    }
  }
}

// First stage for many partials: block b reduces slice b of the build and
// skin partials (clearing the skin partials when asked) into part2[b] and
// skin2[b], which k_rebuild_info then reduces in one block.  A slice total is
// at most per_b waves x 64 atoms x the list capacity, below 2^32.
__global__ __launch_bounds__(1024) void k_rebuild_slices(
    int n_build, const int2 *__restrict__ build_part, int n_skin,
    float *__restrict__ skin_part, int clear_skin, int per_b, int per_s,
    int2 *part2, float *skin2) {
  __shared__ unsigned s_tot[1024 / kWave];
  __shared__ int s_max[1024 / kWave];
  __shared__ float s_d2[1024 / kWave];
  const int b = blockIdx.x;
  int mx = 0;
  unsigned tot = 0;
  float d2 = 0.f;
  const int b1 = min(n_build, (b + 1) * per_b), s1 = min(n_skin, (b + 1) * per_s);
  for (int k = b * per_b + (int)threadIdx.x; k < b1; k += blockDim.x) {
    const int2 v = build_part[k];
    mx = max(mx, v.x);
    tot += (unsigned)v.y;
  }
  for (int k = b * per_s + (int)threadIdx.x; k < s1; k += blockDim.x) {
    d2 = fmaxf(d2, skin_part[k]);
    if (clear_skin) skin_part[k] = 0.f;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mx = max(mx, __shfl_xor(mx, o, kWave));
    tot += __shfl_xor(tot, o, kWave);
    d2 = fmaxf(d2, __shfl_xor(d2, o, kWave));
  }
  const int wv = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_max[wv] = mx;
    s_tot[wv] = tot;
    s_d2[wv] = d2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / kWave); ++k) {
      mx = max(mx, s_max[k]);
      tot += s_tot[k];
      d2 = fmaxf(d2, s_d2[k]);
    }
    part2[b] = make_int2(mx, (int)tot);
    skin2[b] = d2;
  }
}

void launch_rebuild_info(int n_build, const int2 *build_part,
                         unsigned long long *counters, int n_skin,
                         float *skin_part, int clear_skin, unsigned *skin_bits,
                         hipStream_t st, unsigned long long *log,
                         unsigned *log_idx, int log_cap, int2 *slices) {
  const int most = n_build > n_skin ? n_build : n_skin;
  if (slices && most > kInfoSliceLen) {
    int g = (most + kInfoSliceLen - 1) / kInfoSliceLen;
    g = g < kInfoSlices ? g : kInfoSlices;
    const int per_b = (n_build + g - 1) / g, per_s = (n_skin + g - 1) / g;
    float *skin2 = reinterpret_cast<float *>(slices + kInfoSlices);
    hipLaunchKernelGGL(k_rebuild_slices, dim3(g), dim3(1024), 0, st, n_build, build_part,
                       n_skin, skin_part, clear_skin, per_b, per_s, slices, skin2);
    hipLaunchKernelGGL(k_rebuild_info, dim3(1), dim3(1024), 0, st, n_build > 0 ? g : 0,
                       slices, counters, n_skin > 0 ? g : 0, skin2, 0, skin_bits, log,
                       log_idx, log_cap);
    return;
  }
  hipLaunchKernelGGL(k_rebuild_info, dim3(1), dim3(1024), 0, st, n_build,
                     build_part, counters, n_skin, skin_part, clear_skin,
                     skin_bits, log, log_idx, log_cap);
}

// ------------------------------------------------------ wave reductions --

template <typename V>
__device__ __forceinline__ V group_sum(V v, int T) {
  for (int o = T >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// -------------------------------------------------------------- mixed ----

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Renumbering block b to (b % 8) * q + min(b % 8, r) + b / 8, with
// q = G / 8 and r = G % 8, gives each XCD one contiguous range of blocks and
// so one contiguous range of sorted atoms, whose neighbours then share that
// XCD's L2.  The map is a bijection of [0, G).
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int x = b & 7, q = G >> 3, r = G & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Wave-wide maximum of an int (butterfly over the 64 lanes).
__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

// The loop runs a wave-uniform number of groups of U list rows (the wave's
// longest list / (T U)); a lane whose entry k is past its atom's list
// evaluates its own atom (d = 0) and masks the result.  The loop is software
// pipelined: while group g is evaluated, the U position gathers of group g+1
// and the U list rows of group g+2 are in flight, so the gather latency hides
// behind a group's arithmetic.  The row loads issued for the two groups past
// the wave's last one re-read its last group's rows; those values are never
// used.
// P.A12 and P.B6 carry the fixed-point scale 2^32 (see upload_params), so a
// pair component c is already c*2^32 in FP32.
//
// Fixed-point sums: exact FP64 accumulators under round-toward-zero
// (acc_floor, sre_internal.h); the lane value is bits(acc) - bits(C).
//
// One wave-tile (64/T consecutive sorted atoms) of the force evaluation;
// `wave` is the tile index.
template <int T, int U, bool STATS, bool TAIL, bool MULTI, bool SHIFT>
__device__ __forceinline__ void force_tile(const ForceArgs &a, const PairParam *sp,
                                           const int wave) {
  constexpr int A = kWave / T;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  const int s = wave * A + lane / T;
  const bool act = s < a.n;
  const int s0 = wave * A;
  if (s0 >= a.n) return;  // whole wave beyond the atoms

  const int si = act ? s : s0;
  const uint4 pi = a.pos4[si];
  const int rep = (int)(pi.w >> 8);
  const RepInfo &R = a.reps[rep];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  const float rc = a.rcut[si];
  const float rc2 = rc * rc;
  const int nn = act ? a.nnb[si] : 0;
  // entry (g U + u) T + l of this lane exists iff (g U + u) T < lim, a
  // comparison of a wave-uniform index with one per-lane bound
  const int lim = nn - l;
  const int ti = (int)(pi.w & 0xffu) * a.n_types;
  // the wave's longest list: recorded by the list build, else reduced here
  int groups;
  if (a.wave_len) {
    const int m = min(a.wave_len[__builtin_amdgcn_readfirstlane(wave)].x, a.cap);
    groups = __builtin_amdgcn_readfirstlane((m + T * U - 1) / (T * U));
  } else {
    groups = __builtin_amdgcn_readfirstlane(wave_max_int((nn + T * U - 1) / (T * U)));
  }
  PairParam P = a.params[0];

  // fixed-point sums (see above); exact while the lane's partial sums stay
  // within 2^51 of C, i.e. its sum of |pair component| below 2^51 (checked
  // at the end)
  double ax = kAccBase, ay = kAccBase, az = kAccBase;
  float e = 0.f, wx = 0.f, wy = 0.f, wz = 0.f, fsabs = 0.f;
  // single-type energy without shift: the sum of 1/r^6 and the virial
  // trace, combined at the end (sum of A/r^12 - B/r^6 = -trace/12 -
  // (B/2) sum 1/r^6, since each pair's trace term is fs d^2 = 6B/r^6 -
  // 12A/r^12)
  constexpr bool SPLIT_E = STATS && !MULTI && !SHIFT;
  float e6 = 0.f;
  // list rows through a resource based at this wave's tile: the lane's byte
  // offset is fixed and the row offset is a scalar, so a row load needs no
  // per-lane address arithmetic; the resource ends with the list allocation
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
  // software pipeline over two position buffers: while group g is evaluated
  // from one, the gathers of group g+1 fill the other and the list rows of
  // group g+2 are in flight
  auto eval_group = [&](const uint4 *pj, int g) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool v = (g * U + u) * T < lim;
      const float dx = disp(pj[u].x, pi.x, qx);
      const float dy = disp(pj[u].y, pi.y, qy);
      const float dz = disp(pj[u].z, pi.z, qz);
      const float d2 = dist2(dx, dy, dz);
      const bool in = v && (d2 < rc2);
      if (MULTI) P = sp[ti + (int)(pj[u].w & 0xffu)];
      // a pair outside the cutoff (or past the list) gets 1/r^2 = 0, which
      // zeroes its force and its unshifted energy
      const float ri2 = in ? __builtin_amdgcn_rcpf(d2) : 0.f;
      const float ri6 = ri2 * ri2 * ri2;
      const float fs = ri2 * ri6 * __builtin_fmaf(-P.A12, ri6, P.B6);
      fsabs += fabsf(fs);
      const float fpx = fs * dx, fpy = fs * dy, fpz = fs * dz;
      acc_floor(ax, fpx);    // round toward zero: acc + floor(c), exact
      acc_floor(ay, fpy);
      acc_floor(az, fpz);
      if (STATS) {
        if (SPLIT_E) {
          e6 += ri6;
        } else {
          const float t = __builtin_fmaf(P.A, ri6, -P.B);
          e += SHIFT ? __builtin_fmaf(ri6, t, in ? P.shift : 0.f) : ri6 * t;
        }
        wx = __builtin_fmaf(fpx, dx, wx);
        wy = __builtin_fmaf(fpy, dy, wy);
        wz = __builtin_fmaf(fpz, dz, wz);
      }
    }
  };
  // issue the gathers of group g into p from rows jr, then load rows of
  // group g + 2 into jr; the compiler fence keeps these loads ahead of the
  // arithmetic that follows (it would otherwise sink them to their use).
  // A row entry past the lane's list holds whatever an earlier build left
  // there: it is clamped to a valid atom and its pair is masked in
  // eval_group.
  const unsigned nmax = (unsigned)(a.n - 1);
  auto prefetch = [&](uint4 *p, int *jr, int g) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      p[u] = gather_pos_idx<MULTI>(prs, (int)min((unsigned)jr[u], nmax));
    // (rows of groups past the wave's last one: its last rows again, cache
    // hits, so no row past the wave's lists is fetched)
    const int gr = max(min(g + 2, groups - 1), 0);
#pragma unroll
    for (int u = 0; u < U; ++u) jr[u] = row(gr * U + u);
    asm volatile("" ::: "memory");
  };
  int ja[U], jb[U];   // list rows of the groups due in buffers a and b
  uint4 pa[U], pb[U];
#pragma unroll
  for (int u = 0; u < U; ++u) ja[u] = row(u);
#pragma unroll
  for (int u = 0; u < U; ++u) jb[u] = row(U + u);
  // issue order at the loop entry as on the back edge (rows of g+1, then
  // the gathers of g and the rows of g+2), so that the loop waits only for
  // the loads it is about to use
  asm volatile("" ::: "memory");
  prefetch(pa, ja, 0);
  // the accumulators' FP64 adds round toward zero inside the pair loop
  // (FP32 rounding is a separate field of the mode register, untouched)
  round_f64_toward_zero();
  // whole pairs of groups in the loop (no branch inside its body, so the
  // loads of each half are waited for only where they are used), the odd
  // last group after it
  int g = 0;
  for (; g + 1 < groups; g += 2) {
    prefetch(pb, jb, g + 1);
    eval_group(pa, g);
    prefetch(pa, ja, g + 2);
    eval_group(pb, g + 1);
  }
  if (g < groups) eval_group(pa, g);  // wave-uniform
  round_f64_nearest();
  const int64_t acc0 = __double_as_longlong(kAccBase);
  int64_t Fx = __double_as_longlong(ax) - acc0;
  int64_t Fy = __double_as_longlong(ay) - acc0;
  int64_t Fz = __double_as_longlong(az) - acc0;
  Fx = group_sum(Fx, T);
  Fy = group_sum(Fy, T);
  Fz = group_sum(Fz, T);
  // every lane's sum of |pair component| * 2^32 is at most rc * sum |fs|
  // (each component is fs * d with |d| < rc), and must stay below 2^51 for
  // the accumulators to stay in [2^52, 2^53); 2^50 leaves room for the FP32
  // rounding of sum |fs|
  if (__ballot(act && fsabs * rc >= 1125899906842624.0f) &&
      lane == 0)
    atomicOr(a.err_flag, 1u);
  double E = 0.0, Wx = 0.0, Wy = 0.0, Wz = 0.0;
  if (STATS) {
    E = SPLIT_E ? group_sum(-((double)wx + (double)wy + (double)wz) * (kFixInv / 12.0) -
                                0.5 * (double)P.B * (double)e6, T) * 0.5
                : group_sum((double)e, T) * 0.5;
    Wx = group_sum((double)wx, T) * (0.5 * kFixInv);
    Wy = group_sum((double)wy, T) * (0.5 * kFixInv);
    Wz = group_sum((double)wz, T) * (0.5 * kFixInv);
  }
  const bool own = act && l == 0;
  if (own) {
    if (TAIL) {
      double2 tv = make_double2(0.0, 0.0);
      if (a.tail_atom) {
        tv = a.tail_atom[s];
      } else {
        const int tix = a.tail_idx[s];
        if (tix >= 0) tv = a.tail[tix];
      }
      Fz += to_fix(tv.x);
      E += tv.y;
    }
    a.fx[s] = Fx;
    a.fy[s] = Fy;
    a.fz[s] = Fz;
    if (STATS) {
      double2 *d = reinterpret_cast<double2 *>(a.atom_stats + (size_t)s * 4);
      d[0] = make_double2(E, Wx);
      d[1] = make_double2(Wy, Wz);
    }
  }
}

// Flat launch (persist_waves == 0): one wavefront per tile, blocks mapped to
// XCDs by xcd_block.  Persistent launch: persist_waves wavefronts, each
// taking a contiguous chunk of tiles (wavefront order after the same block
// map, so each XCD works through one contiguous range of atoms).
template <int T, int U, bool STATS, bool TAIL, bool MULTI, bool SHIFT, bool NEXT>
__global__ __launch_bounds__(256) void k_force_mixed(ForceArgs a) {
  // the position refresh of this step has read the counter (stream order)
  if (a.ctr_out && blockIdx.x == 0 && threadIdx.x == 0) *a.ctr_out = *a.ctr_in + 1;
  __shared__ PairParam sp[MULTI ? kMaxPairTypes : 1];
  if (MULTI) {
    for (int t = threadIdx.x; t < a.n_types * a.n_types; t += blockDim.x)
      sp[t] = a.params[t];
    __syncthreads();
  }
  int bid, nb;
  if (!force_block<NEXT>(a, &bid, &nb)) return;
  const int blk = a.xcd_remap ? xcd_block(bid, nb) : bid;
  const int wave = __builtin_amdgcn_readfirstlane((blk * blockDim.x + threadIdx.x) / kWave);
  if (a.persist_waves <= 0) {
    force_tile<T, U, STATS, TAIL, MULTI, SHIFT>(a, sp, wave);
    return;
  }
  constexpr int A = kWave / T;
  const int n_tiles = (a.n + A - 1) / A;
  const int per = (n_tiles + a.persist_waves - 1) / a.persist_waves;
  const int t0 = wave * per, t1 = min(n_tiles, t0 + per);
  for (int t = t0; t < t1; ++t) force_tile<T, U, STATS, TAIL, MULTI, SHIFT>(a, sp, t);
}

// ------------------------------------------------- cell-window form ----
//
// The mixed force evaluation without a neighbor list: one wavefront per
// tile of up to kCellTile consecutive sorted atoms of one cell column,
// T = kWave / kCellTile lanes per atom.  The candidates are the cell runs
// within rc_max + skin of the tile's rebuild-time cells -- a window that
// holds every pair a list of the tile's atoms would -- staged 64 at a time
// in LDS; lane l of an atom takes the staged candidates l, l + T, ... (every
// atom of the wave reads the same LDS words at once).  Each lane masks
// |d| < rc_i at the current positions with the list kernel's per-pair
// arithmetic, so the pair set and every pair component are the list
// kernel's and the fixed-point forces (exact sums, summed over the atom's
// lanes) are bitwise the same; the FP32 virial and energy partial sums run
// in another order.  Rebuild-time cells come from the rebuild-time lattice
// copy in cell order (backup_s).
template <bool STATS, bool TAIL, bool MULTI, bool SHIFT>
__global__ __launch_bounds__(kWave) void k_force_cells(ForceArgs a, CellWindow c) {
  if (a.ctr_out && blockIdx.x == 0 && threadIdx.x == 0) *a.ctr_out = *a.ctr_in + 1;
  __shared__ uint4 cand[kWave];
  __shared__ PairParam sp[MULTI ? kMaxPairTypes : 1];
  if (MULTI) {
    for (int t = threadIdx.x; t < a.n_types * a.n_types; t += blockDim.x)
      sp[t] = a.params[t];
    __syncthreads();
  }
  const int tile = blockIdx.x;
  const int n_tiles = (int)(c.tile_off[c.n_cols - 1] + c.tiles_per_col[c.n_cols - 1]);
  if (tile >= n_tiles) return;   // the grid is an upper bound of the tile count
  int lo = 0, hi = c.n_cols - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int)c.tile_off[mid] <= tile) lo = mid;
    else hi = mid - 1;
  }
  const int col_g = lo;
  int rlo = 0, rhi = c.n_rep - 1;
  while (rlo < rhi) {
    const int mid = (rlo + rhi + 1) >> 1;
    if (a.reps[mid].col_off <= col_g) rlo = mid;
    else rhi = mid - 1;
  }
  const RepInfo &R = a.reps[rlo];
  const int ncx = R.nc[0], ncy = R.nc[1], ncz = R.nc[2];
  const float hx = R.h[0], hy = R.h[1], hz = R.h[2];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  constexpr int T = kWave / kCellTile;
  const int col = col_g - R.col_off;
  const int cx = col / ncy, cy = col - cx * ncy;
  const int cbase = R.cell_off + col * ncz;
  const int a0 = c.start[cbase] + (tile - (int)c.tile_off[col_g]) * kCellTile;
  const int a1 = min(c.start[cbase + ncz], a0 + kCellTile);
  const int lane = threadIdx.x;
  const int l = lane % T;
  const int s = a0 + lane / T;
  const bool act = s < a1;
  const int si = act ? s : a0;
  const uint4 pi = a.pos4[si];
  const float rc = a.rcut[si];
  const float rc2 = rc * rc;
  const int ti = (int)(pi.w & 0xffu) * a.n_types;
  float rlmax = act ? rc + c.skin : 0.f;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) rlmax = fmaxf(rlmax, __shfl_xor(rlmax, o, kWave));
  const float rlm = rlmax * 1.0001f + 1e-4f;
  const float rlm2 = rlm * rlm;
  const int cz_lo = cell_of(c.backup_s[a0].z, ncz), cz_hi = cell_of(c.backup_s[a1 - 1].z, ncz);
  const int kx = (int)ceilf(rlm / hx), ky = (int)ceilf(rlm / hy), kz = (int)ceilf(rlm / hz);
  const bool fullx = 2 * kx + 1 >= ncx, fully = 2 * ky + 1 >= ncy;
  const int ox0 = fullx ? -cx : -kx, ox1 = fullx ? ncx - 1 - cx : kx;
  const int oy0 = fully ? -cy : -ky, oy1 = fully ? ncy - 1 - cy : ky;
  PairParam P = a.params[0];
  double ax = kAccBase, ay = kAccBase, az = kAccBase;
  float e = 0.f, wx = 0.f, wy = 0.f, wz = 0.f, fsabs = 0.f;
  constexpr bool SPLIT_E = STATS && !MULTI && !SHIFT;
  float e6 = 0.f;
  round_f64_toward_zero();
  for (int ox = ox0; ox <= ox1; ++ox) {
    const float gx = fullx ? 0.f : fmaxf(0.f, (float)(abs(ox) - 1) * hx);
    if (gx * gx >= rlm2) continue;
    int cxw = cx + ox;
    cxw += cxw < 0 ? ncx : 0;
    cxw -= cxw >= ncx ? ncx : 0;
    for (int oy = oy0; oy <= oy1; ++oy) {
      const float gy = fully ? 0.f : fmaxf(0.f, (float)(abs(oy) - 1) * hy);
      if (gx * gx + gy * gy >= rlm2) continue;
      int cyw = cy + oy;
      cyw += cyw < 0 ? ncy : 0;
      cyw -= cyw >= ncy ? ncy : 0;
      const int colbase = R.cell_off + (cxw * ncy + cyw) * ncz;
      int ra0 = 0, rb0 = ncz - 1, ra1 = 0, rb1 = -1;
      const int zl = cz_lo - kz, zh = cz_hi + kz;
      if (zh - zl + 1 < ncz) {
        if (zl < 0) {
          ra0 = zl + ncz; rb0 = ncz - 1; ra1 = 0; rb1 = zh;
        } else if (zh >= ncz) {
          ra0 = zl; rb0 = ncz - 1; ra1 = 0; rb1 = zh - ncz;
        } else {
          ra0 = zl; rb0 = zh;
        }
      }
      for (int run = 0; run < 2; ++run) {
        const int ca = run ? ra1 : ra0, cb = run ? rb1 : rb0;
        if (cb < ca) continue;
        // runs lie inside [0, n) by construction; the clamps keep every
        // load in bounds whatever the offsets hold
        const int p0 = max(c.start[colbase + ca], 0);
        const int p1 = min(c.start[colbase + cb + 1], a.n);
        for (int p = p0; p < p1; p += kWave) {
          const int q = p + lane;
          cand[lane] = a.pos4[q < p1 ? q : p0];
          __syncthreads();
          const int m = min(kWave, p1 - p);
          for (int k = l; k < m; k += T) {
            const uint4 cj = cand[k];
            const float dx = disp(cj.x, pi.x, qx);
            const float dy = disp(cj.y, pi.y, qy);
            const float dz = disp(cj.z, pi.z, qz);
            const float d2 = dist2(dx, dy, dz);
            const bool in = act && (p + k != s) && (d2 < rc2);
            if (MULTI) P = sp[ti + (int)(cj.w & 0xffu)];
            const float ri2 = in ? __builtin_amdgcn_rcpf(d2) : 0.f;
            const float ri6 = ri2 * ri2 * ri2;
            const float fs = ri2 * ri6 * __builtin_fmaf(-P.A12, ri6, P.B6);
            fsabs += fabsf(fs);
            const float fpx = fs * dx, fpy = fs * dy, fpz = fs * dz;
            acc_floor(ax, fpx);
            acc_floor(ay, fpy);
            acc_floor(az, fpz);
            if (STATS) {
              if (SPLIT_E) {
                e6 += ri6;
              } else {
                const float t = __builtin_fmaf(P.A, ri6, -P.B);
                e += SHIFT ? __builtin_fmaf(ri6, t, in ? P.shift : 0.f) : ri6 * t;
              }
              wx = __builtin_fmaf(fpx, dx, wx);
              wy = __builtin_fmaf(fpy, dy, wy);
              wz = __builtin_fmaf(fpz, dz, wz);
            }
          }
          __syncthreads();
        }
      }
    }
  }
  round_f64_nearest();
  // the partial sums pass through an (empty) asm after the mode switch, so
  // the compiler cannot schedule their FP64 conversions in front of it
  asm volatile("" : "+v"(e), "+v"(e6), "+v"(wx), "+v"(wy), "+v"(wz), "+v"(fsabs));
  const int64_t acc0 = __double_as_longlong(kAccBase);
  int64_t Fx = group_sum(__double_as_longlong(ax) - acc0, T);
  int64_t Fy = group_sum(__double_as_longlong(ay) - acc0, T);
  int64_t Fz = group_sum(__double_as_longlong(az) - acc0, T);
  if (__ballot(act && fsabs * rc >= 1125899906842624.0f) && lane == 0)
    atomicOr(a.err_flag, 1u);
  double E = 0.0, Wx = 0.0, Wy = 0.0, Wz = 0.0;
  if (STATS) {
    E = SPLIT_E ? group_sum(-((double)wx + (double)wy + (double)wz) * (kFixInv / 12.0) -
                                0.5 * (double)P.B * (double)e6, T) * 0.5
                : group_sum((double)e, T) * 0.5;
    Wx = group_sum((double)wx, T) * (0.5 * kFixInv);
    Wy = group_sum((double)wy, T) * (0.5 * kFixInv);
    Wz = group_sum((double)wz, T) * (0.5 * kFixInv);
  }
  if (!act || l != 0) return;
  if (TAIL) {
    double2 tv = make_double2(0.0, 0.0);
    if (a.tail_atom) {
      tv = a.tail_atom[s];
    } else {
      const int tix = a.tail_idx[s];
      if (tix >= 0) tv = a.tail[tix];
    }
    Fz += to_fix(tv.x);
    E += tv.y;
  }
  a.fx[s] = Fx;
  a.fy[s] = Fy;
  a.fz[s] = Fz;
  if (STATS) {
    double2 *d = reinterpret_cast<double2 *>(a.atom_stats + (size_t)s * 4);
    d[0] = make_double2(E, Wx);
    d[1] = make_double2(Wy, Wz);
  }
}

void launch_force_cells(bool stats, bool tail, const ForceArgs &a, const CellWindow &c,
                        hipStream_t st) {
  if (a.n <= 0 || c.n_cols <= 0) return;
  const int max_tiles = (a.n + kCellTile - 1) / kCellTile + c.n_cols;
  const bool multi = a.n_types > 1, shift = a.has_shift != 0;
#define SRE_CELLS(S, T, M, H)                                                       \
  hipLaunchKernelGGL((k_force_cells<S, T, M, H>), dim3(max_tiles), dim3(kWave), 0, st, \
                     a, c)
  if (stats) {
    if (tail) {
      if (multi) { if (shift) SRE_CELLS(true, true, true, true); else SRE_CELLS(true, true, true, false); }
      else { if (shift) SRE_CELLS(true, true, false, true); else SRE_CELLS(true, true, false, false); }
    } else {
      if (multi) { if (shift) SRE_CELLS(true, false, true, true); else SRE_CELLS(true, false, true, false); }
      else { if (shift) SRE_CELLS(true, false, false, true); else SRE_CELLS(true, false, false, false); }
    }
  } else {
    if (tail) {
      if (multi) SRE_CELLS(false, true, true, false); else SRE_CELLS(false, true, false, false);
    } else {
      if (multi) SRE_CELLS(false, false, true, false); else SRE_CELLS(false, false, false, false);
    }
  }
#undef SRE_CELLS
}

// --------------------------------------------------------------- FP64 ----

template <int T, bool STATS, bool TAIL, bool MULTI, bool NEXT>
__global__ __launch_bounds__(256) void k_force_fp64(ForceArgs a) {
  if (a.ctr_out && blockIdx.x == 0 && threadIdx.x == 0) *a.ctr_out = *a.ctr_in + 1;
  __shared__ PairParamD sp[MULTI ? kMaxPairTypes : 1];
  if (MULTI) {
    for (int t = threadIdx.x; t < a.n_types * a.n_types; t += blockDim.x)
      sp[t] = a.paramsd[t];
    __syncthreads();
  }
  constexpr int A = kWave / T;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  int bid, nb;
  if (!force_block<NEXT>(a, &bid, &nb)) return;
  const int blk = a.xcd_remap ? xcd_block(bid, nb) : bid;
  const int wave = (blk * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + lane / T;
  const bool act = s < a.n;
  const int s0 = wave * A;
  if (s0 >= a.n) return;
  const int sa = act ? s : s0;
  const uint4 pi = a.pos4[sa];
  const float4 ri = a.res4[sa];
  const float rc = a.rcut[sa];
  const int nn = act ? a.nnb[sa] : 0;

  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  if (act) {
    const RepInfo &R = a.reps[pi.w >> 8];
    const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
    const double qdx = R.qd[0], qdy = R.qd[1], qdz = R.qd[2];
    const float rc2 = rc * rc;
    const int ti = (int)(pi.w & 0xffu) * a.n_types;
    const int iters = (nn + T - 1) / T;
    PairParamD P = a.paramsd[0];
    const int *lp = a.list + (size_t)wave * (size_t)a.cap * A + lane;
    for (int m = 0; m < iters; ++m) {
      const int k = m * T + l;
      const bool v = k < nn;
      const int j = v ? lp[(size_t)m * kWave] : s;
      const uint4 pj = a.pos4[j];
      // the pair set is decided by the FP32 predicate shared with the mixed path
      const float d2f = dist2(disp(pj.x, pi.x, qx), disp(pj.y, pi.y, qy),
                              disp(pj.z, pi.z, qz));
      const bool in = v && (d2f < rc2);
      if (MULTI) P = sp[ti + (int)(pj.w & 0xffu)];
      const float4 rj = a.res4[j];
      const double dx = (double)(int32_t)(pj.x - pi.x) * qdx + ((double)rj.x - (double)ri.x);
      const double dy = (double)(int32_t)(pj.y - pi.y) * qdy + ((double)rj.y - (double)ri.y);
      const double dz = (double)(int32_t)(pj.z - pi.z) * qdz + ((double)rj.z - (double)ri.z);
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double ri2 = in ? 1.0 / d2 : 0.0;
      const double ri6 = ri2 * ri2 * ri2;
      const double fs = ri2 * ri6 * (P.B6 - P.A12 * ri6);
      fx += fs * dx;
      fy += fs * dy;
      fz += fs * dz;
      if (STATS) {
        e += in ? ri6 * (P.A * ri6 - P.B) + P.shift : 0.0;
        wx += fs * dx * dx;
        wy += fs * dy * dy;
        wz += fs * dz * dz;
      }
    }
  }
  fx = group_sum(fx, T);
  fy = group_sum(fy, T);
  fz = group_sum(fz, T);
  double E = 0.0, Wx = 0.0, Wy = 0.0, Wz = 0.0;
  if (STATS) {
    E = group_sum(e, T) * 0.5;
    Wx = group_sum(wx, T) * 0.5;
    Wy = group_sum(wy, T) * 0.5;
    Wz = group_sum(wz, T) * 0.5;
  }
  const bool own = act && l == 0;
  if (own) {
    if (TAIL) {
      double2 tv = make_double2(0.0, 0.0);
      if (a.tail_atom) {
        tv = a.tail_atom[s];
      } else {
        const int ti = a.tail_idx[s];
        if (ti >= 0) tv = a.tail[ti];
      }
      fz += tv.x;
      E += tv.y;
    }
    a.fxd[s] = fx;
    a.fyd[s] = fy;
    a.fzd[s] = fz;
    if (STATS) {
      double2 *d = reinterpret_cast<double2 *>(a.atom_stats + (size_t)s * 4);
      d[0] = make_double2(E, Wx);
      d[1] = make_double2(Wy, Wz);
    }
  }
}

// The FP64 form in the mixed kernel's structure: the same wave-uniform,
// software-pipelined loop (list rows two groups ahead through a per-wave
// resource, indexed gathers of the next group's lattice positions and
// sub-lattice residuals, one list row per group), the same pair set (the
// FP32 predicate) and FP64 arithmetic and sums as k_force_fp64.  Selected
// for the FP64 form by force_unroll = 1; measured beside the plain form.
template <int T, bool STATS, bool TAIL, bool MULTI, bool NEXT>
__global__ __launch_bounds__(256) void k_force_fp64_piped(ForceArgs a) {
  if (a.ctr_out && blockIdx.x == 0 && threadIdx.x == 0) *a.ctr_out = *a.ctr_in + 1;
  __shared__ PairParamD sp[MULTI ? kMaxPairTypes : 1];
  if (MULTI) {
    for (int t = threadIdx.x; t < a.n_types * a.n_types; t += blockDim.x)
      sp[t] = a.paramsd[t];
    __syncthreads();
  }
  constexpr int A = kWave / T;
  const int lane = threadIdx.x & (kWave - 1);
  const int l = lane % T;
  int bid, nb;
  if (!force_block<NEXT>(a, &bid, &nb)) return;
  const int blk = a.xcd_remap ? xcd_block(bid, nb) : bid;
  const int wave = (blk * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + lane / T;
  const bool act = s < a.n;
  const int s0 = wave * A;
  if (s0 >= a.n) return;
  const int si = act ? s : s0;
  const uint4 pi = a.pos4[si];
  const RepInfo &R = a.reps[pi.w >> 8];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  const double qdx = R.qd[0], qdy = R.qd[1], qdz = R.qd[2];
  const float4 ri = a.res4[si];
  const float rc = a.rcut[si];
  const float rc2 = rc * rc;
  const int nn = act ? a.nnb[si] : 0;
  const int lim = nn - l;
  const int ti = (int)(pi.w & 0xffu) * a.n_types;
  const int groups = __builtin_amdgcn_readfirstlane(wave_max_int((nn + T - 1) / T));
  PairParamD P = a.paramsd[0];
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
  const __amdgpu_buffer_rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float4 *>(a.res4), (short)16, a.n * 16, 0x00020000);
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  struct Pj { uint4 p; float3 r; };
  auto fetch = [&](Pj &q, int &jr, int g) {
    const int j = g * T < lim ? jr : si;
    q.p = gather_pos_idx<MULTI>(prs, j);
    const sre_u32x3 v = sre_struct_load_v3(rrs, j, 0, 0, 0);
    q.r = make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
    jr = row(g + 2);
    asm volatile("" ::: "memory");
  };
  auto eval = [&](const Pj &q, int g) {
    const bool v = g * T < lim;
    const uint4 pj = q.p;
    const float d2f = dist2(disp(pj.x, pi.x, qx), disp(pj.y, pi.y, qy),
                            disp(pj.z, pi.z, qz));
    const bool in = v && (d2f < rc2);
    if (MULTI) P = sp[ti + (int)(pj.w & 0xffu)];
    const double dx = (double)(int32_t)(pj.x - pi.x) * qdx + ((double)q.r.x - (double)ri.x);
    const double dy = (double)(int32_t)(pj.y - pi.y) * qdy + ((double)q.r.y - (double)ri.y);
    const double dz = (double)(int32_t)(pj.z - pi.z) * qdz + ((double)q.r.z - (double)ri.z);
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double ri2 = in ? 1.0 / d2 : 0.0;
    const double ri6 = ri2 * ri2 * ri2;
    const double fs = ri2 * ri6 * (P.B6 - P.A12 * ri6);
    fx += fs * dx;
    fy += fs * dy;
    fz += fs * dz;
    if (STATS) {
      e += in ? ri6 * (P.A * ri6 - P.B) + P.shift : 0.0;
      wx += fs * dx * dx;
      wy += fs * dy * dy;
      wz += fs * dz * dz;
    }
  };
  int ja = row(0), jb = row(1);
  asm volatile("" ::: "memory");
  Pj qa, qb;
  fetch(qa, ja, 0);
  int g = 0;
  for (; g + 1 < groups; g += 2) {
    fetch(qb, jb, g + 1);
    eval(qa, g);
    fetch(qa, ja, g + 2);
    eval(qb, g + 1);
  }
  if (g < groups) eval(qa, g);
  fx = group_sum(fx, T);
  fy = group_sum(fy, T);
  fz = group_sum(fz, T);
  double E = 0.0, Wx = 0.0, Wy = 0.0, Wz = 0.0;
  if (STATS) {
    E = group_sum(e, T) * 0.5;
    Wx = group_sum(wx, T) * 0.5;
    Wy = group_sum(wy, T) * 0.5;
    Wz = group_sum(wz, T) * 0.5;
  }
  if (act && l == 0) {
    if (TAIL) {
      double2 tv = make_double2(0.0, 0.0);
      if (a.tail_atom) {
        tv = a.tail_atom[s];
      } else {
        const int tix = a.tail_idx[s];
        if (tix >= 0) tv = a.tail[tix];
      }
      fz += tv.x;
      E += tv.y;
    }
    a.fxd[s] = fx;
    a.fyd[s] = fy;
    a.fzd[s] = fz;
    if (STATS) {
      double2 *d = reinterpret_cast<double2 *>(a.atom_stats + (size_t)s * 4);
      d[0] = make_double2(E, Wx);
      d[1] = make_double2(Wy, Wz);
    }
  }
}

int force_num_waves(int n, int lanes) {
  const int A = kWave / lanes;
  return (n + A - 1) / A;
}

template <int T, int U, bool STATS, bool TAIL, bool SHIFT, bool NEXT>
static void launch_mixed_n(int grid, int block, bool multi, const ForceArgs &a,
                           hipStream_t st) {
  if (multi)
    hipLaunchKernelGGL((k_force_mixed<T, U, STATS, TAIL, true, SHIFT, NEXT>), dim3(grid),
                       dim3(block), 0, st, a);
  else
    hipLaunchKernelGGL((k_force_mixed<T, U, STATS, TAIL, false, SHIFT, NEXT>), dim3(grid),
                       dim3(block), 0, st, a);
}

template <int T, int U, bool STATS, bool TAIL, bool SHIFT>
static void launch_mixed_s(int grid, int block, bool multi, const ForceArgs &a,
                           hipStream_t st) {
  if (a.next.active)
    launch_mixed_n<T, U, STATS, TAIL, SHIFT, true>(grid, block, multi, a, st);
  else
    launch_mixed_n<T, U, STATS, TAIL, SHIFT, false>(grid, block, multi, a, st);
}

template <int T, bool STATS, bool TAIL, bool MULTI>
static void launch_fp64_n(int grid, int block, bool piped, const ForceArgs &a,
                          hipStream_t st) {
  if (piped) {
    if (a.next.active)
      hipLaunchKernelGGL((k_force_fp64_piped<T, STATS, TAIL, MULTI, true>), dim3(grid),
                         dim3(block), 0, st, a);
    else
      hipLaunchKernelGGL((k_force_fp64_piped<T, STATS, TAIL, MULTI, false>), dim3(grid),
                         dim3(block), 0, st, a);
  } else {
    if (a.next.active)
      hipLaunchKernelGGL((k_force_fp64<T, STATS, TAIL, MULTI, true>), dim3(grid),
                         dim3(block), 0, st, a);
    else
      hipLaunchKernelGGL((k_force_fp64<T, STATS, TAIL, MULTI, false>), dim3(grid),
                         dim3(block), 0, st, a);
  }
}

// the energy shift only enters the statistics
template <int T, int U, bool STATS, bool TAIL>
static void launch_mixed_u(int grid, int block, bool multi, const ForceArgs &a,
                           hipStream_t st) {
  if (STATS && a.has_shift)
    launch_mixed_s<T, U, STATS, TAIL, STATS>(grid, block, multi, a, st);
  else
    launch_mixed_s<T, U, STATS, TAIL, false>(grid, block, multi, a, st);
}

template <int T, bool STATS, bool TAIL>
static void launch_force_t(int precision, int unroll, int block,
                           const ForceArgs &a, hipStream_t st) {
  const int waves = force_num_waves(a.n, T);
  int grid = (waves * kWave + block - 1) / block;
  const bool multi = a.n_types > 1;
  if (precision == 0 && a.persist_waves > 0) grid = a.persist_waves * kWave / block;
  // the next step's refresh and the next rebuild's histogram: blocks of their
  // own in front (multiples of 8, so each force block keeps its XCD)
  if (a.next.active) grid += a.next.n_blocks;
  grid += a.count.n_blocks;
  if (precision == 0) {
    if (unroll == 1) launch_mixed_u<T, 1, STATS, TAIL>(grid, block, multi, a, st);
    else if (unroll == 2) launch_mixed_u<T, 2, STATS, TAIL>(grid, block, multi, a, st);
    else launch_mixed_u<T, 4, STATS, TAIL>(grid, block, multi, a, st);
  } else {   // unroll 1: the FP64 form in the pipelined structure
    if (multi) launch_fp64_n<T, STATS, TAIL, true>(grid, block, unroll == 1, a, st);
    else launch_fp64_n<T, STATS, TAIL, false>(grid, block, unroll == 1, a, st);
  }
}

template <int T>
static void launch_force_lanes(int precision, int unroll, int block, bool stats,
                               bool tail, const ForceArgs &a, hipStream_t st) {
  if (stats) {
    if (tail) launch_force_t<T, true, true>(precision, unroll, block, a, st);
    else launch_force_t<T, true, false>(precision, unroll, block, a, st);
  } else {
    if (tail) launch_force_t<T, false, true>(precision, unroll, block, a, st);
    else launch_force_t<T, false, false>(precision, unroll, block, a, st);
  }
}

void launch_force(int precision, int lanes, int unroll, int block, bool stats,
                  bool tail, const ForceArgs &a, hipStream_t st) {
  if (a.n <= 0) return;
  switch (lanes) {
    case 1: launch_force_lanes<1>(precision, unroll, block, stats, tail, a, st); break;
    case 2: launch_force_lanes<2>(precision, unroll, block, stats, tail, a, st); break;
    case 4: launch_force_lanes<4>(precision, unroll, block, stats, tail, a, st); break;
    case 8: launch_force_lanes<8>(precision, unroll, block, stats, tail, a, st); break;
    default: launch_force_lanes<16>(precision, unroll, block, stats, tail, a, st); break;
  }
}

// ------------------------------------------------------ statistic reduce --

// Replica sums of the per-atom (E, Wxx, Wyy, Wzz) in a fixed order: block
// (b, r) sums the b-th of kStatBlocks contiguous chunks of replica r's atoms
// (contiguous in sorted order; thread k takes atoms k, k + 256, ... of the
// chunk, then a fixed shuffle tree and LDS pass); k_stats_final adds the
// chunks in order.
__global__ __launch_bounds__(256) void k_stats_partial(const RepInfo *reps,
                                                       const double *atom_stats,
                                                       double *part) {
  __shared__ double red[256 / kWave][4];
  const int b = blockIdx.x, r = blockIdx.y;
  const RepInfo &R = reps[r];
  const int chunk = (R.n_atoms + kStatBlocks - 1) / kStatBlocks;
  const int i0 = b * chunk, i1 = min(R.n_atoms, i0 + chunk);
  double acc[4] = {0, 0, 0, 0};
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const double2 *d = reinterpret_cast<const double2 *>(
        atom_stats + (size_t)(R.atom_off + i) * 4);
    const double2 v0 = d[0], v1 = d[1];
    acc[0] += v0.x; acc[1] += v0.y; acc[2] += v1.x; acc[3] += v1.y;
  }
  for (int c = 0; c < 4; ++c) acc[c] = wave_sum(acc[c]);
  const int wv = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0)
    for (int c = 0; c < 4; ++c) red[wv][c] = acc[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / kWave); ++k)
      for (int c = 0; c < 4; ++c) acc[c] += red[k][c];
    for (int c = 0; c < 4; ++c) part[((size_t)r * kStatBlocks + b) * 4 + c] = acc[c];
  }
}

__global__ __launch_bounds__(256) void k_stats_final(int n_rep, const double *part,
                                                     double *stats) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // (replica, component)
  if (t >= 4 * n_rep) return;
  const int r = t / 4, c = t % 4;
  double acc = 0.0;
  for (int b = 0; b < kStatBlocks; ++b) acc += part[((size_t)r * kStatBlocks + b) * 4 + c];
  stats[t] = acc;
}

void launch_stats_reduce(int n_rep, const RepInfo *reps,
                         const double *atom_stats, double *part, double *stats,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_stats_partial, dim3(kStatBlocks, n_rep), dim3(256), 0, st,
                     reps, atom_stats, part);
  hipLaunchKernelGGL(k_stats_final, dim3((4 * n_rep + 255) / 256), dim3(256), 0,
                     st, n_rep, part, stats);
}

// Pairs of the list inside the cutoff (d^2 < rc_i^2, the force kernel's
// mask), counted per atom and summed per replica with integer atomics (one
// per wave and replica run, order independent).  On demand only.
__global__ __launch_bounds__(256) void k_count_in_rc(
    int n, int T, int cap, const uint4 *__restrict__ pos4,
    const float *__restrict__ rcut, const int *__restrict__ nnb,
    const int *__restrict__ list, const RepInfo *__restrict__ reps,
    unsigned long long *counts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int c = 0, rep = -1;
  if (s < n) {
    const uint4 pi = pos4[s];
    rep = (int)(pi.w >> 8);
    const RepInfo &R = reps[rep];
    const float rc = rcut[s];
    const float rc2 = rc * rc;
    const int A = kWave / T, lt = __builtin_ctz((unsigned)T);
    const size_t base = (size_t)(s / A) * (size_t)cap * A + (size_t)((s % A) * T);
    const int nn = nnb[s];
    for (int k = 0; k < nn; ++k) {
      const uint4 pj = pos4[list[base + ((size_t)(k >> lt) << 6) + (k & (T - 1))]];
      const float d2 = dist2(disp(pj.x, pi.x, R.q[0]), disp(pj.y, pi.y, R.q[1]),
                             disp(pj.z, pi.z, R.q[2]));
      c += d2 < rc2 ? 1 : 0;
    }
  }
  // a wave holds at most a few replicas: one atomic per distinct replica
  for (;;) {
    int r0 = rep < 0 ? 0x7fffffff : rep;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) r0 = min(r0, __shfl_xor(r0, o, kWave));
    if (r0 == 0x7fffffff) break;  // wave-uniform
    unsigned long long v = (rep == r0) ? (unsigned long long)c : 0ull;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&counts[r0], v);
    if (rep == r0) rep = -1;
  }
}

void launch_count_in_rc(int n, int lanes, int cap, const uint4 *pos4,
                        const float *rcut, const int *nnb, const int *list,
                        const RepInfo *reps, unsigned long long *counts,
                        hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_count_in_rc, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     lanes, cap, pos4, rcut, nnb, list, reps, counts);
}

// ----------------------------------------------------------------- export --

__global__ __launch_bounds__(256) void k_export_forces(
    int n, const uint32_t *perm, const int64_t *fx, const int64_t *fy,
    const int64_t *fz, const double *fxd, const double *fyd, const double *fzd,
    double *out, int64_t *out_fix) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = (size_t)perm[s] * 3;
  if (fx) {
    const int64_t X = fx[s], Y = fy[s], Z = fz[s];
    if (out) {
      out[i] = (double)X * kFixInv;
      out[i + 1] = (double)Y * kFixInv;
      out[i + 2] = (double)Z * kFixInv;
    }
    if (out_fix) {
      out_fix[i] = X; out_fix[i + 1] = Y; out_fix[i + 2] = Z;
    }
  } else if (out) {
    out[i] = fxd[s];
    out[i + 1] = fyd[s];
    out[i + 2] = fzd[s];
  }
}

void launch_export_forces(int n, const uint32_t *perm, const int64_t *fx,
                          const int64_t *fy, const int64_t *fz,
                          const double *fxd, const double *fyd,
                          const double *fzd, double *out, int64_t *out_fix,
                          hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_export_forces, dim3((n + 255) / 256), dim3(256), 0, st,
                     n, perm, fx, fy, fz, fxd, fyd, fzd, out, out_fix);
}

__global__ __launch_bounds__(256) void k_add_forces_f32(
    int n, const uint32_t *perm, const int64_t *fx, const int64_t *fy,
    const int64_t *fz, const double *fxd, const double *fyd, const double *fzd,
    float *ox, float *oy, float *oz) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t i = perm[s];
  if (fx) {
    ox[i] += (float)((double)fx[s] * kFixInv);
    oy[i] += (float)((double)fy[s] * kFixInv);
    oz[i] += (float)((double)fz[s] * kFixInv);
  } else {
    ox[i] += (float)fxd[s];
    oy[i] += (float)fyd[s];
    oz[i] += (float)fzd[s];
  }
}

void launch_add_forces_f32(int n, const uint32_t *perm, const int64_t *fx,
                           const int64_t *fy, const int64_t *fz,
                           const double *fxd, const double *fyd,
                           const double *fzd, float *ox, float *oy, float *oz,
                           hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_forces_f32, dim3((n + 255) / 256), dim3(256), 0, st,
                     n, perm, fx, fy, fz, fxd, fyd, fzd, ox, oy, oz);
}

__global__ __launch_bounds__(256) void k_export_atom_stats(int n,
                                                           const uint32_t *perm,
                                                           const double *src,
                                                           double *dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = (size_t)perm[s] * 4;
  for (int c = 0; c < 4; ++c) dst[i + c] = src[(size_t)s * 4 + c];
}

void launch_export_atom_stats(int n, const uint32_t *perm, const double *src,
                              double *dst, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_export_atom_stats, dim3((n + 255) / 256), dim3(256), 0,
                     st, n, perm, src, dst);
}

}  // namespace sre
