// k_nlist.hip -- Verlet neighbor-list build over the cell-sorted atoms.
//
// Pair set (H/src/NeighborList.cu:327-420, all-pairs build of the adaptive
// engine; U/src/NeighborList_interface.cu:361-491 for the uniform one):
//     j in L(i)  <=>  j != i  and  |d_ij|^2 < (rc_i + skin)^2
// with the per-atom radius rc_i + skin of atom i (the list is asymmetric).
//
// Candidates: a group of TB lanes per atom walks the cell columns (x, y) that
// intersect its sphere; in each column the cells covering the sphere's chord in
// z are one or two contiguous runs of cell-sorted atoms, streamed TB at a time.
// The exact predicate is applied to every candidate, so the candidate windows
// only need to be conservative.  Accepted neighbours are compacted with a
// group-wide ballot into a deterministic order and written to the wave-tiled
// column-major layout the force kernel reads (list_slot()).  Writes beyond the
// capacity are dropped but still counted; the host regrows and rebuilds.
#include "sre_internal.h"

namespace sre {

template <int TB>
__global__ __launch_bounds__(256) void k_build_list(
    int n, int TF, const uint4 *__restrict__ pos4,
    const float *__restrict__ rcut, const RepInfo *__restrict__ reps,
    const int *__restrict__ start, float skin, int cap, int *__restrict__ list,
    int *__restrict__ nnb, unsigned long long *counters) {
  constexpr int A = kWave / TB;
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / TB, l = lane % TB;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + g;
  const int shift = g * TB;
  const uint64_t gmask = (TB == 64) ? ~0ull : ((1ull << TB) - 1ull);
  const uint64_t below = (1ull << l) - 1ull;
  int count = 0;
  if (s < n) {
    const uint4 pi = pos4[s];
    const RepInfo &R = reps[pi.w >> 8];
    const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
    const int ncx = R.nc[0], ncy = R.nc[1], ncz = R.nc[2];
    const float hx = R.h[0], hy = R.h[1], hz = R.h[2];
    const float rl = rcut[s] + skin;
    const float rl2 = rl * rl;
    const float rlm = rl * 1.0001f + 1e-4f;  // conservative window radius
    const float rlm2 = rlm * rlm;
    const int cx = cell_of(pi.x, ncx), cy = cell_of(pi.y, ncy),
              cz = cell_of(pi.z, ncz);
    // position inside the own cell from the low 32 bits of u * nc
    const float fx = (float)(uint32_t)((uint64_t)pi.x * (uint64_t)ncx) * (hx * 0x1p-32f);
    const float fy = (float)(uint32_t)((uint64_t)pi.y * (uint64_t)ncy) * (hy * 0x1p-32f);
    const float fz = (float)(uint32_t)((uint64_t)pi.z * (uint64_t)ncz) * (hz * 0x1p-32f);
    const int kx = (int)ceilf(rlm / hx), ky = (int)ceilf(rlm / hy),
              kz = (int)ceilf(rlm / hz);
    const bool fullx = 2 * kx + 1 >= ncx, fully = 2 * ky + 1 >= ncy,
               fullz = 2 * kz + 1 >= ncz;
    const int ox0 = fullx ? -cx : -kx, ox1 = fullx ? ncx - 1 - cx : kx;
    const int oy0 = fully ? -cy : -ky, oy1 = fully ? ncy - 1 - cy : ky;
    for (int ox = ox0; ox <= ox1; ++ox) {
      float gx = 0.f;
      if (!fullx) gx = ox > 0 ? ox * hx - fx : (ox < 0 ? fx + (-ox - 1) * hx : 0.f);
      gx = fmaxf(gx, 0.f);
      if (gx * gx >= rlm2) continue;
      int cxw = cx + ox;
      cxw += cxw < 0 ? ncx : 0;
      cxw -= cxw >= ncx ? ncx : 0;
      for (int oy = oy0; oy <= oy1; ++oy) {
        float gy = 0.f;
        if (!fully) gy = oy > 0 ? oy * hy - fy : (oy < 0 ? fy + (-oy - 1) * hy : 0.f);
        gy = fmaxf(gy, 0.f);
        const float gxy2 = gx * gx + gy * gy;
        if (gxy2 >= rlm2) continue;
        int cyw = cy + oy;
        cyw += cyw < 0 ? ncy : 0;
        cyw -= cyw >= ncy ? ncy : 0;
        const int colbase = R.cell_off + (cxw * ncy + cyw) * ncz;
        int ra0 = 0, rb0 = ncz - 1, ra1 = 0, rb1 = -1;
        if (!fullz) {
          const float dz = sqrtf(fmaxf(rlm2 - gxy2, 0.f));
          const int lo = cz + (int)floorf((fz - dz) / hz);
          const int hi = cz + (int)floorf((fz + dz) / hz);
          if (hi - lo + 1 < ncz) {
            if (lo < 0) {
              ra0 = lo + ncz; rb0 = ncz - 1; ra1 = 0; rb1 = hi;
            } else if (hi >= ncz) {
              ra0 = lo; rb0 = ncz - 1; ra1 = 0; rb1 = hi - ncz;
            } else {
              ra0 = lo; rb0 = hi;
            }
          }
        }
        for (int run = 0; run < 2; ++run) {
          const int ca = run ? ra1 : ra0, cb = run ? rb1 : rb0;
          if (cb < ca) continue;
          const int p0 = start[colbase + ca], p1 = start[colbase + cb + 1];
          for (int p = p0; p < p1; p += TB) {
            const int pj = p + l;
            bool ok = false;
            if (pj < p1 && pj != s) {
              const uint4 pj4 = pos4[pj];
              const float dx = disp(pj4.x, pi.x, qx);
              const float dy = disp(pj4.y, pi.y, qy);
              const float dz2 = disp(pj4.z, pi.z, qz);
              ok = dist2(dx, dy, dz2) < rl2;
            }
            const uint64_t gb = (__ballot(ok) >> shift) & gmask;
            if (ok) {
              const int k = count + __popcll(gb & below);
              if (k < cap) list[list_slot(s, k, TF, cap)] = pj;
            }
            count += __popcll(gb);
          }
        }
      }
    }
    if (l == 0) nnb[s] = count;
  }
  // counters[0]: longest list, counters[1]: total list length (integer
  // atomics, one per wave: order independent)
  int m = count;
  int t = (l == 0) ? count : 0;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    m = max(m, __shfl_xor(m, o, kWave));
    t += __shfl_xor(t, o, kWave);
  }
  if (lane == 0 && m > 0) {
    atomicMax(&counters[0], (unsigned long long)m);
    atomicAdd(&counters[1], (unsigned long long)t);
  }
}

void launch_build_list(int lanes, int force_lanes, int n, const uint4 *pos4,
                       const float *rcut, const RepInfo *reps, const int *start,
                       float skin, int cap, int *list, int *nnb,
                       unsigned long long *counters, hipStream_t st) {
  if (n <= 0) return;
  const int A = kWave / lanes;
  const int waves = (n + A - 1) / A;
  const int block = 256;
  const int grid = (waves * kWave + block - 1) / block;
#define SRE_BUILD(TB)                                                        \
  hipLaunchKernelGGL(k_build_list<TB>, dim3(grid), dim3(block), 0, st, n,   \
                     force_lanes, pos4, rcut, reps, start, skin, cap, list, \
                     nnb, counters)
  switch (lanes) {
    case 1: SRE_BUILD(1); break;
    case 2: SRE_BUILD(2); break;
    case 4: SRE_BUILD(4); break;
    case 8: SRE_BUILD(8); break;
    case 16: SRE_BUILD(16); break;
    case 32: SRE_BUILD(32); break;
    default: SRE_BUILD(64); break;
  }
#undef SRE_BUILD
}

}  // namespace sre
