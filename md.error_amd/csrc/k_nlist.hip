// k_nlist.hip -- Verlet neighbor-list build over the cell-sorted atoms.
//
// Pair set (H/src/NeighborList.cu:327-420, all-pairs build of the adaptive
// engine; U/src/NeighborList_interface.cu:361-491 for the uniform one):
//     j in L(i)  <=>  j != i  and  |d_ij|^2 < (rc_i + skin)^2
// with the per-atom radius rc_i + skin of atom i (the list is asymmetric).
//
// Candidates: a group of TB lanes per atom.  The (x, y) cell columns that
// intersect the atom's sphere are taken kBuildCols * TB at a time ("chunk");
// lane l of the group sets up kBuildCols of them: the z cells covering the
// sphere's chord in that column are one or two contiguous runs of
// cell-sorted atoms, whose bounds come from the cell start array (all the
// lane's loads are independent, so they are in flight together).  A prefix
// sum over the group turns the chunk's runs into one flat candidate sequence
// held in LDS; the group then streams it kBuildUnroll * TB candidates per
// round (every lane issues its kBuildUnroll gathers before testing any), so
// short runs cost no partially filled batches.  The exact predicate is
// applied to every candidate, so the windows only need to be conservative.
// Each round's accepted neighbours are placed by an exclusive prefix over the
// group's lanes (deterministic (lane, unroll) order) and written to the
// wave-tiled column-major layout the force kernel reads (list_slot()).
// Writes beyond the capacity are dropped but still counted; the host regrows
// and rebuilds.
#include "sre_internal.h"

namespace sre {

constexpr int kBuildCols = 4;     // columns per lane per chunk
constexpr int kBuildUnroll = 6;   // candidates per lane per round

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

template <int TB>
__global__ __launch_bounds__(256) void k_build_list(
    int n, int TF, const uint4 *__restrict__ pos4,
    const float *__restrict__ rcut, const RepInfo *__restrict__ reps,
    const int *__restrict__ start, float skin, int cap, int *__restrict__ list,
    int *__restrict__ nnb, int2 *__restrict__ part) {
  constexpr int A = kWave / TB;
  constexpr int CC = kBuildCols * TB;   // columns per chunk
  constexpr int NS = 2 * CC;            // runs per chunk
  constexpr int U = kBuildUnroll;
  __shared__ int s_p0[256 / TB][NS];    // first sorted atom of each run
  __shared__ int s_end[256 / TB][NS];   // inclusive prefix of run lengths
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / TB, l = lane % TB;
  const int gb = threadIdx.x / TB;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int s = wave * A + g;
  const bool act = s < n;
  const int si = act ? s : n - 1;

  const uint4 pi = pos4[si];
  const __amdgpu_buffer_rsrc_t prs = pos_rsrc_idx(pos4, n);
  const RepInfo &R = reps[pi.w >> 8];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  const int ncx = R.nc[0], ncy = R.nc[1], ncz = R.nc[2];
  const float hx = R.h[0], hy = R.h[1], hz = R.h[2];
  const float ihz = 1.f / hz;
  const float rl = rcut[si] + skin;
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
  const int noy = oy1 - oy0 + 1;
  const int ncol = act ? (ox1 - ox0 + 1) * noy : 0;
  const int ncol_w = __builtin_amdgcn_readfirstlane(wave_max_i(ncol));

  // list slots of this atom in the force kernel's layout (T = TF lanes):
  // slot(k) = sbase + (k / TF) * 64 + k % TF, TF a power of two
  const int AF = kWave / TF;
  const int lt = __builtin_ctz((unsigned)TF);
  const size_t sbase =
      (size_t)(si / AF) * (size_t)cap * (size_t)AF + (size_t)((si % AF) * TF);
  int *const lrow = list + sbase;
  int *const sp0 = s_p0[gb];
  int *const send = s_end[gb];
  int count = 0;
  for (int cbase = 0; cbase < ncol_w; cbase += CC) {
    // runs of this lane's kBuildCols columns -> (first atom, length) in LDS
    const int slot0 = l * 2 * kBuildCols;
    int tot = 0;
    // this lane's first column; the following ones step oy, then ox
    int c = cbase + l * kBuildCols;
    int ox = ox0 + c / noy, oy = oy0 + c % noy;
#pragma unroll 2
    for (int m = 0; m < kBuildCols; ++m, ++c) {
      int ra0 = 0, rb0 = -1, ra1 = 0, rb1 = -1, colbase = 0;
      if (c < ncol) {
        float gx = 0.f, gy = 0.f;
        if (!fullx) gx = ox > 0 ? ox * hx - fx : (ox < 0 ? fx + (-ox - 1) * hx : 0.f);
        if (!fully) gy = oy > 0 ? oy * hy - fy : (oy < 0 ? fy + (-oy - 1) * hy : 0.f);
        gx = fmaxf(gx, 0.f);
        gy = fmaxf(gy, 0.f);
        const float gxy2 = gx * gx + gy * gy;
        if (gxy2 < rlm2) {
          int cxw = cx + ox;
          cxw += cxw < 0 ? ncx : 0;
          cxw -= cxw >= ncx ? ncx : 0;
          int cyw = cy + oy;
          cyw += cyw < 0 ? ncy : 0;
          cyw -= cyw >= ncy ? ncy : 0;
          colbase = R.cell_off + (cxw * ncy + cyw) * ncz;
          rb0 = ncz - 1;  // whole column unless the chord is shorter
          if (!fullz) {
            // (the window radius carries a 1e-4 margin, far above the
            // rounding of a multiplication by 1/hz)
            const float dz = sqrtf(fmaxf(rlm2 - gxy2, 0.f));
            const int lo = cz + (int)floorf((fz - dz) * ihz);
            const int hi = cz + (int)floorf((fz + dz) * ihz);
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
        }
      }
      const int b0 = rb0 >= ra0 ? start[colbase + ra0] : 0;
      const int e0 = rb0 >= ra0 ? start[colbase + rb0 + 1] : 0;
      const int b1 = rb1 >= ra1 ? start[colbase + ra1] : 0;
      const int e1 = rb1 >= ra1 ? start[colbase + rb1 + 1] : 0;
      // runs lie inside [0, n) by construction; the clamps keep every gather
      // in bounds whatever the cell offsets hold
      const int q0 = min(max(b0, 0), n), q1 = min(max(b1, 0), n);
      const int n0 = max(min(e0, n) - q0, 0), n1 = max(min(e1, n) - q1, 0);
      sp0[slot0 + 2 * m] = q0;
      send[slot0 + 2 * m] = n0;       // length for now, prefix below
      sp0[slot0 + 2 * m + 1] = q1;
      send[slot0 + 2 * m + 1] = n1;
      if (++oy > oy1) {
        oy = oy0;
        ++ox;
      }
      tot += n0 + n1;
    }
    // flat sequence: prefix over the group (lane-major slot order)
    int incl = tot;
#pragma unroll
    for (int o = 1; o < TB; o <<= 1) {
      const int t = __shfl_up(incl, o, TB);
      if (l >= o) incl += t;
    }
    const int T = __shfl(incl, TB - 1, TB);
    int run = incl - tot;
    for (int r = 0; r < 2 * kBuildCols; ++r) {
      run += send[slot0 + r];
      send[slot0 + r] = run;
    }
    __builtin_amdgcn_wave_barrier();
    // stream the flat sequence: in round k0 lane l takes the U consecutive
    // candidates v = (k0 TB + l) U + u
    const int rounds = __builtin_amdgcn_readfirstlane(wave_max_i((T + TB * U - 1) / (TB * U)));
    // cursor over the runs: run cur covers flat indices [send[cur-1],
    // end) and candidate v of it is sorted atom off + v (kept in registers;
    // LDS is read only when the cursor moves to the next run)
    int cur = 0, end = send[0], off = sp0[0];
    // candidates of round k0 and their gathers (a lane past the sequence
    // takes its own atom si, which the self test below rejects)
    auto load = [&](int k0, int *jj, uint4 *cj) {
      const int v0 = (k0 * TB + l) * U;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int v = v0 + u;
        int j = si;
        if (v < T) {
          while (v >= end) {
            ++cur;
            const int b = end;
            end = send[cur];
            off = sp0[cur] - b;
          }
          j = off + v;
        }
        jj[u] = j;
        cj[u] = gather_pos_idx<false>(prs, j);
      }
      asm volatile("" ::: "memory");
    };
    // the accepted candidates of a round, placed in flat-sequence order
    // (lane by lane, each lane's U candidates in turn: an exclusive prefix
    // of the lanes' accepted counts over the group), so that consecutive
    // list entries are consecutive candidates -- neighbouring atoms of one
    // cell column -- and the force kernel's lanes of one atom gather from
    // the same few cache lines
    auto accept = [&](const int *jj, const uint4 *cj) {
      bool ok[U];
      int c = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float dx = disp(cj[u].x, pi.x, qx);
        const float dy = disp(cj[u].y, pi.y, qy);
        const float dz2 = disp(cj[u].z, pi.z, qz);
        // (an inactive group, s >= n, has no candidates: all its lanes
        // hold j = si)
        ok[u] = (jj[u] != si) & (dist2(dx, dy, dz2) < rl2);
        c += ok[u] ? 1 : 0;
      }
      int incl = c;
#pragma unroll
      for (int o = 1; o < TB; o <<= 1) {
        const int t = __shfl_up(incl, o, TB);
        if (l >= o) incl += t;
      }
      int k = count + incl - c;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (ok[u] && k < cap) lrow[(unsigned)(((k >> lt) << 6) + (k & (TF - 1)))] = jj[u];
        k += ok[u] ? 1 : 0;
      }
      count += __shfl(incl, TB - 1, TB);
    };
    // software pipeline: the gathers of round k0 + 1 are in flight while
    // round k0 is tested; whole pairs of rounds in the loop (no branch in its
    // body, so each gather is waited for only where it is used), the odd
    // last round after it (every branch is wave-uniform)
    int ja[U], jb[U];
    uint4 ca[U], cb[U];
    load(0, ja, ca);
    int k0 = 0;
    for (; k0 + 1 < rounds; k0 += 2) {
      load(k0 + 1, jb, cb);
      accept(ja, ca);
      load(k0 + 2, ja, ca);
      accept(jb, cb);
    }
    if (k0 < rounds) accept(ja, ca);
    __builtin_amdgcn_wave_barrier();  // LDS reuse by the next chunk
  }
  // an overflowing list keeps its first cap entries (the partials below carry
  // the full count, from which the host regrows), so no reader of the list
  // ever runs past its rows
  if (act && l == 0) nnb[s] = min(count, cap);
  // this wave's longest list and total list length
  int m = act ? count : 0;
  int t = (act && l == 0) ? count : 0;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    m = max(m, __shfl_xor(m, o, kWave));
    t += __shfl_xor(t, o, kWave);
  }
  if (lane == 0) part[wave] = make_int2(m, t);
}

int build_part_count(int mode, int n, int lanes, int n_cols) {
  if (n <= 0) return 0;
  if (mode == 1) return (n + kWave - 1) / kWave + n_cols;  // tile grid
  const int A = kWave / lanes;
  const int waves = (n + A - 1) / A;
  return (waves * kWave + 255) / 256 * (256 / kWave);     // whole 256 blocks
}

void launch_build_list(int lanes, int force_lanes, int n, const uint4 *pos4,
                       const float *rcut, const RepInfo *reps, const int *start,
                       float skin, int cap, int *list, int *nnb, int2 *part,
                       hipStream_t st) {
  if (n <= 0) return;
  const int A = kWave / lanes;
  const int waves = (n + A - 1) / A;
  const int block = 256;
  const int grid = (waves * kWave + block - 1) / block;
#define SRE_BUILD(TB)                                                        \
  hipLaunchKernelGGL(k_build_list<TB>, dim3(grid), dim3(block), 0, st, n,   \
                     force_lanes, pos4, rcut, reps, start, skin, cap, list, \
                     nnb, part)
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

// ------------------------------------------------------------ tile build --
//
// One wave per tile: up to 64 consecutive cell-sorted atoms of one (x, y)
// cell column.  The tile's candidate region is the union of its atoms'
// spheres: the columns within (rl_max + margin) of the tile's column and, in
// each, the z cells from the tile's lowest cell - kz to its highest + kz --
// one or two contiguous runs of sorted atoms.  Runs are staged 64 atoms at a
// time in LDS and every lane tests its own atom against each staged candidate
// (a broadcast LDS read), so all loops are wave-uniform and the only
// divergence is the list write of an accepted pair.  The predicate is the
// same exact one as above; entries are written in candidate order.

__device__ __forceinline__ int find_replica_by_col(const RepInfo *reps, int n_rep,
                                                   int c) {
  int lo = 0, hi = n_rep - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reps[mid].col_off <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_col_tiles(int n_cols, int n_rep,
                                                   const RepInfo *reps,
                                                   const int *start, int tile,
                                                   uint32_t *tiles_per_col) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  const RepInfo &R = reps[find_replica_by_col(reps, n_rep, c)];
  const int base = R.cell_off + (c - R.col_off) * R.nc[2];
  const int cnt = start[base + R.nc[2]] - start[base];
  tiles_per_col[c] = (uint32_t)((cnt + tile - 1) / tile);
}

__global__ __launch_bounds__(kWave) void k_build_tiles(
    int n_cols, int n_rep, int TF, const uint32_t *__restrict__ tile_off,
    const uint32_t *__restrict__ tiles_per_col, const uint4 *__restrict__ pos4,
    const float *__restrict__ rcut, const RepInfo *__restrict__ reps,
    const int *__restrict__ start, float skin, int cap, int *__restrict__ list,
    int *__restrict__ nnb, int2 *__restrict__ part) {
  __shared__ uint4 cand[kWave];
  const int t = blockIdx.x;
  const int n_tiles = (int)(tile_off[n_cols - 1] + tiles_per_col[n_cols - 1]);
  if (t >= n_tiles) {  // the grid is an upper bound of the tile count
    if (threadIdx.x == 0) part[t] = make_int2(0, 0);
    return;
  }
  // column of this tile: last c with tile_off[c] <= t
  int lo = 0, hi = n_cols - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int)tile_off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int c = lo;
  const RepInfo &R = reps[find_replica_by_col(reps, n_rep, c)];
  const int ncx = R.nc[0], ncy = R.nc[1], ncz = R.nc[2];
  const float hx = R.h[0], hy = R.h[1], hz = R.h[2];
  const float qx = R.q[0], qy = R.q[1], qz = R.q[2];
  const int col = c - R.col_off;
  const int cx = col / ncy, cy = col - cx * ncy;
  const int cbase = R.cell_off + col * ncz;
  const int a0 = start[cbase] + (t - (int)tile_off[c]) * kWave;
  const int a1 = min(start[cbase + ncz], a0 + kWave);
  const int lane = threadIdx.x;
  const int s = a0 + lane;
  const bool act = s < a1;
  const uint4 pi = pos4[act ? s : a0];
  const float rl = rcut[act ? s : a0] + skin;
  const float rl2 = rl * rl;
  float rlmax = act ? rl : 0.f;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
    rlmax = fmaxf(rlmax, __shfl_xor(rlmax, o, kWave));
  const float rlm = rlmax * 1.0001f + 1e-4f;
  const float rlm2 = rlm * rlm;
  const int cz_lo = cell_of(pos4[a0].z, ncz), cz_hi = cell_of(pos4[a1 - 1].z, ncz);
  const int kx = (int)ceilf(rlm / hx), ky = (int)ceilf(rlm / hy),
            kz = (int)ceilf(rlm / hz);
  const bool fullx = 2 * kx + 1 >= ncx, fully = 2 * ky + 1 >= ncy;
  const int ox0 = fullx ? -cx : -kx, ox1 = fullx ? ncx - 1 - cx : kx;
  const int oy0 = fully ? -cy : -ky, oy1 = fully ? ncy - 1 - cy : ky;
  // list slots of this lane's atom (layout of the force kernel, T = TF)
  const int A = kWave / TF;
  const int lt = __builtin_ctz((unsigned)TF);
  const size_t sbase = act ? (size_t)(s / A) * (size_t)cap * A + (size_t)((s % A) * TF) : 0;
  int count = 0;
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
        const int p0 = start[colbase + ca], p1 = start[colbase + cb + 1];
        for (int p = p0; p < p1; p += kWave) {
          const int q = p + lane;
          cand[lane] = pos4[q < p1 ? q : p0];
          __syncthreads();
          const int m = min(kWave, p1 - p);
          for (int k = 0; k < m; ++k) {
            const uint4 cj = cand[k];
            const float d2 = dist2(disp(cj.x, pi.x, qx), disp(cj.y, pi.y, qy),
                                   disp(cj.z, pi.z, qz));
            const int j = p + k;
            if (act && j != s && d2 < rl2) {
              if (count < cap)
                list[sbase + ((size_t)(count >> lt) << 6) + (count & (TF - 1))] = j;
              ++count;
            }
          }
          __syncthreads();
        }
      }
    }
  }
  if (act) nnb[s] = min(count, cap);   // as in k_build_list
  int mx = count, tot = count;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mx = max(mx, __shfl_xor(mx, o, kWave));
    tot += __shfl_xor(tot, o, kWave);
  }
  if (lane == 0) part[t] = make_int2(mx, tot);
}

void launch_tile_map(int n_cols, int n_rep, const RepInfo *reps, const int *start,
                     int tile, uint32_t *tiles_per_col, uint32_t *tile_off,
                     uint32_t *scan_tmp, hipStream_t st) {
  if (n_cols <= 0) return;
  hipLaunchKernelGGL(k_col_tiles, dim3((n_cols + 255) / 256), dim3(256), 0, st,
                     n_cols, n_rep, reps, start, tile, tiles_per_col);
  device_exscan_u32((size_t)n_cols, tiles_per_col, tile_off, scan_tmp, st);
}

void launch_build_tiles(int n, int n_cols, int n_rep, int force_lanes,
                        const uint4 *pos4, const float *rcut,
                        const RepInfo *reps, const int *start, float skin,
                        int cap, int *list, int *nnb, int2 *part,
                        uint32_t *tiles_per_col, uint32_t *tile_off,
                        uint32_t *scan_tmp, hipStream_t st) {
  if (n <= 0 || n_cols <= 0) return;
  launch_tile_map(n_cols, n_rep, reps, start, kWave, tiles_per_col, tile_off, scan_tmp,
                  st);
  // every column adds at most one partial tile
  const int max_tiles = (n + kWave - 1) / kWave + n_cols;
  hipLaunchKernelGGL(k_build_tiles, dim3(max_tiles), dim3(kWave), 0, st, n_cols,
                     n_rep, force_lanes, tile_off, tiles_per_col, pos4, rcut,
                     reps, start, skin, cap, list, nnb, part);
}

}  // namespace sre
