// k_sort.hip -- cell keys, binning of the atoms by cell, cell offsets, and
// the gather of per-atom data into cell order.
//
// The reference bins atoms with fixed-capacity cells and atomicInc slot
// insertion (H/src/NeighborList.cu:84-167), whose in-cell order depends on
// scheduling.  Here the cell order is always the stable order (cell, caller
// index), a pure function of the input, produced two ways:
//   * a counting sort by cell (the rebuild's default, described below at
//     k_cell_count), whose atomics only place atoms in a cell's slots before
//     a rank by caller index fixes their order;
//   * an atomic-free LSD radix sort of the cell key over digits of at most 8
//     bits (the allocation-free path and a tuning option), each pass built
//     from (1) a per-block stable ranking of 1024 keys: lanes holding the
//     same digit find each other with one ballot per digit bit (their rank is
//     the number of such lanes below them), per-wave digit counts go through
//     LDS, and the keys are written in (digit, original position) order;
//     (2) a digit-major exclusive scan of the per-block digit counts; (3) a
//     stable scatter to the global positions.
#include "sre_internal.h"

namespace sre {

constexpr int kSortE = 1024;      // elements per sort block (one per thread)
constexpr int kSortT = 1024;      // threads per sort block
constexpr int kMaxDigitBits = 8;
constexpr int kMaxPasses = 4;
constexpr int kScanTile = 4096;   // elements per scan tile (1024 threads x 4)
// up to this many sort blocks the scatter derives its offsets itself (each
// block reads all nblocks x 2^bits counts) instead of a separate scan
constexpr int kFusedScanBlocks = 128;

// ------------------------------------------------------------------ keys --


// ------------------------------------------------------- block scan 1024 --

// Exclusive scan of one value per thread over a 1024-thread block.
__device__ uint32_t block_exscan_1024(uint32_t v, uint32_t *tmp,
                                      uint32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < kSortT / 64; ++i) {
      uint32_t t = tmp[i];
      tmp[i] = s;
      s += t;
    }
    tmp[kSortT / 64] = s;
  }
  __syncthreads();
  uint32_t r = x - v + tmp[w];
  if (total) *total = tmp[kSortT / 64];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ rank block --

// Stable local ranking of one block of keys by digit (key >> shift) & mask:
// writes the block's keys/values to tkey/tval in (digit, position) order and
// the block's digit counts digit-major to counts[d * nblocks + b].  The first
// pass (pos != NULL) computes the cell keys from the positions (value = atom
// index) instead of reading key/val.
__global__ __launch_bounds__(kSortT) void k_sort_block(
    int n, int shift, int bits, int nblocks, const uint32_t *key,
    const uint32_t *val, const float *pos, int stride, const int *rep_of,
    const RepInfo *reps, uint32_t *tkey, uint32_t *tval, uint32_t *counts) {
  constexpr int W = kSortT / 64;
  __shared__ uint32_t wcnt[W][1 << kMaxDigitBits];  // per wave, then prefix
  __shared__ uint32_t dstart[1 << kMaxDigitBits];
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  const int b = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int base = b * kSortE;
  const bool valid = base + t < n;
  const uint32_t R = 1u << bits, mask = R - 1u;
  uint32_t k = 0u, v = 0u;
  if (valid) {
    if (pos) {
      k = cell_key(pos, stride, rep_of, reps, base + t);
      v = (uint32_t)(base + t);
    } else {
      k = key[base + t];
      v = val[base + t];
    }
  }
  const uint32_t d = (k >> shift) & mask;
  for (int x = t; x < W * (1 << kMaxDigitBits); x += kSortT) (&wcnt[0][0])[x] = 0u;
  // lanes of this wave with the same digit
  uint64_t peers = __ballot(valid);
  for (int bit = 0; bit < bits; ++bit) {
    const bool on = (d >> bit) & 1u;
    const uint64_t bal = __ballot(on);
    peers &= on ? bal : ~bal;
  }
  const int rank = __popcll(peers & ((1ull << lane) - 1ull));
  __syncthreads();
  if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
  __syncthreads();
  uint32_t bc = 0;
  if ((uint32_t)t < R) {
    for (int ww = 0; ww < W; ++ww) {
      const uint32_t c = wcnt[ww][t];
      wcnt[ww][t] = bc;
      bc += c;
    }
    counts[(size_t)t * nblocks + b] = bc;
  }
  const uint32_t ds = block_exscan_1024((uint32_t)t < R ? bc : 0u, tmp, nullptr);
  if ((uint32_t)t < R) dstart[t] = ds;
  __syncthreads();
  if (valid) {
    const uint32_t pos = dstart[d] + wcnt[w][d] + (uint32_t)rank;
    tkey[base + pos] = k;
    tval[base + pos] = v;
  }
}

// Stable scatter of block b's locally ranked keys.  Global offsets come from
// the digit-major exclusive scan (offsets), or, when offsets is NULL (few
// blocks), are computed here: every block sums each digit's row of counts
// over all blocks and over the blocks before it (one wave per digit row,
// coalesced, fixed-order wave reduction) and scans the digit totals.
__global__ __launch_bounds__(kSortT) void k_sort_scatter(
    int n, int shift, int bits, int nblocks, const uint32_t *tkey,
    const uint32_t *tval, const uint32_t *counts, const uint32_t *offsets,
    uint32_t *okey, uint32_t *oval) {
  __shared__ uint32_t lstart[1 << kMaxDigitBits];
  __shared__ uint32_t goff[1 << kMaxDigitBits];
  __shared__ uint32_t s_pre[1 << kMaxDigitBits];
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  const int b = blockIdx.x;
  const int base = b * kSortE;
  const int nv = min(kSortE, n - base);
  const int R = 1 << bits;
  const uint32_t mask = (uint32_t)R - 1u;
  if (!offsets) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = w; d < R; d += kSortT / 64) {
      uint32_t tot = 0, pre = 0;
      for (int bb = lane; bb < nblocks; bb += 64) {
        const uint32_t c = counts[(size_t)d * nblocks + bb];
        tot += c;
        pre += bb < b ? c : 0u;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o, 64);
        pre += __shfl_xor(pre, o, 64);
      }
      if (lane == 0) {
        goff[d] = tot;
        s_pre[d] = pre;
      }
    }
    __syncthreads();
    const uint32_t tot = (int)threadIdx.x < R ? goff[threadIdx.x] : 0u;
    const uint32_t ds = block_exscan_1024(tot, tmp, nullptr);
    if ((int)threadIdx.x < R) goff[threadIdx.x] = ds + s_pre[threadIdx.x];
  }
  const int per = (R + kSortT - 1) / kSortT;
  const int lo = min(R, (int)threadIdx.x * per), hi = min(R, lo + per);
  uint32_t local = 0;
  for (int d = lo; d < hi; ++d) local += counts[(size_t)d * nblocks + b];
  uint32_t run = block_exscan_1024(local, tmp, nullptr);
  for (int d = lo; d < hi; ++d) {
    lstart[d] = run;
    run += counts[(size_t)d * nblocks + b];
    if (offsets) goff[d] = offsets[(size_t)d * nblocks + b];
  }
  __syncthreads();
  for (int p = threadIdx.x; p < nv; p += kSortT) {
    uint32_t k = tkey[base + p];
    uint32_t d = (k >> shift) & mask;
    uint32_t dst = goff[d] + ((uint32_t)p - lstart[d]);
    okey[dst] = k;
    oval[dst] = tval[base + p];
  }
}

// ------------------------------------------------------------ device scan --

__global__ __launch_bounds__(kSortT) void k_scan_reduce(size_t m,
                                                        const uint32_t *in,
                                                        uint32_t *sums) {
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t v = 0;
  for (int k = 0; k < 4; ++k) {
    size_t idx = base + (size_t)threadIdx.x * 4 + k;
    if (idx < m) v += in[idx];
  }
  uint32_t total;
  block_exscan_1024(v, tmp, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// single-block exclusive scan of m elements, kScanTile at a time with a
// running carry (in place allowed: a tile is read before it is written)
__global__ __launch_bounds__(kSortT) void k_scan_single(size_t m,
                                                        const uint32_t *in,
                                                        uint32_t *out) {
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  uint32_t carry = 0;
  for (size_t base = 0; base < m; base += kScanTile) {
    const size_t i0 = base + (size_t)threadIdx.x * 4;
    uint32_t v[4];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = (i0 + k < m) ? in[i0 + k] : 0u;
      local += v[k];
    }
    uint32_t total;
    uint32_t run = block_exscan_1024(local, tmp, &total) + carry;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < m) out[i0 + k] = run;
      run += v[k];
    }
    carry += total;
  }
}

__global__ __launch_bounds__(kSortT) void k_scan_down(size_t m,
                                                      const uint32_t *in,
                                                      const uint32_t *sums,
                                                      uint32_t *out) {
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * 4;
  uint32_t v[4];
  uint32_t local = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < m) ? in[base + k] : 0u;
    local += v[k];
  }
  uint32_t run = block_exscan_1024(local, tmp, nullptr) + sums[blockIdx.x];
  for (int k = 0; k < 4; ++k) {
    if (base + k < m) out[base + k] = run;
    run += v[k];
  }
}

void device_exscan_u32(size_t m, const uint32_t *in, uint32_t *out,
                       uint32_t *tmp, hipStream_t st) {
  if (m <= 4 * (size_t)kScanTile) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kSortT), 0, st, m, in,
                       out);
    return;
  }
  size_t nt = (m + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(kSortT), 0, st,
                     m, in, tmp);
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kSortT), 0, st, nt, tmp,
                     tmp);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nt), dim3(kSortT), 0, st, m,
                     in, tmp, out);
}

static void digit_plan(int key_bits, int *np, int bits[kMaxPasses]) {
  if (key_bits < 1) key_bits = 1;
  int passes = (key_bits + kMaxDigitBits - 1) / kMaxDigitBits;
  int rem = key_bits;
  for (int p = 0; p < passes; ++p) {
    bits[p] = (rem + (passes - p) - 1) / (passes - p);
    rem -= bits[p];
  }
  *np = passes;
}

size_t radix_counts_capacity(int n, int key_bits) {
  int np, bits[kMaxPasses];
  digit_plan(key_bits, &np, bits);
  int mb = 0;
  for (int p = 0; p < np; ++p) mb = bits[p] > mb ? bits[p] : mb;
  size_t nblocks = (size_t)(n + kSortE - 1) / kSortE;
  return ((size_t)1 << mb) * (nblocks > 0 ? nblocks : 1);
}

int radix_passes(int key_bits) {
  int np, bits[kMaxPasses];
  digit_plan(key_bits, &np, bits);
  return np;
}

int launch_radix_sort(int n, int key_bits, const float *pos, int stride,
                      const int *rep_of, const RepInfo *reps, const SortBufs &b,
                      hipStream_t st) {
  if (n <= 0) return 0;
  if (key_bits > kMaxPasses * kMaxDigitBits) return -1;
  int np, bits[kMaxPasses];
  digit_plan(key_bits, &np, bits);
  const int nblocks = (n + kSortE - 1) / kSortE;
  // pass p writes key_a/val_a when np - 1 - p is even, so the last pass
  // always ends there; pass 0 reads the positions
  const bool odd = (np & 1) != 0;
  uint32_t *kin = nullptr, *vin = nullptr;
  uint32_t *kout = odd ? b.key_a : b.key_b, *vout = odd ? b.val_a : b.val_b;
  int shift = 0;
  for (int p = 0; p < np; ++p) {
    const size_t m = ((size_t)1 << bits[p]) * nblocks;
    hipLaunchKernelGGL(k_sort_block, dim3(nblocks), dim3(kSortT), 0, st, n,
                       shift, bits[p], nblocks, kin, vin, p == 0 ? pos : nullptr,
                       stride, rep_of, reps, b.tkey, b.tval, b.counts);
    // few blocks: each scatter block derives its offsets from the counts
    const bool fused = nblocks <= kFusedScanBlocks;
    if (!fused) device_exscan_u32(m, b.counts, b.offsets, b.scan_tmp, st);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(kSortT), 0, st, n,
                       shift, bits[p], nblocks, b.tkey, b.tval, b.counts,
                       fused ? nullptr : b.offsets, kout, vout);
    shift += bits[p];
    kin = kout;
    vin = vout;
    kout = (kout == b.key_a) ? b.key_b : b.key_a;
    vout = (vout == b.val_a) ? b.val_b : b.val_a;
  }
  return 0;
}

// ------------------------------------------------------------- cell start --

// Atom i into sorted slot d: lattice coordinates, type/replica word, FP64
// sub-lattice residual, rc from the table (the reference's AssignRCut rule)
// and tail-table index; records inv_perm[i] = d and the rebuild-time lattice
// position backup[i] (caller order, read by the position refresh) and, when
// given, its cell-order copy and the cell-order anchor.
__device__ __forceinline__ void gather_atom(
    int d, int i, const float *pos, int stride, const int *types,
    const int *rep_of, const RepInfo *reps, const float *rc_table, int n_types,
    uint4 *pos4, uint4 *backup, uint32_t *inv_perm, float4 *res4, float *rcut,
    int *tail_idx, const float *rc_atom, uint4 *backup_s, const float *anchor,
    float4 *anchor_s) {
  const int r = rep_of[i];
  const RepInfo &R = reps[r];
  const float *p = pos + (size_t)i * stride;
  const float x = p[0], y = p[1], z = p[2];
  const int ty = types[i];
  uint4 v = make_uint4(quantize(x, R.invq[0]), quantize(y, R.invq[1]),
                       quantize(z, R.invq[2]),
                       (uint32_t)ty | ((uint32_t)r << 8));
  pos4[d] = v;
  backup[i] = v;
  inv_perm[i] = (uint32_t)d;
  if (backup_s) backup_s[d] = v;
  if (anchor_s) {
    const float *a = anchor + (size_t)i * stride;
    anchor_s[d] = make_float4(a[0], a[1], a[2], 0.f);
  }
  if (res4)
    res4[d] = make_float4(quant_residual(x, R.invq[0], R.qd[0]),
                          quant_residual(y, R.invq[1], R.qd[1]),
                          quant_residual(z, R.invq[2], R.qd[2]), 0.f);
  const int ix = rc_bin(x, R.L[0], R.rc_n[0]);
  const int iy = rc_bin(y, R.L[1], R.rc_n[1]);
  const int iz = rc_bin(z, R.L[2], R.rc_n[2]);
  rcut[d] = rc_atom ? rc_atom[i]
                    : rc_table[R.rc_off + iz + R.rc_n[2] * (iy + R.rc_n[1] * ix)];
  tail_idx[d] = (R.tail_off >= 0) ? (R.tail_off + iz) * n_types + ty : -1;
}

// One thread per sorted slot s (plus one): gathers atom perm[s] into slot s
// and fills the cell offsets: start[c] = first sorted position with key >= c,
// thread s covering the cells from key[s-1] + 1 to key[s] (thread n: the
// cells after the last key), so every entry of start[0..n_cells] is written
// exactly once.
__global__ __launch_bounds__(256) void k_gather_sorted(
    int n, int n_cells, const uint32_t *keys, const uint32_t *perm,
    const float *pos, int stride, const int *types, const int *rep_of,
    const RepInfo *reps, const float *rc_table, int n_types, int *start,
    uint4 *pos4, uint4 *backup, uint32_t *inv_perm, float4 *res4, float *rcut,
    int *tail_idx, const float *rc_atom, uint4 *backup_s, const float *anchor,
    float4 *anchor_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > n) return;
  {
    // keys are < n_cells by construction; the clamp keeps the writes inside
    // start[] whatever they hold
    const int lo = s == 0 ? 0 : (int)min(keys[s - 1], (uint32_t)n_cells) + 1;
    const int hi = s == n ? n_cells : (int)min(keys[s], (uint32_t)n_cells);
    for (int c = lo; c <= hi; ++c) start[c] = s;
  }
  if (s == n) return;
  gather_atom(s, (int)perm[s], pos, stride, types, rep_of, reps, rc_table, n_types,
              pos4, backup, inv_perm, res4, rcut, tail_idx, rc_atom, backup_s,
              anchor, anchor_s);
}

// ------------------------------------------------------ cell counting sort --
//
// The rebuild's default binning.  Keys are cell indices, so one count per
// cell fits in memory and a counting sort needs one histogram, one scan and
// one scatter instead of a radix pass per 8 key bits:
//   (1) k_cell_count: the key of atom i (caller order), count[key] += 1;
//   (2) the exclusive scan of the counts, start[c] = first slot of cell c;
//   (3) the scatter: cursor[key] += 1 gives each atom a distinct slot of its
//       cell, in an order that depends on scheduling;
//   (4) k_gather_cells: the atom in a slot of cell c moves to slot
//       start[c] + (number of atoms of c with a smaller caller index), which
//       is the stable order (cell, caller index) -- the radix sort's order --
//       so the result is a pure function of the input; it also sets count
//       and cursor of its cell back to zero for the next rebuild (every cell
//       with a non-zero entry holds an atom).
// Up to kPlaceCells cells, (2) and (3) are one launch (k_cell_place): every
// block of 1024 atoms scans all the counts in LDS itself (up to 60 KB read
// from L2 per block) instead of waiting for a separate scan launch; larger
// grids use k_cell_scan and k_cell_scatter.  A cell holds a
// few atoms (its volume is a fraction of (rc + skin)^3 and the atoms
// repel), so the rank loop of (4) is short.

// kPlaceCells: cells scanned by one 1024-thread block; its LDS holds the
// offsets with one pad word per 16 (thread t reads and writes words 17 t + e,
// so the 64 lanes of a wave spread over the banks), 16320 words of the 16384
// a block may hold
static_assert(kPlaceCells + kPlaceCells / 16 <= 16384, "placement offsets exceed 64 KB");
__device__ __forceinline__ int place_idx(int c) { return c + (c >> 4); }

__global__ __launch_bounds__(256) void k_cell_count(int n, int n_cells, const float *pos,
                                                    int stride, const int *rep_of,
                                                    const RepInfo *reps, uint32_t *keys,
                                                    uint32_t *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // keys are < n_cells by construction; the clamp keeps the counts in bounds
  const uint32_t k = min(cell_key(pos, stride, rep_of, reps, i), (uint32_t)(n_cells - 1));
  keys[i] = k;
  atomicAdd(&count[k], 1u);
}

// One tile of the scan of counts [base, base + 16384) (count holds zeros up
// to a multiple of 16 past m): thread t takes 16 consecutive counts (four
// 16-byte loads); returns the tile total, writes the exclusive prefix plus
// carry through out(c, value).
template <typename Out>
__device__ __forceinline__ uint32_t scan_tile16(int m, int base, const uint32_t *count,
                                                uint32_t carry, uint32_t *tmp, Out out) {
  constexpr int E = 16;
  const int i0 = base + (int)threadIdx.x * E;
  uint32_t v[E];
  if (i0 < m) {
    const uint4 *c4 = reinterpret_cast<const uint4 *>(count + i0);
#pragma unroll
    for (int q = 0; q < E / 4; ++q) {
      const uint4 w = c4[q];
      v[4 * q] = w.x; v[4 * q + 1] = w.y; v[4 * q + 2] = w.z; v[4 * q + 3] = w.w;
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = 0u;
  }
  uint32_t sum = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) sum += v[e];
  uint32_t tot = 0;
  uint32_t ex = block_exscan_1024(sum, tmp, &tot) + carry;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (i0 + e < m) out(i0 + e, ex);
    ex += v[e];
  }
  return tot;
}

// start[c] for c < m (any m), by one block, tile after tile.
__global__ __launch_bounds__(kSortT) void k_cell_scan(int m, const uint32_t *count,
                                                      int *start) {
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  uint32_t carry = 0;
  for (int base = 0; base < m; base += kSortT * 16)
    carry += scan_tile16(m, base, count, carry, tmp,
                         [&](int c, uint32_t v) { start[c] = (int)v; });
}

__global__ __launch_bounds__(256) void k_cell_scatter(int n, int n_cells,
                                                      const uint32_t *keys,
                                                      int *start, uint32_t *cursor,
                                                      uint32_t *skey, uint32_t *sval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) start[n_cells] = n;
  if (i >= n) return;
  const uint32_t k = keys[i];
  const int slot = start[k] + (int)atomicAdd(&cursor[k], 1u);  // in [start[k], start[k+1])
  if ((unsigned)slot < (unsigned)n) {
    skey[slot] = k;
    sval[slot] = (uint32_t)i;
  }
}

// (2) + (3) for m <= kPlaceCells: block b scans all counts in LDS (loaded
// and, by block 0, written to start[0..m] with coalesced accesses; each
// thread scans 16 consecutive counts in between), then each thread places
// one atom.
__global__ __launch_bounds__(kSortT) void k_cell_place(int n, int m, const uint32_t *keys,
                                                       const uint32_t *count,
                                                       uint32_t *cursor, int *start,
                                                       uint32_t *skey, uint32_t *sval) {
  constexpr int E = 16;
  __shared__ int s_start[kPlaceCells + kPlaceCells / 16];
  __shared__ uint32_t tmp[kSortT / 64 + 1];
  for (int c = threadIdx.x; c < m; c += kSortT) s_start[place_idx(c)] = (int)count[c];
  __syncthreads();
  const int i0 = (int)threadIdx.x * E;
  uint32_t v[E];
  uint32_t sum = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    v[e] = i0 + e < m ? (uint32_t)s_start[place_idx(i0 + e)] : 0u;
    sum += v[e];
  }
  uint32_t ex = block_exscan_1024(sum, tmp, nullptr);   // ends with a barrier
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (i0 + e < m) s_start[place_idx(i0 + e)] = (int)ex;
    ex += v[e];
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int c = threadIdx.x; c < m; c += kSortT) start[c] = s_start[place_idx(c)];
    if (threadIdx.x == 0) start[m] = n;
  }
  const int i = blockIdx.x * kSortT + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];     // < m (clamped when counted)
  const int slot = s_start[place_idx((int)k)] + (int)atomicAdd(&cursor[k], 1u);
  if ((unsigned)slot < (unsigned)n) {
    skey[slot] = k;
    sval[slot] = (uint32_t)i;
  }
}

__global__ __launch_bounds__(256) void k_gather_cells(
    int n, const uint32_t *skey, const uint32_t *sval, const int *start,
    uint32_t *count, uint32_t *cursor, uint32_t *perm, const float *pos,
    int stride, const int *types, const int *rep_of, const RepInfo *reps,
    const float *rc_table, int n_types, uint4 *pos4, uint4 *backup,
    uint32_t *inv_perm, float4 *res4, float *rcut, int *tail_idx,
    const float *rc_atom, uint4 *backup_s, const float *anchor, float4 *anchor_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = skey[s], i = sval[s];
  count[k] = 0u;
  cursor[k] = 0u;
  const int lo = start[k], hi = min(start[k + 1], n);
  int d = lo;
  for (int t = lo; t < hi; ++t) d += sval[t] < i ? 1 : 0;
  if ((unsigned)d >= (unsigned)n) return;       // only for inconsistent offsets
  perm[d] = i;
  gather_atom(d, (int)i, pos, stride, types, rep_of, reps, rc_table, n_types, pos4,
              backup, inv_perm, res4, rcut, tail_idx, rc_atom, backup_s, anchor,
              anchor_s);
}

static size_t cell_count_half(int n_cells) {
  return ((size_t)n_cells + 15) / 16 * 16 + 16;
}

size_t cell_count_capacity(int n_cells) { return 2 * cell_count_half(n_cells); }

void launch_cell_sort_gather(int n, int n_cells, const float *pos, int stride,
                             const int *types, const int *rep_of,
                             const RepInfo *reps, const float *rc_table,
                             int n_types, const SortBufs &b, uint32_t *count,
                             int *start, uint4 *pos4, uint4 *backup,
                             uint32_t *inv_perm, float4 *res4, float *rcut,
                             int *tail_idx, const float *rc_atom, uint4 *backup_s,
                             const float *anchor, float4 *anchor_s, hipStream_t st,
                             bool counted) {
  if (n <= 0) return;
  uint32_t *cursor = count + cell_count_half(n_cells);
  const dim3 g((n + 255) / 256), t(256);
  if (!counted)
    hipLaunchKernelGGL(k_cell_count, g, t, 0, st, n, n_cells, pos, stride, rep_of, reps,
                       b.tkey, count);
  if (n_cells <= kPlaceCells) {
    hipLaunchKernelGGL(k_cell_place, dim3((n + kSortT - 1) / kSortT), dim3(kSortT), 0, st,
                       n, n_cells, b.tkey, count, cursor, start, b.key_b, b.val_b);
  } else {
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(kSortT), 0, st, n_cells, count, start);
    hipLaunchKernelGGL(k_cell_scatter, g, t, 0, st, n, n_cells, b.tkey, start, cursor,
                       b.key_b, b.val_b);
  }
  hipLaunchKernelGGL(k_gather_cells, g, t, 0, st, n, b.key_b, b.val_b, start, count,
                     cursor, b.val_a, pos, stride, types, rep_of, reps, rc_table,
                     n_types, pos4, backup, inv_perm, res4, rcut, tail_idx, rc_atom,
                     backup_s, anchor, anchor ? anchor_s : nullptr);
}

// AssignRCut::assign between rebuilds (H/src/AssignRCut.cu:77-117): every
// atom's rc from the table at its current position, kept per atom in caller
// order (rc_atom, which later rebuilds carry into cell order) and written to
// the atom's slot of the current list order (inv_perm, when a list exists).
__global__ __launch_bounds__(256) void k_assign_rc(int n, const float *pos, int stride,
                                                   const int *rep_of,
                                                   const RepInfo *reps,
                                                   const float *rc_table,
                                                   const uint32_t *inv_perm,
                                                   float *rc_atom, float *rcut) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const RepInfo &R = reps[rep_of[i]];
  const float *p = pos + (size_t)i * stride;
  const int ix = rc_bin(p[0], R.L[0], R.rc_n[0]);
  const int iy = rc_bin(p[1], R.L[1], R.rc_n[1]);
  const int iz = rc_bin(p[2], R.L[2], R.rc_n[2]);
  const float v = rc_table[R.rc_off + iz + R.rc_n[2] * (iy + R.rc_n[1] * ix)];
  rc_atom[i] = v;
  if (inv_perm) rcut[inv_perm[i]] = v;
}

void launch_assign_rc(int n, const float *pos, int stride, const int *rep_of,
                      const RepInfo *reps, const float *rc_table,
                      const uint32_t *inv_perm, float *rc_atom, float *rcut,
                      hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_assign_rc, dim3((n + 255) / 256), dim3(256), 0, st, n, pos,
                     stride, rep_of, reps, rc_table, inv_perm, rc_atom, rcut);
}

void launch_gather_sorted(int n, int n_cells, const uint32_t *keys,
                          const uint32_t *perm, const float *pos, int stride,
                          const int *types, const int *rep_of,
                          const RepInfo *reps, const float *rc_table,
                          int n_types, int *start, uint4 *pos4, uint4 *backup,
                          uint32_t *inv_perm, float4 *res4, float *rcut,
                          int *tail_idx, const float *rc_atom, hipStream_t st,
                          uint4 *backup_s, const float *anchor, float4 *anchor_s) {
  hipLaunchKernelGGL(k_gather_sorted, dim3((n + 1 + 255) / 256), dim3(256), 0,
                     st, n, n_cells, keys, perm, pos, stride, types, rep_of,
                     reps, rc_table, n_types, start, pos4, backup, inv_perm,
                     res4, rcut, tail_idx, rc_atom, backup_s, anchor,
                     anchor ? anchor_s : nullptr);
}

__global__ void k_gather_anchor(int n, const uint32_t *perm, const float *anchor,
                                int stride, float4 *anchor_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float *a = anchor + (size_t)perm[s] * stride;
  anchor_s[s] = make_float4(a[0], a[1], a[2], 0.f);
}

void launch_gather_anchor(int n, const uint32_t *perm, const float *anchor, int stride,
                          float4 *anchor_s, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_anchor, dim3((n + 255) / 256), dim3(256), 0, st, n, perm,
                     anchor, stride, anchor_s);
}

// ------------------------------------------------------------------ fills --

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i64(int64_t *p, int64_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, st, p, v, n);
}
void launch_fill_i64(int64_t *p, int64_t v, size_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, st, p, v, n);
}

}  // namespace sre
