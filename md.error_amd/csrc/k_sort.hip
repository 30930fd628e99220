// k_sort.hip -- cell keys, atomic-free LSD radix sort, cell offsets, and the
// gather of per-atom data into cell order.
//
// The reference bins atoms with fixed-capacity cells and atomicInc slot
// insertion (H/src/NeighborList.cu:84-167), whose in-cell order depends on
// scheduling.  Here binning is a stable LSD radix sort of the cell key over
// digits of at most 8 bits, each pass built from
//   (1) a per-block stable ranking of 1024 keys: lanes holding the same digit
//       find each other with one ballot per digit bit (their rank is the
//       number of such lanes below them), per-wave digit counts go through
//       LDS, and the keys are written in (digit, original position) order;
//   (2) a digit-major exclusive scan of the per-block digit counts;
//   (3) a stable scatter to the global positions.
// No atomics are used, so the cell order (and everything derived from it) is
// a pure function of the input.
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

// Cell key of atom i: replica cell offset + (cx, cy, cz) with z fastest, so a
// column of cells is one contiguous range of sorted atoms.
__device__ __forceinline__ uint32_t cell_key(const float *pos, int stride,
                                             const int *rep_of,
                                             const RepInfo *reps, int i) {
  const RepInfo &R = reps[rep_of[i]];
  const float *p = pos + (size_t)i * stride;
  const int cx = cell_of(quantize(p[0], R.invq[0]), R.nc[0]);
  const int cy = cell_of(quantize(p[1], R.invq[1]), R.nc[1]);
  const int cz = cell_of(quantize(p[2], R.invq[2]), R.nc[2]);
  return (uint32_t)(R.cell_off + (cx * R.nc[1] + cy) * R.nc[2] + cz);
}

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

// One thread per sorted slot s (plus one): gathers atom i = perm[s] into
// cell order -- lattice coordinates, type/replica word, FP64 sub-lattice
// residual, rc from the table (the reference's AssignRCut rule) and
// tail-table index -- records inv_perm[i] = s and the rebuild-time lattice
// position backup[i] (caller order, read by the position refresh), and fills
// the cell offsets: start[c] = first sorted
// position with key >= c, thread s covering the cells from key[s-1] + 1 to
// key[s] (thread n: the cells after the last key), so every entry of
// start[0..n_cells] is written exactly once.
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
  const int i = (int)perm[s];
  const int r = rep_of[i];
  const RepInfo &R = reps[r];
  const float *p = pos + (size_t)i * stride;
  const float x = p[0], y = p[1], z = p[2];
  const int ty = types[i];
  uint4 v = make_uint4(quantize(x, R.invq[0]), quantize(y, R.invq[1]),
                       quantize(z, R.invq[2]),
                       (uint32_t)ty | ((uint32_t)r << 8));
  pos4[s] = v;
  backup[i] = v;
  inv_perm[i] = (uint32_t)s;
  if (backup_s) backup_s[s] = v;
  if (anchor_s) {
    const float *a = anchor + (size_t)i * stride;
    anchor_s[s] = make_float4(a[0], a[1], a[2], 0.f);
  }
  if (res4)
    res4[s] = make_float4(quant_residual(x, R.invq[0], R.qd[0]),
                          quant_residual(y, R.invq[1], R.qd[1]),
                          quant_residual(z, R.invq[2], R.qd[2]), 0.f);
  const int ix = rc_bin(x, R.L[0], R.rc_n[0]);
  const int iy = rc_bin(y, R.L[1], R.rc_n[1]);
  const int iz = rc_bin(z, R.L[2], R.rc_n[2]);
  rcut[s] = rc_atom ? rc_atom[i]
                    : rc_table[R.rc_off + iz + R.rc_n[2] * (iy + R.rc_n[1] * ix)];
  tail_idx[s] = (R.tail_off >= 0) ? (R.tail_off + iz) * n_types + ty : -1;
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
