// sre_internal.h -- types and device helpers shared by the SRE HIP sources.
//
// Numerical definitions used by every kernel (the oracle restates them):
//
//  * Fixed-point lattice coordinate of an FP32 position x in a box edge L:
//        u = (floor(double(x) * (2^32 / double(L)) + 0.5) mod 2^32)
//    evaluated in binary64.  Minimum image is then plain int32 wraparound of
//    u_j - u_i, and the displacement is  d = float(int32(u_j - u_i)) * q,
//    q = L * 2^-32 (exact in FP32).  The reference computes
//    d -= floorf(d/L + 0.5) L in FP32 (H/include/BoxGeometry.h:171-186); both
//    pick the nearest image, ours without rounding in the wrap.
//  * Squared distance: d2 = fmaf(dz, dz, fmaf(dy, dy, dx*dx)) (FP32, exactly
//    this operation order; the library is compiled with -ffp-contract=off).
//  * List predicate  j != i && d2 < rl*rl,  rl = rc_i + skin  (FP32), the
//    reference's H/src/NeighborList.cu:391.  Force predicate d2 < rc_i*rc_i
//    (H/src/InteractionEngine_interface.cu:1224-1245).
//  * rc bin of an atom: one-image wrap, ix = int((x*nx)/L) in FP32 with an
//    IEEE division (H/src/AssignRCut.cu:92-101), clamped to [0, nx-1].
//  * Forces: int64 fixed point, scale 2^32, each pair component rounded
//    down to an integer (floor) before the integer sum (order-independent).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sre {

constexpr int kWave = 64;
constexpr int kMaxTypes = 8;
constexpr int kMaxPairTypes = kMaxTypes * kMaxTypes;
constexpr double kFixScale = 4294967296.0;          // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;
constexpr double kMagic = 6755399441055744.0;       // 1.5 * 2^52
constexpr float kPairForceLimit = 524288.0f;        // 2^19: |pair component| limit of the magic rounding
constexpr int kForceUnroll = 4;   // most list rows per group in the mixed force kernel
// list rows allocated past the last tile: the force kernel prefetches up to
// 2 * kForceUnroll rows beyond a tile's end
constexpr int kListPadRows = 2 * kForceUnroll + 2;

// Geometry and table offsets of one replica (device-resident array).
struct RepInfo {
  float L[3];       // box edges (FP32, as given)
  float q[3];       // L * 2^-32
  double invq[3];   // 2^32 / L
  double qd[3];     // L * 2^-32 in FP64
  int nc[3];        // cell grid
  float h[3];       // cell edges L / nc
  int cell_off;     // first global cell id of this replica
  int col_off;      // first global cell column (x, y) of this replica
  int atom_off;     // first atom (caller order and sorted order)
  int n_atoms;
  int rc_off;       // offset of the rc lattice in the concatenated table
  int rc_n[3];      // rc lattice dims nx, ny, nz
  int tail_off;     // offset (in entries of n_types) of the tail table, -1 if none
  int rho_off;      // offset of rho[n_types][rho_nz], -1 if none
  int rho_nz;
  float rc_min, rc_max;
  int fc_off;       // first cell of the force-correction grid, -1 if none
  int fc_n[3];      // force-correction grid dims
};

// Cell grid of a replica: edges of about cell_xy (x, y) and cell_z (z) times
// rc_min + skin, coarsened by 10 % steps while there are more than
// 8 N + 4096 cells; sets R.nc and R.h (R.L, R.n_atoms and R.rc_min set).
inline void choose_cell_grid(RepInfo &R, float skin, float cell_xy, float cell_z) {
  const double rlmin = R.rc_min + skin;
  double hxy = (cell_xy > 0.05f ? cell_xy : 0.05) * rlmin;
  double hz = (cell_z > 0.05f ? cell_z : 0.05) * rlmin;
  int nc[3];
  for (int tries = 0;; ++tries) {
    for (int d = 0; d < 3; ++d) {
      const int v = (int)floor(R.L[d] / (d < 2 ? hxy : hz));
      nc[d] = v > 1 ? v : 1;
    }
    const long long c = (long long)nc[0] * nc[1] * nc[2];
    if (c <= 8ll * R.n_atoms + 4096 || tries > 40) break;
    hxy *= 1.1;
    hz *= 1.1;
  }
  for (int d = 0; d < 3; ++d) {
    R.nc[d] = nc[d];
    R.h[d] = R.L[d] / (float)nc[d];
  }
}

// LJ pair constants in the form used by the kernels:
//   u = ri6*(A*ri6 - B) + shift,  F_i = d * ri2*ri6*(B6 - A12*ri6)
// with A = 4 eps sig^12, B = 4 eps sig^6, A12 = 12 A, B6 = 6 B.
struct PairParam {
  float A12, B6, A, B;   // A12 and B6 are stored multiplied by 2^32
  float shift, eps, sig, pad;
};
struct PairParamD {
  double A12, B6, A, B;
  double shift, eps, sig, pad;
};

// The kernels' constants of one type pair from (eps, sigma, shift); the
// mixed kernel's force constants carry the fixed-point scale 2^32 (an exact
// power-of-two factor), so its pair components come out pre-scaled.
inline void make_pair_params(double ep, double sg, double sh, PairParam *p,
                             PairParamD *pd) {
  const double s6 = sg * sg * sg * sg * sg * sg, s12 = s6 * s6;
  const double A = 4.0 * ep * s12, B = 4.0 * ep * s6;
  *pd = PairParamD{12.0 * A, 6.0 * B, A, B, sh, ep, sg, 0.0};
  *p = PairParam{(float)(12.0 * A * kFixScale), (float)(6.0 * B * kFixScale),
                 (float)A, (float)B, (float)sh, (float)ep, (float)sg, 0.f};
}

// ---------------------------------------------------------------- helpers --

__host__ __device__ inline uint32_t quantize(float x, double invq) {
  double t = (double)x * invq;
  double f = floor(t + 0.5);
  double w = f - floor(f * 0x1p-32) * 0x1p32;
  return (uint32_t)w;
}

// Sub-lattice residual r = x - u*q (mod L) of the same coordinate, in FP64
// length units (|r| <= q/2); the FP64 form adds r_j - r_i to the lattice
// displacement so that it does not inherit the 2^-32 quantization.
__host__ __device__ inline float quant_residual(float x, double invq, double qd) {
  double t = (double)x * invq;
  double f = floor(t + 0.5);
  return (float)((t - f) * qd);
}

__host__ __device__ inline int cell_of(uint32_t u, int nc) {
  return (int)(((uint64_t)u * (uint64_t)nc) >> 32);
}

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

// Position gathers through a buffer resource over pos4 in index mode: a
// resource with a 16-byte record stride takes the atom index itself
// (buffer_load ... idxen), so a gather needs neither a 64-bit address nor a
// per-pair shift (pos4 holds fewer than 2^27 atoms, checked when the system
// is set).  gfx9 resource word 3 = 0x00020000 (32-bit data format, no
// swizzle).  These are the LLVM structured buffer loads, which this compiler
// exposes under their intrinsic names only.
typedef unsigned int sre_u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int sre_u32x4 __attribute__((ext_vector_type(4)));
extern "C" __device__ sre_u32x3 sre_struct_load_v3(__amdgpu_buffer_rsrc_t rsrc, int vindex,
                                                   int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v3i32");
extern "C" __device__ sre_u32x4 sre_struct_load_v4(__amdgpu_buffer_rsrc_t rsrc, int vindex,
                                                   int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");

__device__ __forceinline__ __amdgpu_buffer_rsrc_t pos_rsrc_idx(const uint4 *p, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(p), (short)16, n * 16,
                                           0x00020000);
}
template <bool W>
__device__ __forceinline__ uint4 gather_pos_idx(__amdgpu_buffer_rsrc_t r, int j) {
  if (W) {
    const sre_u32x4 v = sre_struct_load_v4(r, j, 0, 0, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
  } else {
    const sre_u32x3 v = sre_struct_load_v3(r, j, 0, 0, 0);
    return make_uint4(v[0], v[1], v[2], 0u);
  }
}

__host__ __device__ inline int rc_bin(float x, float L, int n) {
  if (x < 0.f) x = x + L;
  else if (x >= L) x = x - L;
  int ix = (int)((x * (float)n) / L);
  ix = ix < 0 ? 0 : ix;
  ix = ix > n - 1 ? n - 1 : ix;
  return ix;
}

// Counter-based Gaussian deviates of (seed, step, atom): SplitMix64 (Steele,
// Lea and Flood, "Fast splittable pseudorandom number generators", OOPSLA
// 2014) hashes followed by the Box-Muller transform, so a perturbation does
// not depend on launch shape or on the order atoms are visited in.  Two
// 64-bit hashes per atom give four 24-bit uniforms in (0, 1]; the logarithm and the
// sine and cosine are the hardware's (v_log_f32 is log2; v_sin_f32 and
// v_cos_f32 take their argument in turns), accurate to a few ulp, which a
// synthetic thermal displacement does not need more of.
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline float u24(uint32_t bits24) {
  // (0, 1]: never zero, so the logarithm is finite
  return ((float)bits24 + 1.0f) * (1.0f / 16777216.0f);
}

// The step's key, the same for every atom (so computed once per wavefront,
// on the scalar unit); atom i's Box-Muller pair p (p = 0: components 0 and
// 1, p = 1: component 2) draws its two uniforms from one hash of
// key ^ (2i + p).
__device__ __forceinline__ uint64_t gauss_key(uint64_t seed, uint64_t step) {
  return splitmix64(seed ^ splitmix64(step * 0x632BE59BD9B4E019ull));
}

__device__ inline float3 gauss3(uint64_t seed, uint64_t step, uint64_t i) {
  const uint64_t key = gauss_key(seed, step);
  const uint64_t h1 = splitmix64(key ^ (i << 1)), h2 = splitmix64(key ^ (i << 1 | 1));
  const float u1 = u24((uint32_t)(h1 >> 40)), u2 = u24((uint32_t)(h1 >> 8) & 0xffffffu);
  const float u3 = u24((uint32_t)(h2 >> 40)), u4 = u24((uint32_t)(h2 >> 8) & 0xffffffu);
  // sqrt(-2 ln u) = sqrt(-2 ln2 log2 u)
  const float r1 = __builtin_sqrtf(-1.3862943611f * __builtin_amdgcn_logf(u1));
  const float r2 = __builtin_sqrtf(-1.3862943611f * __builtin_amdgcn_logf(u3));
  return make_float3(r1 * __builtin_amdgcn_cosf(u2), r1 * __builtin_amdgcn_sinf(u2),
                     r2 * __builtin_amdgcn_cosf(u4));
}

// Component c of gauss3 (key = gauss_key(seed, step)), with the same
// operations.
__device__ inline float gauss_component(uint64_t key, uint64_t i, int c) {
  const uint64_t h = splitmix64(key ^ (i << 1 | (uint64_t)(c >> 1)));
  const float ua = u24((uint32_t)(h >> 40)), ub = u24((uint32_t)(h >> 8) & 0xffffffu);
  const float r = __builtin_sqrtf(-1.3862943611f * __builtin_amdgcn_logf(ua));
  return c == 1 ? r * __builtin_amdgcn_sinf(ub) : r * __builtin_amdgcn_cosf(ub);
}

// Cross-lane moves without LDS (DPP): the value of lane k of this lane's
// quad, and the maximum of a non-negative float over the wavefront (read on
// every lane; all 64 lanes must be active).
template <int K>
__device__ __forceinline__ float quad_lane(float v) {
  // quad_perm [K, K, K, K]
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), K * 0x55, 0xF, 0xF,
                                                    false));
}
__device__ __forceinline__ float wave_max_nonneg(float v) {
  // non-negative floats order as their bit patterns; quad, half row, row,
  // then rows 0 -> 1 and 2 -> 3, then row 1 -> rows 2 and 3: lane 63 ends
  // with the maximum
  int b = __float_as_int(v);
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0xB1, 0xF, 0xF, false));   // quad [1,0,3,2]
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0x4E, 0xF, 0xF, false));   // quad [2,3,0,1]
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0x141, 0xF, 0xF, false));  // row_half_mirror
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0x140, 0xF, 0xF, false));  // row_mirror
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0x142, 0xA, 0xF, false));  // row_bcast:15
  b = max(b, __builtin_amdgcn_update_dpp(b, b, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(b, 63));
}

__device__ __forceinline__ float disp(uint32_t uj, uint32_t ui, float q) {
  return (float)(int32_t)(uj - ui) * q;
}

__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Round v*2^32 to the nearest integer (ties to even) and return the bit
// pattern of (v*2^32 + 1.5*2^52); subtracting bits(1.5*2^52) recovers the
// integer.  Valid for |v| < 2^19.
__device__ __forceinline__ uint64_t fix_bits(double v) {
  double t = __builtin_fma(v, kFixScale, kMagic);
  return (uint64_t)__double_as_longlong(t);
}
__device__ __forceinline__ uint64_t magic_bits() {
  return (uint64_t)__double_as_longlong(kMagic);
}
// Exact conversion of an FP64 value (|v| < 2^19) to fixed point.
__device__ __forceinline__ int64_t to_fix(double v) {
  return (int64_t)(fix_bits(v) - magic_bits());
}

// Exact integer sums of FP32 terms: a lane keeps an FP64 accumulator
// started at C = 1.5*2^52.  While it stays in [2^52, 2^53) the representable
// values are exactly the integers, and with the FP64 rounding mode set to
// round toward zero, acc + c rounds to acc + floor(c): one conversion and
// one add per term give the exact integer sum C + sum floor(c), the same in
// any order (floor instead of round-to-nearest: a bias of half a unit per
// term).  The lane value is bits(acc) - bits(C).
// The rounding mode is switched by hand around the pair loop, and the
// conversions and adds that depend on it are written as one asm statement
// each, so that the compiler, which assumes the default mode for every FP64
// instruction it emits, neither places its own FP64 code under the switched
// mode nor switches the mode back in front of these adds.
constexpr double kAccBase = 6755399441055744.0;   // 1.5 * 2^52
__device__ __forceinline__ void acc_floor(double &acc, float c) {
  double t;
  asm("v_cvt_f64_f32 %1, %2\n\tv_add_f64 %0, %0, %1" : "+v"(acc), "=&v"(t) : "v"(c));
}
// MODE register field 3:2 (FP64/FP16 rounding): 3 = toward zero, 0 = nearest even
__device__ __forceinline__ void round_f64_toward_zero() {
  asm volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 3" ::: "memory");
}
__device__ __forceinline__ void round_f64_nearest() {
  asm volatile("s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0" ::: "memory");
}

// list slot of entry k of sorted atom s: waves of 64/T atoms, entry rows of 64
__host__ __device__ inline size_t list_slot(int s, int k, int T, int cap) {
  int A = kWave / T;
  int w = s / A, a = s - w * A;
  return (size_t)w * (size_t)cap * (size_t)A + (size_t)(k / T) * kWave +
         (size_t)(a * T + (k % T));
}

// ------------------------------------------------------------ launchers ----
// Implemented in the k_*.hip files; called from engine.hip.

struct SortBufs {
  uint32_t *key_a, *val_a, *key_b, *val_b, *tkey, *tval;
  uint32_t *counts, *offsets, *scan_tmp;
};

// Sorts the atoms by cell key (stable; key < 2^key_bits, computed from the
// positions in the first pass): sorted keys in key_a, the permutation
// (sorted slot -> atom) in val_a, always there, so it has a fixed address
// that captured graphs read.  Returns 0 on success.
int radix_passes(int key_bits);
int launch_radix_sort(int n, int key_bits, const float *pos, int stride,
                      const int *rep_of, const RepInfo *reps, const SortBufs &b,
                      hipStream_t st);
size_t radix_counts_capacity(int n, int key_bits);
// Counting sort by cell and gather into cell order (+ rc assignment): the
// same permutation (val_a), cell offsets start[0..n_cells] and gathered data
// as the radix sort followed by launch_gather_sorted.  count (the cell
// counts and the placement cursors) holds cell_count_capacity(n_cells) zeros
// before the call and again after it; uses b.tkey, b.key_b, b.val_b as
// scratch.
size_t cell_count_capacity(int n_cells);
// up to this many cells the scan of the counts is fused into the placement
constexpr int kPlaceCells = 15360;
// counted: the keys (b.tkey) and counts were made by a CountFold of the last
// force launch, so the histogram launch is skipped.
void launch_cell_sort_gather(int n, int n_cells, const float *pos, int stride,
                             const int *types, const int *rep_of,
                             const RepInfo *reps, const float *rc_table,
                             int n_types, const SortBufs &b, uint32_t *count,
                             int *start, uint4 *pos4, uint4 *backup,
                             uint32_t *inv_perm, float4 *res4, float *rcut,
                             int *tail_idx, const float *rc_atom, uint4 *backup_s,
                             const float *anchor, float4 *anchor_s, hipStream_t st,
                             bool counted = false);
// Gather into cell order (+ rc assignment) fused with the cell offsets
// start[0..n_cells] computed from the sorted keys.
void launch_gather_sorted(int n, int n_cells, const uint32_t *keys,
                          const uint32_t *perm, const float *pos, int stride,
                          const int *types, const int *rep_of,
                          const RepInfo *reps, const float *rc_table,
                          int n_types, int *start, uint4 *pos4, uint4 *backup,
                          uint32_t *inv_perm, float4 *res4, float *rcut,
                          int *tail_idx, const float *rc_atom, hipStream_t st,
                          uint4 *backup_s = nullptr, const float *anchor = nullptr,
                          float4 *anchor_s = nullptr);
// The same rebuild-time data in cell order for the next-step refresh of
// force launches: backup_s[s] (the gather above) and anchor_s[s] =
// anchor[perm[s]] (x, y, z, 0), also after a new anchor.
void launch_gather_anchor(int n, const uint32_t *perm, const float *anchor, int stride,
                          float4 *anchor_s, hipStream_t st);
// rc_atom[i] (caller order) from the table at the atom's current position,
// also into rcut[inv_perm[i]] when inv_perm is given
void launch_assign_rc(int n, const float *pos, int stride, const int *rep_of,
                      const RepInfo *reps, const float *rc_table,
                      const uint32_t *inv_perm, float *rc_atom, float *rcut,
                      hipStream_t st);
// Both builders write one (longest list, total length) pair per wavefront to
// part[] (build_part_count entries, every one written by every launch);
// launch_rebuild_info reduces them.
int build_part_count(int mode, int n, int lanes, int n_cols);
void launch_build_list(int lanes, int force_lanes, int n, const uint4 *pos4,
                       const float *rcut, const RepInfo *reps, const int *start,
                       float skin, int cap, int *list, int *nnb, int2 *part,
                       hipStream_t st);
// Tiles of <= tile atoms inside one cell column: tiles_per_col and tile_off
// (n_cols each) filled on the device from the cell offsets.
void launch_tile_map(int n_cols, int n_rep, const RepInfo *reps, const int *start,
                     int tile, uint32_t *tiles_per_col, uint32_t *tile_off,
                     uint32_t *scan_tmp, hipStream_t st);
// Tile build: tiles of <= 64 atoms inside one cell column.  tiles_per_col and
// tile_off (n_cols each) are scratch arrays filled on the device.
void launch_build_tiles(int n, int n_cols, int n_rep, int force_lanes,
                        const uint4 *pos4, const float *rcut,
                        const RepInfo *reps, const int *start, float skin,
                        int cap, int *list, int *nnb, int2 *part,
                        uint32_t *tiles_per_col, uint32_t *tile_off,
                        uint32_t *scan_tmp, hipStream_t st);
// counters[0] = longest list, counters[1] = total list length over the
// n_build partials (skipped when n_build == 0); *skin_bits = the largest
// squared displacement (float bits) over the n_skin partials, which are
// zeroed after reading when clear_skin (skipped when n_skin == 0).  Fixed
// order: no atomics.  With a log (device-mapped host memory of log_cap
// entries of three words), the three values (longest, total, displacement^2
// bits) also go to entry (*log_idx % log_cap), and *log_idx is advanced.
// With slices (2 kInfoSlices int2 of scratch) and more than kInfoSliceLen
// partials of either kind, slices of kInfoSliceLen are first reduced by one
// block each (same results: max and sum are exact in any order).
constexpr int kInfoSliceLen = 16384;
constexpr int kInfoSlices = 1024;     // enough for 2^24 partials
void launch_rebuild_info(int n_build, const int2 *build_part,
                         unsigned long long *counters, int n_skin,
                         float *skin_part, int clear_skin, unsigned *skin_bits,
                         hipStream_t st, unsigned long long *log = nullptr,
                         unsigned *log_idx = nullptr, int log_cap = 0,
                         int2 *slices = nullptr);
void device_exscan_u32(size_t m, const uint32_t *in, uint32_t *out,
                       uint32_t *tmp, hipStream_t st);
// Box geometry the position refresh needs; passed by value when all atoms
// share one replica, so that no load of the replica table precedes the
// atom's loads.
struct Geom {
  float L[3], q[3];
  double invq[3], qd[3];
};

// step_ctr (optional) is advanced by one thread after the positions are read
// Optional perturbation folded into the position refresh:
// pos = (anchor ? anchor : pos) + sigma * gauss3(seed, *step_ctr, atom).
struct PerturbArgs {
  float sigma;
  uint64_t seed;
  const float *anchor;
  const uint64_t *step_ctr;
};

// MD integrator state on the device (k_md.hip): two-link Nose-Hoover chain
// (reference NoseHoover_Chains2: xi init 1, vxi init 0, Q1 = tau^2 T LL /
// (4 pi^2), Q2 = Q1 / LL, LL = 3N) and the velocity factors of the next
// drift: v <- a2 (a1 v - vcom).
struct MDState {
  double xi1, xi2, vxi1, vxi2;
  double Q1, Q2, LL, refT;
  double a1, a2;
  double vcom[3];
  double K;            // kinetic energy after the last factors/finish
  double dt, total_mass;
  int thermostat, remove_com, have_end, pad;
};
// Optional velocity Verlet kick + drift folded into the position refresh:
// v <- a2 (a1 v - vcom) + F dt / (2 m); x <- x + v dt (caller order; F from
// fcall, saved by the previous kick).
struct DriftArgs {
  float *vel;
  const float *fcall;
  const float *mass;   // per type
  const int *types;
  const MDState *state;
  float dt;
};
// Grid force correction (AssignForceCorr::assign + the force add of the
// force-correction engine): per atom the cell of its current position by the
// reference's FP32 rule, F += grid[cell].
void launch_add_force_grid(int n, const float *pos, int stride, const uint4 *backup,
                           const uint32_t *inv_perm, const RepInfo *reps,
                           const float4 *grid, int64_t *fx, int64_t *fy, int64_t *fz,
                           double *fxd, double *fyd, double *fzd, hipStream_t st);
constexpr int kCopySegs = 6;
struct CopySegs {
  const uint32_t *src[kCopySegs];
  uint32_t *dst[kCopySegs];
  size_t words[kCopySegs];
  int count;
};
void launch_md_copy(const CopySegs &c, hipStream_t st);
int md_part_count(int n);
void launch_md_kick(int n, const uint32_t *inv_perm, const int *types,
                    const int64_t *fx, const int64_t *fy, const int64_t *fz,
                    const double *fxd, const double *fyd, const double *fzd,
                    const float *mass, float hdt, float *vel, float *fcall,
                    double *part, hipStream_t st);
void launch_md_factors(int n, const double *part, MDState *state,
                       hipStream_t st);
void launch_md_finish(int n, const double *part, MDState *state, float *vel,
                      hipStream_t st);
// Position refresh + skin test: each wavefront keeps the largest squared
// displacement it has seen since the last rebuild in skin_part[wave]
// (skin_part_count(n) entries).
inline int skin_part_count(int n) { return (n + kWave - 1) / kWave; }
// backup and inv_perm are in caller order (written by the sorted gather)
// host_rep0: the host copy of the only replica's geometry (NULL with several)
void launch_update_positions(int n, const uint32_t *inv_perm, float *pos,
                             int stride, const RepInfo *reps, const RepInfo *host_rep0,
                             uint4 *pos4, const uint4 *backup, float4 *res4,
                             float *skin_part, const PerturbArgs *pert,
                             const DriftArgs *drift, hipStream_t st);

// The next step's position refresh folded into a force launch (perturbation
// mode, no integrator): n_blocks blocks in front of the force blocks
// refresh every atom into a second coordinate buffer -- the arithmetic of
// k_update_positions with a perturbation at step *ctr_in + 1 -- so that the
// launch of step s also produces the coordinates of step s + 1 while its
// force wavefronts read those of step s.  Each refresh wavefront keeps its
// running maximum squared displacement in skin_part (no atomics).
struct NextRefresh {
  int active;
  int n_blocks;                // refresh blocks (a multiple of 8)
  const uint32_t *perm;        // sorted slot -> caller index
  float *pos;                  // caller positions (the new ones written back)
  int stride;
  const float4 *anchor_s;      // positions perturbed about, cell order, or NULL
                               // (then pos itself, caller order)
  const uint4 *backup_s;       // rebuild-time lattice positions (cell order)
  uint4 *pos4;                 // the next step's lattice positions (sorted)
  float4 *res4;                // FP64 form: the next step's residuals, or NULL
  float *skin_part;            // running maximum per 64 cell-order atoms
  float sigma;
  uint64_t seed;
  Geom g1;                     // the geometry when one_rep
  int one_rep;
};

// The histogram step of the counting sort of the rebuild that follows a
// force launch, done by blocks in front of that launch when it is the last
// of a segment (it does not move the atoms, so the keys are those the
// rebuild would compute): one atom per thread, caller order, keys[i] and
// count[key] += 1 as k_cell_count.
struct CountFold {
  int n_blocks;                // 0 = none; else a multiple of 8
  const float *pos;            // caller positions
  int stride;
  const int *rep_of;
  int n_cells;
  uint32_t *keys, *count;
};

struct ForceArgs {
  int n;
  int n_types;
  int cap;
  const uint4 *pos4;
  const float4 *res4;          // FP64 form: sub-lattice residuals (sorted order)
  const float *rcut;
  const int *nnb;
  const int *list;
  size_t list_len;             // ints allocated for list
  const RepInfo *reps;
  const PairParam *params;     // n_types^2
  const PairParamD *paramsd;   // n_types^2 (FP64 form)
  const double2 *tail;         // per-atom-index table (F_z, E) or NULL
  const int *tail_idx;         // per sorted atom, -1 = none
  const double2 *tail_atom;    // exact mode: per sorted atom (F_z, E) or NULL
  int64_t *fx, *fy, *fz;       // fixed-point output (sorted order)
  double *fxd, *fyd, *fzd;     // FP64 output (sorted order)
  double *atom_stats;          // optional per-atom (E, Wxx, Wyy, Wzz)
  unsigned *err_flag;
  // perturbation step counter: one thread writes *ctr_out = *ctr_in + 1
  // (ctr_in == ctr_out, or two slots when the launch refreshes the next
  // step from *ctr_in); NULL = none
  const uint64_t *ctr_in;
  uint64_t *ctr_out;
  const int2 *wave_len;        // per force wavefront: (longest list, total) from a
                               // list build of the same tiling, or NULL
  int xcd_remap;               // 1: contiguous block ranges per XCD
  int has_shift;               // some pair type has a nonzero energy shift
  int persist_waves;           // mixed form: > 0 = persistent launch of that many
                               // wavefronts, each taking a contiguous chunk of tiles
  NextRefresh next;            // the next step's refresh (perturbation mode)
  CountFold count;             // the next rebuild's cell histogram, or none
};
// The cell-window form of the mixed force evaluation (no list): tiles of
// up to kCellTile sorted atoms of one cell column (tile_off / tiles_per_col
// from launch_tile_map after the rebuild), kWave / kCellTile lanes per atom,
// candidates from the rebuild's cell offsets, rebuild-time cells from
// backup_s.
constexpr int kCellTile = 16;
struct CellWindow {
  int n_cols, n_rep;
  const uint32_t *tile_off, *tiles_per_col;
  const int *start;
  const uint4 *backup_s;
  float skin;
};
void launch_force_cells(bool stats, bool tail, const ForceArgs &a, const CellWindow &c,
                        hipStream_t st);
// unroll: list rows per group of the mixed kernel (1, 2 or 4)
void launch_force(int precision, int lanes, int unroll, int block, bool stats,
                  bool tail,
                  const ForceArgs &a, hipStream_t st);
int force_num_waves(int n, int lanes);
// stats[4 r + c] = sum over replica r's atoms of atom_stats[4 s + c];
// part holds n_rep * kStatBlocks * 4 doubles of chunk partials
constexpr int kStatBlocks = 32;
void launch_stats_reduce(int n_rep, const RepInfo *reps,
                         const double *atom_stats, double *part, double *stats,
                         hipStream_t st);
// counts[r] += pairs of replica r's lists with d^2 < rc_i^2 (counts zeroed
// by the caller)
void launch_count_in_rc(int n, int lanes, int cap, const uint4 *pos4,
                        const float *rcut, const int *nnb, const int *list,
                        const RepInfo *reps, unsigned long long *counts,
                        hipStream_t st);
void launch_export_forces(int n, const uint32_t *perm, const int64_t *fx,
                          const int64_t *fy, const int64_t *fz,
                          const double *fxd, const double *fyd,
                          const double *fzd, double *out, int64_t *out_fix,
                          hipStream_t st);
void launch_add_forces_f32(int n, const uint32_t *perm, const int64_t *fx,
                           const int64_t *fy, const int64_t *fz,
                           const double *fxd, const double *fyd,
                           const double *fzd, float *ox, float *oy, float *oz,
                           hipStream_t st);
void launch_export_atom_stats(int n, const uint32_t *perm, const double *src,
                              double *dst, hipStream_t st);

void launch_tail_table(int n_entries, const RepInfo *reps,
                       const int *entry_rep, int n_types,
                       const PairParamD *params, const float *rc_table,
                       const double *rho, double2 *tail, hipStream_t st);
void launch_tail_exact(int n, int n_types, const uint4 *pos4,
                       const float *rcut, const RepInfo *reps,
                       const PairParamD *params, const double *rho,
                       double2 *out, hipStream_t st);

struct AdjointArgs {
  int n, n_types, cap, n_rep;
  size_t list_len;         // ints allocated for list
  const int2 *wave_len;    // per wavefront (longest list, total) of the same tiling, or NULL
  const uint4 *pos4;
  const float *rcut;
  const int *nnb;
  const int *list;
  const uint32_t *perm;
  const RepInfo *reps;
  const PairParam *params;
  const double *dLdE;      // per replica (device)
  const double *dLdF;      // caller order [N][3] or NULL
  const int *tail_idx;
  int use_tail;
  long long *wave_part;    // value-major: [4 * n_types^2][launched waves], scale 2^32
  int64_t *tail_w;         // per tail entry: (sum g_z, sum dLdE), scale 2^32
};
int adjoint_grid_waves(int n, int lanes);
// pair part: out[4*npt], the integer totals (scale 2^32) of the four pair
// terms per ordered type pair, term-major (k_adjoint.hip); with them
// d/d eps = 4 T1 - 24 T2 and d/d sigma = eps / sigma (24 T3 - 144 T4)
void launch_adjoint_pairs(const AdjointArgs &a, int precision, int lanes,
                          long long *out, hipStream_t st);
// tail weights (sum of dLdF_z, sum of dLdE per tail entry, scale 2^32) into
// tail_w, exact; scratch holds tail_w_scratch(n_entries) integers
size_t tail_w_scratch(int n_entries);
void launch_tail_weights(int n, int n_entries, const int *tail_idx, const uint32_t *perm,
                         const uint4 *pos4, const double *dLdF, const double *dLdE,
                         long long *scratch, int64_t *tail_w, hipStream_t st);
// tail part: g_rc per rc lattice entry, out[2*npt] as above
void launch_adjoint_tail(int n_bins, const int *entry_rep, const RepInfo *reps,
                         int n_types, const PairParamD *params,
                         const float *rc_table, const double *rho,
                         const int64_t *tail_w, double *contrib, double *g_rc,
                         double *out, hipStream_t st);

void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);
void launch_fill_i64(int64_t *p, int64_t v, size_t n, hipStream_t st);
void launch_perturb(int n, float *pos, int stride, float sigma, uint64_t seed,
                    uint64_t step, hipStream_t st);

}  // namespace sre
