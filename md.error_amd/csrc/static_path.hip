// static_path.hip -- one force evaluation that allocates no device memory.
//
// sre_forces_static runs the engine's own kernels (cell sort, cutoff
// assignment, list build, mixed-precision force kernel, statistic
// reduction, force export) for one periodic system, with every buffer they
// need carved out of one file-scope device array, `scratch`, instead of
// allocations.  It is for callers that cannot call hipMalloc (code inside
// stream capture, serving loops, harnesses that forbid it).  The array is
// held by each GPU the library runs on until the process ends; its address
// on a device is looked up once (hipGetSymbolAddress) and only that address
// is handed to kernels and copies.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race.
#include <mutex>
#include <vector>

#include "sre_host.h"
#include "sre_internal.h"

namespace sre {

constexpr size_t kScratchBytes = (size_t)40 << 20;   // 40 MiB per GPU
constexpr int kStaticMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kStaticMaxDevices];

// Device address of this device's scratch array (looked up on the first call
// on the device, with that device current).
static sre_status scratch_address(int dev, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p)
      return fail(SRE_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(r));
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  return SRE_OK;
}

// Bump carving of the scratch array: 256-byte aligned pieces; `used` may
// exceed the array (then nothing is launched: the caller checks fits()).
struct Carve {
  unsigned char *base;
  size_t used = 0;
  template <typename T>
  T *take(size_t count) {
    used = (used + 255) & ~(size_t)255;
    T *p = reinterpret_cast<T *>(base + used);
    used += (count ? count : 1) * sizeof(T);
    return p;
  }
  bool fits() const { return used <= kScratchBytes; }
};

}  // namespace sre

using namespace sre;

extern "C" sre_status sre_forces_static(int n_atoms, const float *d_xyz, int stride,
                                        const int32_t *d_types, const float box[3],
                                        int n_types, const float *eps, const float *sigma,
                                        int nz, const float *rc_z, float skin,
                                        void *stream, double *d_forces, double *stats) {
  if (n_atoms < 1 || !d_xyz || stride < 3 || !box || n_types < 1 ||
      n_types > kMaxTypes || !eps || !sigma || nz < 1 || !rc_z || !(skin >= 0.f) ||
      !d_forces)
    return fail(SRE_ERR_INVALID, "sre_forces_static: bad arguments");
  if (n_atoms > SRE_STATIC_MAX_ATOMS)
    return fail(SRE_ERR_INVALID, "sre_forces_static: %d atoms exceed the static limit %d",
                n_atoms, SRE_STATIC_MAX_ATOMS);
  if (nz > SRE_STATIC_MAX_RC_BINS)
    return fail(SRE_ERR_INVALID, "sre_forces_static: %d rc bins exceed the static limit %d",
                nz, SRE_STATIC_MAX_RC_BINS);
  for (int a = 0; a < n_types; ++a)
    for (int b = 0; b < n_types; ++b) {
      const int ab = a * n_types + b, ba = b * n_types + a;
      if (eps[ab] != eps[ba] || sigma[ab] != sigma[ba])
        return fail(SRE_ERR_INVALID, "pair parameters must be symmetric (%d,%d)", a, b);
      if (!(sigma[ab] > 0.f) || !(eps[ab] >= 0.f))
        return fail(SRE_ERR_INVALID, "need sigma > 0 and eps >= 0 (%d,%d)", a, b);
    }
  for (int k = 0; k < nz; ++k)
    if (!(rc_z[k] > 0.f))
      return fail(SRE_ERR_INVALID, "rc bin %d must be > 0", k);
  for (int d = 0; d < 3; ++d)
    if (!(box[d] > 0.f)) return fail(SRE_ERR_INVALID, "box edge %d must be > 0", d);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int N = n_atoms;
  if (d_types) {    // device data indexes the parameter tables: check it here
    std::vector<int32_t> t(N);
    HIP_TRY(hipMemcpyAsync(t.data(), d_types, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < N; ++i)
      if (t[i] < 0 || t[i] >= n_types)
        return fail(SRE_ERR_INVALID, "type of atom %d is %d, outside [0,%d)", i, t[i],
                    n_types);
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kStaticMaxDevices)
    return fail(SRE_ERR_INVALID, "device %d beyond the static table", dev);
  unsigned char *base = nullptr;
  sre_status s = scratch_address(dev, &base);
  if (s != SRE_OK) return s;

  // geometry of the one replica, as sre_set_system and the rebuild set it
  RepInfo R{};
  for (int d = 0; d < 3; ++d) {
    R.L[d] = box[d];
    R.q[d] = ldexpf(R.L[d], -32);
    R.invq[d] = 4294967296.0 / (double)R.L[d];
    R.qd[d] = ldexp((double)R.L[d], -32);
  }
  R.atom_off = 0;
  R.n_atoms = N;
  R.rc_off = 0;
  R.rc_n[0] = R.rc_n[1] = 1;
  R.rc_n[2] = nz;
  R.tail_off = R.rho_off = R.fc_off = -1;
  R.rc_min = R.rc_max = rc_z[0];
  for (int k = 1; k < nz; ++k) {
    R.rc_min = std::min(R.rc_min, rc_z[k]);
    R.rc_max = std::max(R.rc_max, rc_z[k]);
  }
  for (int d = 0; d < 3; ++d)
    if (!(2.f * (R.rc_max + skin) < R.L[d]))
      return fail(SRE_ERR_BOX_TOO_SMALL,
                  "rc_max + skin = %g is not below half the box edge %d (%g)",
                  (double)(R.rc_max + skin), d, (double)R.L[d]);
  // the engine's default tuning: 4 lanes per atom (build and force), cells
  // of 1.0 x 0.35 (rc_min + skin), 256-thread force blocks, 2 rows per group
  const int T = 4, U = 2, block = 256;
  choose_cell_grid(R, skin, 1.0f, 0.35f);
  R.cell_off = 0;
  R.col_off = 0;
  const long long cells = (long long)R.nc[0] * R.nc[1] * R.nc[2];
  int key_bits = 1;
  while ((1ll << key_bits) < cells) ++key_bits;
  if (key_bits > 32 || radix_passes(key_bits) > 4)
    return fail(SRE_ERR_CELL, "radix sort: key of %d bits unsupported", key_bits);
  const int cap = SRE_STATIC_MAX_LIST;
  const int A = kWave / T;
  const size_t waves = (size_t)(N + A - 1) / A;
  const size_t list_len = waves * (size_t)cap * A + (size_t)kListPadRows * kWave;
  const int n_parts = build_part_count(0, N, T, 0);
  const size_t cc = radix_counts_capacity(N, key_bits);

  Carve c{base};
  RepInfo *d_reps = c.take<RepInfo>(1);
  PairParam *d_params = c.take<PairParam>(kMaxPairTypes);
  PairParamD *d_paramsd = c.take<PairParamD>(kMaxPairTypes);
  float *d_rc = c.take<float>(nz);
  int *d_rep_of = c.take<int>(N);
  int *d_type0 = d_types ? nullptr : c.take<int>(N);
  SortBufs sb{};
  sb.key_a = c.take<uint32_t>(N);
  sb.val_a = c.take<uint32_t>(N);
  sb.key_b = c.take<uint32_t>(N);
  sb.val_b = c.take<uint32_t>(N);
  sb.tkey = c.take<uint32_t>(N);
  sb.tval = c.take<uint32_t>(N);
  sb.counts = c.take<uint32_t>(cc);
  sb.offsets = c.take<uint32_t>(cc);
  sb.scan_tmp = c.take<uint32_t>(cc / 4096 + 8);
  int *d_start = c.take<int>((size_t)cells + 1);
  uint4 *d_pos4 = c.take<uint4>(N);
  uint4 *d_backup = c.take<uint4>(N);
  uint32_t *d_inv_perm = c.take<uint32_t>(N);
  float *d_rcut = c.take<float>(N);
  int *d_tail_idx = c.take<int>(N);
  int *d_nnb = c.take<int>(N);
  int *d_list = c.take<int>(list_len);
  int2 *d_build_part = c.take<int2>(n_parts);
  unsigned long long *d_counters = c.take<unsigned long long>(3);
  int64_t *d_fx = c.take<int64_t>(N);
  int64_t *d_fy = c.take<int64_t>(N);
  int64_t *d_fz = c.take<int64_t>(N);
  double *d_atom_stats = c.take<double>((size_t)4 * N);
  double *d_stat_part = c.take<double>((size_t)4 * kStatBlocks);
  double *d_stats = c.take<double>(4);
  unsigned *d_flags = c.take<unsigned>(4);
  if (!c.fits())
    return fail(SRE_ERR_INVALID,
                "sre_forces_static: the system needs %zu bytes of scratch (limit %zu; "
                "%lld cells)", c.used, kScratchBytes, cells);

  // parameters (host -> scratch), cleared words, then the rebuild sequence
  std::vector<PairParam> p(n_types * n_types);
  std::vector<PairParamD> pd(n_types * n_types);
  for (int a = 0; a < n_types * n_types; ++a)
    make_pair_params(eps[a], sigma[a], 0.0, &p[a], &pd[a]);
  HIP_TRY(hipMemcpyAsync(d_reps, &R, sizeof(RepInfo), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_params, p.data(), sizeof(PairParam) * p.size(),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_paramsd, pd.data(), sizeof(PairParamD) * pd.size(),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_rc, rc_z, sizeof(float) * nz, hipMemcpyHostToDevice, st));
  launch_fill_u32(reinterpret_cast<uint32_t *>(d_rep_of), 0u, N, st);
  if (d_type0) launch_fill_u32(reinterpret_cast<uint32_t *>(d_type0), 0u, N, st);
  launch_fill_u32(d_flags, 0u, 4, st);
  const int *types = d_types ? d_types : d_type0;
  if (launch_radix_sort(N, key_bits, d_xyz, stride, d_rep_of, d_reps, sb, st) != 0)
    return fail(SRE_ERR_CELL, "radix sort rejected %d key bits", key_bits);
  launch_gather_sorted(N, (int)cells, sb.key_a, sb.val_a, d_xyz, stride, types, d_rep_of,
                       d_reps, d_rc, n_types, d_start, d_pos4, d_backup, d_inv_perm,
                       nullptr, d_rcut, d_tail_idx, nullptr, st);
  launch_build_list(T, T, N, d_pos4, d_rcut, d_reps, d_start, skin, cap, d_list, d_nnb,
                    d_build_part, st);
  launch_rebuild_info(n_parts, d_build_part, d_counters, 0, nullptr, 0, nullptr, st);
  HIP_TRY(hipGetLastError());
  unsigned long long info[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(info, d_counters, sizeof info, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (info[0] > (unsigned long long)cap)
    return fail(SRE_ERR_LIST_OVERFLOW,
                "sre_forces_static: an atom has %llu list entries, above the static "
                "capacity %d", info[0], cap);

  ForceArgs a{};
  a.n = N;
  a.n_types = n_types;
  a.cap = cap;
  a.pos4 = d_pos4;
  a.rcut = d_rcut;
  a.nnb = d_nnb;
  a.list = d_list;
  a.list_len = list_len;
  a.reps = d_reps;
  a.params = d_params;
  a.paramsd = d_paramsd;
  a.tail_idx = d_tail_idx;
  a.fx = d_fx;
  a.fy = d_fy;
  a.fz = d_fz;
  a.atom_stats = d_atom_stats;
  a.err_flag = d_flags + 1;
  a.xcd_remap = 1;
  launch_force(SRE_PREC_MIXED, T, U, block, true, false, a, st);
  launch_export_forces(N, sb.val_a, d_fx, d_fy, d_fz, nullptr, nullptr, nullptr, d_forces,
                       nullptr, st);
  launch_stats_reduce(1, d_reps, d_atom_stats, d_stat_part, d_stats, st);
  HIP_TRY(hipGetLastError());
  double host_stats[4];
  unsigned flags[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(host_stats, d_stats, sizeof host_stats, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(flags, d_flags, sizeof flags, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (flags[1] & 1u)
    return fail(SRE_ERR_FORCE_RANGE,
                "pair forces beyond the exact fixed-point range (overlapping atoms?)");
  if (stats)
    for (int k = 0; k < 4; ++k) stats[k] = host_stats[k];
  return SRE_OK;
}
