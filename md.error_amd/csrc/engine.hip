// engine.hip -- host side of the short-range engine and its C ABI.
//
// One sre_engine bundles what the reference spreads over MDSystem (device
// data, H/include/MDSystem.h:43-78), AssignRCut (H/include/AssignRCut.h:10-42),
// CellList (H/include/CellList_interface.h:18-73), NeighborList
// (H/include/NeighborList_interface.h:20-82), InteractionEngine
// (H/include/InteractionEngine_interface.h:17-155), MDStatistic
// (H/include/Statistic_interface.h:30-117) and Displacement_max
// (H/include/Displacement_interface.h:15-43).  Unlike the reference it keeps
// no global state (no __constant__ tables, no texture bindings), so several
// engines can coexist; each launches only on its own stream.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sre.h"
#include "sre_host.h"
#include "sre_internal.h"

using namespace sre;

// ------------------------------------------------------------ error state --

static thread_local std::string g_err;

sre_status sre::fail(sre_status s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

// ----------------------------------------------------------------- engine --

constexpr int kStepGraphSlots = 4;   // cached step graphs (flags, step count)
constexpr int kMaxGraphSteps = 64;   // steps per captured graph
constexpr int kRebuildLog = 1024;    // rebuild-log entries (segments per queue)
// internal step-graph flag: the K steps are MD steps (thermostat factors,
// kick + drift refresh, force, kick) and SRE_STATS applies to the last only
constexpr unsigned kGraphMD = 1u << 16;

struct sre_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int n_types = 1;
  float skin = 0.3f;
  int precision = SRE_PREC_MIXED;
  sre_tuning tune{4, 4, 256, 1.0f, 0.35f, 0, 1, 2, 0, 0, 0};
  int n_simd = 1024;               // SIMDs of the device (4 per CU)

  // system
  int n_rep = 0, n_atoms = 0;
  std::vector<int> rep_n;
  std::vector<float> box;
  std::vector<RepInfo> reps;
  std::vector<RepInfo> reps_uploaded;  // device copy of reps, to skip unchanged uploads
  RepInfo *d_reps = nullptr;
  int *d_rep_of = nullptr;
  int *d_types = nullptr;
  std::vector<int> h_types;
  bool types_set = false;

  // force field
  std::vector<float> eps, sig, shift;
  bool has_shift = false;
  PairParam *d_params = nullptr;
  PairParamD *d_paramsd = nullptr;
  bool params_set = false;

  // rc lattices
  std::vector<std::vector<float>> rc_tab;
  std::vector<int> rc_dims;  // 3 per replica
  float *d_rc = nullptr;
  int rc_total = 0;
  bool rc_dirty = true;

  // density profiles and tail tables
  std::vector<std::vector<double>> rho;
  std::vector<int> rho_nz;
  double *d_rho = nullptr;
  double2 *d_tail = nullptr;
  int *d_tail_entry_rep = nullptr;
  int tail_bins = 0;
  double2 *d_tail_atom = nullptr;
  int tail_mode = SRE_TAIL_BIN;
  // grid force correction, per replica f[3 * cell + a]; concatenated float4
  std::vector<std::vector<float>> fc_tab;
  std::vector<int> fc_dims;
  float4 *d_fc = nullptr;
  size_t fc_cap = 0;
  int fc_cells = 0;
  bool tail_dirty = true;          // profile/rc layout changed: rebuild tables
  bool tail_values_dirty = false;  // only the LJ parameters changed

  // positions
  const float *d_pos = nullptr;
  int pos_stride = 4;
  float *d_pos_own = nullptr;
  size_t pos_own_floats = 0;

  // cells and sort
  SortBufs sb{};
  size_t counts_cap = 0;
  uint32_t *d_cell_count = nullptr;  // counting sort: zero between rebuilds
  size_t cell_count_cap = 0;
  int key_bits = 1;
  int n_cells = 0;
  int *d_start = nullptr;
  int start_cap = 0;
  int n_cols = 0;                  // (x, y) cell columns over all replicas
  uint32_t *d_tiles_per_col = nullptr, *d_tile_off = nullptr, *d_tile_tmp = nullptr;
  int tile_cap = 0;
  uint4 *d_pos4 = nullptr, *d_backup = nullptr;  // backup: caller order
  uint32_t *d_inv_perm = nullptr;  // caller order -> sorted slot
  float4 *d_res4 = nullptr;   // FP64 form only
  float *d_rcut = nullptr;
  int *d_tail_idx = nullptr;
  // rc per atom in caller order, set by sre_assign_rcut; once it has been
  // called, rebuilds carry these values instead of reading the table
  float *d_rc_atom = nullptr;
  bool rc_explicit = false;

  // neighbor list
  int cap = 0;
  int list_lanes = 0;
  size_t list_size = 0;
  int *d_list = nullptr;
  int *d_nnb = nullptr;
  // [0] longest list, [1] total length, [2] low word: max displacement^2
  unsigned long long *d_counters = nullptr;
  uint64_t *h_info = nullptr;               // pinned readback (4 words)
  double *h_stats = nullptr;                // pinned: 4 per replica + the force-range flag
  size_t h_stats_cap = 0;
  // rebuild log: every launch of the captured rebuild writes its (longest,
  // total, displacement^2 bits) to entry (launch number % kRebuildLog) of
  // this pinned host array, mapped for the device (d_log); d_log_idx is the
  // device's launch counter and log_launches the host's count of the same
  // launches, so a caller reads an entry after one synchronisation, with no
  // copy queued per rebuild
  uint64_t *h_log = nullptr;
  unsigned long long *d_log = nullptr;
  unsigned *d_log_idx = nullptr;
  uint64_t log_launches = 0;
  float last_disp2 = 0.f;                   // max displacement^2 before the last rebuild
  int max_count = 0;
  int64_t n_list = 0;

  // outputs
  int64_t *d_fx = nullptr, *d_fy = nullptr, *d_fz = nullptr;
  double *d_fxd = nullptr, *d_fyd = nullptr, *d_fzd = nullptr;
  double *d_stats = nullptr;       // per replica (E, Wxx, Wyy, Wzz)
  double *d_stat_part = nullptr;   // per replica and chunk, 4 partial sums
  unsigned long long *d_nin = nullptr;  // per replica pairs inside rc
  double *d_atom_stats = nullptr;
  double *d_tmp_out = nullptr;
  unsigned *d_flags = nullptr;  // [0] max displacement^2 bits, [1] error bits
  uint64_t *d_step_ctr = nullptr;  // perturbation step counter (graph mode)
  float pert_sigma = 0.f;
  uint64_t pert_seed = 0;
  float *d_skin_part = nullptr;    // per update wavefront / force tile: max displacement^2
  // the second coordinate buffer of graphs that refresh the next step inside
  // the force launch (and its FP64-form residuals)
  uint4 *d_pos4_alt = nullptr;
  float4 *d_res4_alt = nullptr;
  // the refresh inputs of those graphs in cell order, written by each
  // rebuild: the rebuild-time lattice positions and the perturbation anchor
  uint4 *d_backup_s = nullptr;
  float4 *d_anchor_s = nullptr;
  int2 *d_build_part = nullptr;    // per build wavefront: (longest, total)
  int2 *d_info_slices = nullptr;   // scratch of the two-stage partial reduction
  int build_part_cap = 0;
  float *d_anchor = nullptr;       // anchored perturbation: positions at set-up
  size_t anchor_cap = 0, anchor_count = 0;
  int anchor_stride = 0;
  bool pert_anchored = false;
  bool stats_valid = false;
  bool atom_stats_valid = false;
  // the force kernel leaves per-wave partial sums; their fixed-order
  // reduction runs when the statistics are read
  bool stats_pending = false;
  bool forces_valid = false;

  bool built = false;
  bool system_set = false;
  int graph_epoch = 0;
  double *d_adj = nullptr;         // adjoint workspace (grown on demand)
  size_t adj_cap = 0;
  int64_t *d_adj_tw = nullptr;
  size_t adj_tw_cap = 0;
  // step graphs: K steps of [refresh + force] each, cached by (flags, K)
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    unsigned flags = 0;
    int steps = 0;
    int epoch = -1;
    uint64_t last_use = 0;
  };
  StepGraph step_graphs[kStepGraphSlots];
  uint64_t graph_uses = 0;
  // rebuild graph: the device part of a rebuild, cached by its arguments
  hipGraphExec_t rebuild_exec = nullptr;
  std::vector<unsigned char> rebuild_sig;
  // segment graphs: m x [steps graph steps + rebuild] as one graph for
  // m = 2^k, k < kSegGraphs (slot k), all captured together for one (flags,
  // steps) and the step and rebuild arguments
  static constexpr int kSegGraphs = 5;
  hipGraphExec_t seg_exec[kSegGraphs] = {};
  unsigned seg_flags = 0;
  int seg_steps = 0, seg_epoch = -1;
  std::vector<unsigned char> seg_sig;

  // MD integrator (caller order, one replica)
  std::vector<float> mass;         // per type (default 1)
  float *d_mass = nullptr;
  float *d_vel = nullptr;          // 3 per atom
  float *d_fcall = nullptr;        // force of the last kick, 3 per atom
  double *d_md_part = nullptr;     // kick partial sums (K, Px, Py, Pz) per block
  MDState *d_md = nullptr;
  MDState md{};                    // host copy of the parameters
  bool vel_set = false, md_set = false;
  int md_since_rebuild = 0;        // MD steps since the last rebuild
  // adaptive stretches: state at the start of the running stretch
  float *d_snap_pos = nullptr, *d_snap_vel = nullptr, *d_snap_fcall = nullptr;
  double *d_snap_part = nullptr;
  MDState *d_snap_md = nullptr;
  size_t snap_pos_cap = 0;
  int md_adapt_k = 8;              // steps of the next adaptive stretch (reset by sre_set_system)
  int md_rollbacks = 0;            // stretches redone after a skin violation

  ~sre_engine();
};

sre_engine::~sre_engine() {
  for (StepGraph &g : step_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (rebuild_exec) (void)hipGraphExecDestroy(rebuild_exec);
  for (hipGraphExec_t &g : seg_exec)
    if (g) (void)hipGraphExecDestroy(g);
  dfree(d_reps); dfree(d_rep_of); dfree(d_types); dfree(d_params);
  dfree(d_paramsd); dfree(d_rc); dfree(d_rho); dfree(d_tail);
  dfree(d_tail_entry_rep); dfree(d_tail_atom); dfree(d_pos_own);
  dfree(sb.key_a); dfree(sb.val_a); dfree(sb.key_b); dfree(sb.val_b);
  dfree(sb.tkey); dfree(sb.tval); dfree(sb.counts); dfree(sb.offsets);
  dfree(sb.scan_tmp); dfree(d_cell_count); dfree(d_start); dfree(d_pos4); dfree(d_backup);
  dfree(d_inv_perm);
  dfree(d_rcut); dfree(d_rc_atom); dfree(d_tail_idx); dfree(d_list); dfree(d_nnb); dfree(d_res4);
  dfree(d_tiles_per_col); dfree(d_tile_off); dfree(d_tile_tmp);
  dfree(d_counters); dfree(d_fx); dfree(d_fy); dfree(d_fz); dfree(d_fxd);
  if (h_info) (void)hipHostFree(h_info);
  if (h_stats) (void)hipHostFree(h_stats);
  if (h_log) (void)hipHostFree(h_log);
  dfree(d_log_idx);
  dfree(d_fyd); dfree(d_fzd); dfree(d_nin); dfree(d_stats); dfree(d_stat_part);
  dfree(d_atom_stats); dfree(d_tmp_out); dfree(d_flags); dfree(d_step_ctr);
  dfree(d_anchor); dfree(d_skin_part); dfree(d_build_part); dfree(d_info_slices);
  dfree(d_pos4_alt); dfree(d_res4_alt); dfree(d_backup_s); dfree(d_anchor_s);
  dfree(d_adj); dfree(d_adj_tw);
  dfree(d_mass); dfree(d_vel); dfree(d_fcall); dfree(d_md_part); dfree(d_md);
  dfree(d_fc);
  dfree(d_snap_pos); dfree(d_snap_vel); dfree(d_snap_fcall); dfree(d_snap_part);
  dfree(d_snap_md);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

static bool pow2_lanes(int t) {
  return t == 1 || t == 2 || t == 4 || t == 8 || t == 16;
}

// ------------------------------------------------------------- lifetime --

extern "C" const char *sre_last_error(void) { return g_err.c_str(); }
extern "C" const char *sre_version(void) { return "sre 0.1 gfx950"; }

extern "C" sre_status sre_create(sre_engine **out, const sre_options *opt) {
  if (!out || !opt) return fail(SRE_ERR_INVALID, "sre_create: null argument");
  *out = nullptr;
  if (opt->n_types < 1 || opt->n_types > kMaxTypes)
    return fail(SRE_ERR_INVALID, "n_types must be in [1, %d]", kMaxTypes);
  if (!(opt->skin >= 0.f))
    return fail(SRE_ERR_INVALID, "skin must be >= 0");
  if (opt->precision != SRE_PREC_MIXED && opt->precision != SRE_PREC_FP64)
    return fail(SRE_ERR_INVALID, "unknown precision %d", opt->precision);
  sre_engine *e = new sre_engine();
  e->n_types = opt->n_types;
  e->skin = opt->skin;
  e->precision = opt->precision;
  // measured best lanes per atom of the force kernel on the 64k slab: the
  // FP64 form's longer pair arithmetic prefers more lanes
  if (e->precision == SRE_PREC_FP64) e->tune.force_lanes = 8;
  if (opt->device >= 0) {
    hipError_t r = hipSetDevice(opt->device);
    if (r != hipSuccess) {
      delete e;
      return fail(SRE_ERR_HIP, "hipSetDevice(%d): %s", opt->device,
                  hipGetErrorString(r));
    }
  }
  if (hipGetDevice(&e->device) != hipSuccess) {
    delete e;
    return fail(SRE_ERR_HIP, "hipGetDevice failed (no GPU?)");
  }
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) ==
            hipSuccess && cus > 0)
      e->n_simd = 4 * cus;
  }
  if (opt->stream) {
    e->stream = (hipStream_t)opt->stream;
  } else {
    hipError_t r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r != hipSuccess) {
      delete e;
      return fail(SRE_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(r));
    }
    e->own_stream = true;
  }
  e->eps.assign(e->n_types * e->n_types, 1.f);
  e->sig.assign(e->n_types * e->n_types, 1.f);
  e->shift.assign(e->n_types * e->n_types, 0.f);
  hipError_t r = dalloc(&e->d_flags, 4);
  if (r == hipSuccess) r = hipMemsetAsync(e->d_flags, 0, 4 * sizeof(unsigned), e->stream);
  if (r == hipSuccess) r = dalloc(&e->d_counters, 3);
  if (r == hipSuccess) r = hipMemsetAsync(e->d_counters, 0, 3 * sizeof(uint64_t), e->stream);
  if (r == hipSuccess) r = hipHostMalloc((void **)&e->h_info, 4 * sizeof(uint64_t), hipHostMallocDefault);
  if (r == hipSuccess)
    r = hipHostMalloc((void **)&e->h_log, sizeof(uint64_t) * 3 * kRebuildLog,
                      hipHostMallocMapped | hipHostMallocCoherent);
  if (r == hipSuccess) r = hipHostGetDevicePointer((void **)&e->d_log, e->h_log, 0);
  if (r == hipSuccess) r = dalloc(&e->d_log_idx, 1);
  if (r == hipSuccess) r = hipMemsetAsync(e->d_log_idx, 0, sizeof(unsigned), e->stream);
  if (r == hipSuccess) r = dalloc(&e->d_params, kMaxPairTypes);
  if (r == hipSuccess) r = dalloc(&e->d_paramsd, kMaxPairTypes);
  if (r != hipSuccess) {
    delete e;
    return fail(SRE_ERR_HIP, "allocation failed: %s", hipGetErrorString(r));
  }
  *out = e;
  return SRE_OK;
}

extern "C" void sre_destroy(sre_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  delete e;
}

// ------------------------------------------------------ system description --

// The force-range flag (d_flags[1]) is set by a force kernel whose fixed-point
// sums left the exact range.  It describes the forces computed since it was
// last cleared: it is cleared (a stream-ordered fill, safe inside captured
// work) when it is reported, when the system is set and when new positions
// are uploaded, bound or restored from an MD snapshot.  A report also marks
// the forces and statistics invalid, so they cannot be read afterwards.
static void clear_force_flag(sre_engine *e) {
  launch_fill_u32(e->d_flags + 1, 0u, 1, e->stream);
}

static sre_status report_force_range(sre_engine *e) {
  clear_force_flag(e);
  e->forces_valid = e->stats_valid = e->atom_stats_valid = false;
  return fail(SRE_ERR_FORCE_RANGE,
              "pair forces beyond the exact fixed-point range (overlapping atoms?)");
}


extern "C" sre_status sre_set_system(sre_engine *e, int n_replicas,
                                     const int *n_atoms, const float *box) {
  if (!e || n_replicas < 1 || !n_atoms || !box)
    return fail(SRE_ERR_INVALID, "sre_set_system: bad arguments");
  HIP_TRY(hipSetDevice(e->device));
  long long total = 0;
  for (int r = 0; r < n_replicas; ++r) {
    if (n_atoms[r] < (n_replicas > 1 ? kWave : 1))
      return fail(SRE_ERR_INVALID,
                  "replica %d has %d atoms; batched replicas need >= %d", r,
                  n_atoms[r], kWave);
    for (int d = 0; d < 3; ++d)
      if (!(box[3 * r + d] > 0.f))
        return fail(SRE_ERR_INVALID, "replica %d: box edge %d must be > 0", r, d);
    total += n_atoms[r];
  }
  // pos4 is addressed with 32-bit byte offsets (16 B per atom)
  if (total >= (1ll << 27))
    return fail(SRE_ERR_INVALID, "too many atoms (%lld, limit 2^27 - 1)", total);
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->n_rep = n_replicas;
  e->n_atoms = (int)total;
  e->rep_n.assign(n_atoms, n_atoms + n_replicas);
  e->box.assign(box, box + 3 * n_replicas);
  e->reps.assign(n_replicas, RepInfo{});
  e->reps_uploaded.clear();
  int off = 0;
  for (int r = 0; r < n_replicas; ++r) {
    RepInfo &R = e->reps[r];
    for (int d = 0; d < 3; ++d) {
      R.L[d] = box[3 * r + d];
      R.q[d] = ldexpf(R.L[d], -32);
      R.invq[d] = 4294967296.0 / (double)R.L[d];
      R.qd[d] = ldexp((double)R.L[d], -32);
    }
    R.atom_off = off;
    R.n_atoms = n_atoms[r];
    R.tail_off = -1;
    R.rho_off = -1;
    R.fc_off = -1;
    off += n_atoms[r];
  }
  e->fc_tab.assign(n_replicas, std::vector<float>());
  e->fc_dims.assign(3 * n_replicas, 0);
  e->fc_cells = 0;
  // host-side replica map, uploaded once
  std::vector<int> rep_of(e->n_atoms);
  for (int r = 0; r < n_replicas; ++r)
    for (int i = 0; i < n_atoms[r]; ++i) rep_of[e->reps[r].atom_off + i] = r;
  const int N = e->n_atoms;
  dfree(e->d_rep_of); dfree(e->d_types); dfree(e->d_reps);
  dfree(e->d_pos4); dfree(e->d_backup); dfree(e->d_rcut); dfree(e->d_tail_idx);
  dfree(e->d_rc_atom);
  e->rc_explicit = false;
  dfree(e->d_inv_perm);
  dfree(e->d_res4);
  dfree(e->d_nnb); dfree(e->d_fx); dfree(e->d_fy); dfree(e->d_fz);
  dfree(e->d_fxd); dfree(e->d_fyd); dfree(e->d_fzd); dfree(e->d_atom_stats);
  dfree(e->d_stats); dfree(e->d_nin); dfree(e->d_tmp_out); dfree(e->d_tail_atom);
  dfree(e->d_stat_part);
  dfree(e->d_skin_part);
  dfree(e->d_pos4_alt); dfree(e->d_res4_alt); dfree(e->d_backup_s); dfree(e->d_anchor_s);
  dfree(e->d_vel); dfree(e->d_fcall); dfree(e->d_md_part);
  dfree(e->d_snap_pos); dfree(e->d_snap_vel); dfree(e->d_snap_fcall);
  dfree(e->d_snap_part);
  e->snap_pos_cap = 0;
  e->vel_set = false;
  // the integrator's state (3N degrees of freedom, chain masses) belongs to
  // the old system: the caller sets it again with sre_set_integrator
  e->md_set = false;
  e->md_since_rebuild = 0;
  e->md_adapt_k = 8;
  clear_force_flag(e);
  dfree(e->sb.key_a); dfree(e->sb.val_a); dfree(e->sb.key_b); dfree(e->sb.val_b);
  dfree(e->sb.tkey); dfree(e->sb.tval);
  HIP_TRY(dalloc(&e->d_rep_of, N));
  HIP_TRY(dalloc(&e->d_types, N));
  HIP_TRY(dalloc(&e->d_reps, n_replicas));
  HIP_TRY(dalloc(&e->d_pos4, N));
  HIP_TRY(dalloc(&e->d_backup, N));
  HIP_TRY(dalloc(&e->d_inv_perm, N));
  HIP_TRY(dalloc(&e->d_rcut, N));
  HIP_TRY(dalloc(&e->d_rc_atom, N));
  HIP_TRY(dalloc(&e->d_tail_idx, N));
  HIP_TRY(dalloc(&e->d_nnb, N));
  if (e->precision == SRE_PREC_MIXED) {
    HIP_TRY(dalloc(&e->d_fx, N));
    HIP_TRY(dalloc(&e->d_fy, N));
    HIP_TRY(dalloc(&e->d_fz, N));
  } else {
    HIP_TRY(dalloc(&e->d_res4, N));
    HIP_TRY(dalloc(&e->d_fxd, N));
    HIP_TRY(dalloc(&e->d_fyd, N));
    HIP_TRY(dalloc(&e->d_fzd, N));
  }
  HIP_TRY(dalloc(&e->d_atom_stats, (size_t)4 * N));
  HIP_TRY(dalloc(&e->d_stats, (size_t)4 * n_replicas));
  HIP_TRY(dalloc(&e->d_nin, n_replicas));
  HIP_TRY(dalloc(&e->d_stat_part, (size_t)4 * kStatBlocks * n_replicas));
  HIP_TRY(dalloc(&e->d_tmp_out, (size_t)4 * N));
  HIP_TRY(dalloc(&e->sb.key_a, N));
  HIP_TRY(dalloc(&e->sb.val_a, N));
  HIP_TRY(dalloc(&e->sb.key_b, N));
  HIP_TRY(dalloc(&e->sb.val_b, N));
  HIP_TRY(dalloc(&e->sb.tkey, N));
  HIP_TRY(dalloc(&e->sb.tval, N));
  HIP_TRY(dalloc(&e->d_skin_part, skin_part_count(N)));
  if (!e->d_info_slices) HIP_TRY(dalloc(&e->d_info_slices, 2 * kInfoSlices));
  launch_fill_u32((uint32_t *)e->d_skin_part, 0u, skin_part_count(N), e->stream);
  HIP_TRY(dalloc(&e->d_pos4_alt, N));
  HIP_TRY(dalloc(&e->d_backup_s, N));
  if (e->precision != SRE_PREC_MIXED) HIP_TRY(dalloc(&e->d_res4_alt, N));
  HIP_TRY(hipMemcpy(e->d_rep_of, rep_of.data(), sizeof(int) * N,
                    hipMemcpyHostToDevice));
  // defaults: every atom type 0, uniform rc lattice of 1 bin (set later)
  e->rc_tab.assign(n_replicas, std::vector<float>());
  e->rc_dims.assign(3 * n_replicas, 0);
  e->rho.assign(n_replicas, std::vector<double>());
  e->rho_nz.assign(n_replicas, 0);
  e->types_set = false;
  e->rc_dirty = true;
  e->tail_dirty = true;
  e->built = false;
  e->system_set = true;
  e->d_pos = nullptr;
  e->stats_valid = e->forces_valid = e->atom_stats_valid = false;
  e->graph_epoch++;
  return SRE_OK;
}

static sre_status md_update_total_mass(sre_engine *e);

// The host geometry of the only replica (position refresh argument), or NULL
static const RepInfo *rep0(const sre_engine *e) {
  return e->n_rep == 1 ? &e->reps[0] : nullptr;
}

extern "C" sre_status sre_set_types(sre_engine *e, const int32_t *types) {
  if (!e || !types || !e->system_set)
    return fail(SRE_ERR_INVALID, "sre_set_types: system not set");
  for (int i = 0; i < e->n_atoms; ++i)
    if (types[i] < 0 || types[i] >= e->n_types)
      return fail(SRE_ERR_INVALID, "type of atom %d is %d, outside [0,%d)", i,
                  types[i], e->n_types);
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipMemcpy(e->d_types, types, sizeof(int) * e->n_atoms,
                    hipMemcpyHostToDevice));
  e->h_types.assign(types, types + e->n_atoms);
  e->types_set = true;
  e->built = false;
  if (e->md_set) return md_update_total_mass(e);
  return SRE_OK;
}

static sre_status upload_params(sre_engine *e) {
  const int nt = e->n_types;
  std::vector<PairParam> p(nt * nt);
  std::vector<PairParamD> pd(nt * nt);
  for (int a = 0; a < nt * nt; ++a)
    make_pair_params(e->eps[a], e->sig[a], e->shift[a], &p[a], &pd[a]);
  HIP_TRY(hipMemcpy(e->d_params, p.data(), sizeof(PairParam) * nt * nt,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->d_paramsd, pd.data(), sizeof(PairParamD) * nt * nt,
                    hipMemcpyHostToDevice));
  return SRE_OK;
}

extern "C" sre_status sre_set_pair_params(sre_engine *e, const float *eps,
                                          const float *sigma,
                                          const float *shift) {
  if (!e || !eps || !sigma)
    return fail(SRE_ERR_INVALID, "sre_set_pair_params: null argument");
  const int nt = e->n_types;
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const int ab = a * nt + b, ba = b * nt + a;
      if (eps[ab] != eps[ba] || sigma[ab] != sigma[ba] ||
          (shift && shift[ab] != shift[ba]))
        return fail(SRE_ERR_INVALID, "pair parameters must be symmetric (%d,%d)", a, b);
      if (!(sigma[ab] > 0.f) || !(eps[ab] >= 0.f))
        return fail(SRE_ERR_INVALID, "need sigma > 0 and eps >= 0 (%d,%d)", a, b);
    }
  HIP_TRY(hipSetDevice(e->device));
  e->eps.assign(eps, eps + nt * nt);
  e->sig.assign(sigma, sigma + nt * nt);
  if (shift) e->shift.assign(shift, shift + nt * nt);
  else e->shift.assign(nt * nt, 0.f);
  const bool has_shift =
      std::any_of(e->shift.begin(), e->shift.end(), [](float v) { return v != 0.f; });
  if (has_shift != e->has_shift) {
    e->has_shift = has_shift;
    e->graph_epoch++;   // the captured force kernel variant depends on it
  }
  sre_status s = upload_params(e);
  if (s != SRE_OK) return s;
  e->params_set = true;
  e->tail_values_dirty = true;
  return SRE_OK;
}

extern "C" sre_status sre_set_rc_table(sre_engine *e, int replica, int nx,
                                       int ny, int nz, const float *rc) {
  if (!e || !e->system_set || replica < 0 || replica >= e->n_rep || !rc ||
      nx < 1 || ny < 1 || nz < 1)
    return fail(SRE_ERR_INVALID, "sre_set_rc_table: bad arguments");
  const size_t m = (size_t)nx * ny * nz;
  for (size_t k = 0; k < m; ++k)
    if (!(rc[k] > 0.f))
      return fail(SRE_ERR_INVALID, "rc lattice entry %zu must be > 0", k);
  e->rc_tab[replica].assign(rc, rc + m);
  e->rc_dims[3 * replica + 0] = nx;
  e->rc_dims[3 * replica + 1] = ny;
  e->rc_dims[3 * replica + 2] = nz;
  e->rc_dirty = true;
  e->tail_dirty = true;
  e->built = false;
  return SRE_OK;
}

// Concatenates the replicas' force grids into d_fc (float4 per cell), sets
// fc_off / fc_n and uploads the replica table (stream-ordered, synchronous).
static sre_status upload_force_grids(sre_engine *e) {
  std::vector<float4> all;
  for (int r = 0; r < e->n_rep; ++r) {
    RepInfo &R = e->reps[r];
    if (e->fc_tab[r].empty()) {
      R.fc_off = -1;
      continue;
    }
    R.fc_off = (int)all.size();
    for (int d = 0; d < 3; ++d) R.fc_n[d] = e->fc_dims[3 * r + d];
    const std::vector<float> &t = e->fc_tab[r];
    for (size_t c = 0; c < t.size() / 3; ++c)
      all.push_back(make_float4(t[3 * c], t[3 * c + 1], t[3 * c + 2], 0.f));
  }
  if (all.size() > e->fc_cap) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    dfree(e->d_fc);
    HIP_TRY(dalloc(&e->d_fc, all.size()));
    e->fc_cap = all.size();
    e->graph_epoch++;   // captured launches read d_fc by address
  }
  e->fc_cells = (int)all.size();
  if (!all.empty())
    HIP_TRY(hipMemcpyAsync(e->d_fc, all.data(), sizeof(float4) * all.size(),
                           hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_reps, e->reps.data(), sizeof(RepInfo) * e->n_rep,
                         hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->reps_uploaded = e->reps;
  return SRE_OK;
}

extern "C" sre_status sre_set_force_grid(sre_engine *e, int replica, int nx, int ny,
                                         int nz, const float *f) {
  if (!e || !e->system_set || replica < 0 || replica >= e->n_rep)
    return fail(SRE_ERR_INVALID, "sre_set_force_grid: bad replica");
  HIP_TRY(hipSetDevice(e->device));
  if (!f) {
    e->fc_tab[replica].clear();
    return upload_force_grids(e);
  }
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz > (1ll << 26))
    return fail(SRE_ERR_INVALID, "sre_set_force_grid: bad grid %d x %d x %d", nx, ny, nz);
  const size_t n = (size_t)3 * nx * ny * nz;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(f[i]))
      return fail(SRE_ERR_INVALID, "sre_set_force_grid: value %zu is not finite", i);
  e->fc_tab[replica].assign(f, f + n);
  e->fc_dims[3 * replica] = nx;
  e->fc_dims[3 * replica + 1] = ny;
  e->fc_dims[3 * replica + 2] = nz;
  return upload_force_grids(e);
}

extern "C" sre_status sre_set_density_profile(sre_engine *e, int replica,
                                              int nz, const double *rho) {
  if (!e || !e->system_set || replica < 0 || replica >= e->n_rep)
    return fail(SRE_ERR_INVALID, "sre_set_density_profile: bad arguments");
  if (!rho || nz <= 0) {
    e->rho[replica].clear();
    e->rho_nz[replica] = 0;
  } else {
    e->rho[replica].assign(rho, rho + (size_t)nz * e->n_types);
    e->rho_nz[replica] = nz;
  }
  e->tail_dirty = true;
  e->built = false;
  return SRE_OK;
}

extern "C" sre_status sre_set_tail_mode(sre_engine *e, int mode) {
  if (!e || (mode != SRE_TAIL_BIN && mode != SRE_TAIL_EXACT))
    return fail(SRE_ERR_INVALID, "sre_set_tail_mode: bad mode");
  e->tail_mode = mode;
  e->tail_dirty = true;
  e->built = false;
  return SRE_OK;
}

// --------------------------------------------------------------- positions --

extern "C" sre_status sre_upload_positions(sre_engine *e, const float *xyz,
                                           int stride) {
  if (!e || !e->system_set || !xyz || stride < 3)
    return fail(SRE_ERR_INVALID, "sre_upload_positions: bad arguments");
  HIP_TRY(hipSetDevice(e->device));
  const size_t nf = (size_t)e->n_atoms * stride;
  if (e->pos_own_floats < nf) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    dfree(e->d_pos_own);
    HIP_TRY(dalloc(&e->d_pos_own, nf));
    e->pos_own_floats = nf;
    e->graph_epoch++;
  }
  HIP_TRY(hipMemcpyAsync(e->d_pos_own, xyz, nf * sizeof(float),
                         hipMemcpyHostToDevice, e->stream));
  clear_force_flag(e);
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->d_pos != e->d_pos_own || e->pos_stride != stride) e->graph_epoch++;
  e->d_pos = e->d_pos_own;
  e->pos_stride = stride;
  return SRE_OK;
}

extern "C" sre_status sre_bind_positions(sre_engine *e, const void *d_xyz,
                                         int stride) {
  if (!e || !e->system_set || !d_xyz || stride < 3)
    return fail(SRE_ERR_INVALID, "sre_bind_positions: bad arguments");
  HIP_TRY(hipSetDevice(e->device));
  if (e->d_pos != d_xyz || e->pos_stride != stride) e->graph_epoch++;
  e->d_pos = (const float *)d_xyz;
  e->pos_stride = stride;
  clear_force_flag(e);
  return SRE_OK;
}

// ---------------------------------------------------------------- rebuild --

static int bitlen(unsigned v) {
  int b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

// Uploads rc lattices, chooses cell grids, uploads RepInfo.
static sre_status prepare_tables(sre_engine *e) {
  // rc lattices
  if (e->rc_dirty) {
    int total = 0;
    for (int r = 0; r < e->n_rep; ++r) {
      if (e->rc_tab[r].empty())
        return fail(SRE_ERR_INVALID, "replica %d has no rc table", r);
      total += (int)e->rc_tab[r].size();
    }
    std::vector<float> all;
    all.reserve(total);
    for (int r = 0; r < e->n_rep; ++r) {
      RepInfo &R = e->reps[r];
      R.rc_off = (int)all.size();
      for (int d = 0; d < 3; ++d) R.rc_n[d] = e->rc_dims[3 * r + d];
      R.rc_min = *std::min_element(e->rc_tab[r].begin(), e->rc_tab[r].end());
      R.rc_max = *std::max_element(e->rc_tab[r].begin(), e->rc_tab[r].end());
      all.insert(all.end(), e->rc_tab[r].begin(), e->rc_tab[r].end());
    }
    if (total > e->rc_total) {
      dfree(e->d_rc);
      HIP_TRY(dalloc(&e->d_rc, total));
      e->graph_epoch++;
    }
    e->rc_total = total;
    HIP_TRY(hipMemcpy(e->d_rc, all.data(), sizeof(float) * total,
                      hipMemcpyHostToDevice));
    e->rc_dirty = false;
  }
  // cell grids
  long long cells = 0, cols = 0;
  for (int r = 0; r < e->n_rep; ++r) {
    RepInfo &R = e->reps[r];
    const float rlmax = R.rc_max + e->skin;
    for (int d = 0; d < 3; ++d)
      if (!(2.f * rlmax < R.L[d]))
        return fail(SRE_ERR_BOX_TOO_SMALL,
                    "replica %d: rc_max + skin = %g is not below half the box "
                    "edge %d (%g)", r, rlmax, d, R.L[d]);
    choose_cell_grid(R, e->skin, e->tune.cell_xy, e->tune.cell_z);
    R.cell_off = (int)cells;
    R.col_off = (int)cols;
    cells += (long long)R.nc[0] * R.nc[1] * R.nc[2];
    cols += (long long)R.nc[0] * R.nc[1];
  }
  if (cells >= (1ll << 31) - 2)
    return fail(SRE_ERR_CELL, "cell grid too large (%lld cells)", cells);
  e->n_cells = (int)cells;
  e->n_cols = (int)cols;
  if (e->tile_cap < e->n_cols) {
    dfree(e->d_tiles_per_col);
    dfree(e->d_tile_off);
    dfree(e->d_tile_tmp);
    HIP_TRY(dalloc(&e->d_tiles_per_col, e->n_cols));
    HIP_TRY(dalloc(&e->d_tile_off, e->n_cols));
    HIP_TRY(dalloc(&e->d_tile_tmp, e->n_cols / 4096 + 8));
    e->tile_cap = e->n_cols;
  }
  e->key_bits = std::max(1, bitlen((unsigned)(cells - 1)));
  if (e->key_bits > 32) return fail(SRE_ERR_CELL, "cell key exceeds 32 bits");
  if (e->start_cap < e->n_cells + 1) {
    dfree(e->d_start);
    HIP_TRY(dalloc(&e->d_start, e->n_cells + 1));
    e->start_cap = e->n_cells + 1;
    e->graph_epoch++;
  }
  if (cell_count_capacity(e->n_cells) > e->cell_count_cap) {
    const size_t m = cell_count_capacity(e->n_cells);
    dfree(e->d_cell_count);
    HIP_TRY(dalloc(&e->d_cell_count, m));
    launch_fill_u32(e->d_cell_count, 0u, m, e->stream);
    e->cell_count_cap = m;
  }
  const size_t cc = radix_counts_capacity(e->n_atoms, e->key_bits);
  if (cc > e->counts_cap) {
    dfree(e->sb.counts);
    dfree(e->sb.offsets);
    dfree(e->sb.scan_tmp);
    HIP_TRY(dalloc(&e->sb.counts, cc));
    HIP_TRY(dalloc(&e->sb.offsets, cc));
    HIP_TRY(dalloc(&e->sb.scan_tmp, cc / 4096 + 8));
    e->counts_cap = cc;
  }
  // density profiles / tail tables
  if (e->tail_dirty) {
    std::vector<double> rho_all;
    std::vector<int> entry_rep;
    int bins = 0;
    for (int r = 0; r < e->n_rep; ++r) {
      RepInfo &R = e->reps[r];
      R.tail_off = -1;
      R.rho_off = -1;
      R.rho_nz = 0;
      if (e->rho_nz[r] > 0) {
        R.rho_off = (int)rho_all.size();
        R.rho_nz = e->rho_nz[r];
        rho_all.insert(rho_all.end(), e->rho[r].begin(), e->rho[r].end());
        if (R.rc_n[0] == 1 && R.rc_n[1] == 1) {
          R.tail_off = bins;
          for (int b = 0; b < R.rc_n[2]; ++b) entry_rep.push_back(r);
          bins += R.rc_n[2];
        } else {
          return fail(SRE_ERR_INVALID,
                      "replica %d: the tail correction needs an rc(z) table "
                      "(nx = ny = 1)", r);
        }
      }
    }
    dfree(e->d_rho);
    dfree(e->d_tail);
    dfree(e->d_tail_entry_rep);
    e->tail_bins = bins;
    if (bins > 0) {
      HIP_TRY(dalloc(&e->d_rho, rho_all.size()));
      HIP_TRY(hipMemcpy(e->d_rho, rho_all.data(), sizeof(double) * rho_all.size(),
                        hipMemcpyHostToDevice));
      HIP_TRY(dalloc(&e->d_tail, (size_t)bins * e->n_types));
      HIP_TRY(dalloc(&e->d_tail_entry_rep, bins));
      HIP_TRY(hipMemcpy(e->d_tail_entry_rep, entry_rep.data(), sizeof(int) * bins,
                        hipMemcpyHostToDevice));
    }
    e->graph_epoch++;
  }
  if (e->reps_uploaded.size() != e->reps.size() ||
      memcmp(e->reps_uploaded.data(), e->reps.data(),
             sizeof(RepInfo) * e->reps.size()) != 0) {
    HIP_TRY(hipMemcpy(e->d_reps, e->reps.data(), sizeof(RepInfo) * e->n_rep,
                      hipMemcpyHostToDevice));
    e->reps_uploaded = e->reps;
  }
  if (e->tail_dirty || e->tail_values_dirty) {
    if (e->tail_bins > 0)
      launch_tail_table(e->tail_bins * e->n_types, e->d_reps,
                        e->d_tail_entry_rep, e->n_types, e->d_paramsd, e->d_rc,
                        e->d_rho, e->d_tail, e->stream);
    e->tail_dirty = false;
    e->tail_values_dirty = false;
  }
  return SRE_OK;
}

// New LJ parameters after the last rebuild: recompute the tail values in place.
static void refresh_tail_values(sre_engine *e) {
  if (!e->tail_values_dirty) return;
  if (e->tail_bins > 0) {
    launch_tail_table(e->tail_bins * e->n_types, e->d_reps, e->d_tail_entry_rep,
                      e->n_types, e->d_paramsd, e->d_rc, e->d_rho, e->d_tail,
                      e->stream);
    if (e->tail_mode == SRE_TAIL_EXACT && e->d_tail_atom)
      launch_tail_exact(e->n_atoms, e->n_types, e->d_pos4, e->d_rcut, e->d_reps,
                        e->d_paramsd, e->d_rho, e->d_tail_atom, e->stream);
  }
  e->tail_values_dirty = false;
}

// Per-wave statistic partials of the force kernel for the current list
// layout; allocated at rebuild so that no timed or captured region allocates.
static sre_status ensure_list(sre_engine *e, int cap) {
  const int T = e->tune.force_lanes;
  const int A = kWave / T;
  const size_t waves = (size_t)(e->n_atoms + A - 1) / A;
  // + padding rows: the force kernel prefetches rows past a tile's end
  const size_t need = waves * (size_t)cap * A + (size_t)kListPadRows * kWave;
  if (need > e->list_size) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    dfree(e->d_list);
    HIP_TRY(dalloc(&e->d_list, need));
    e->list_size = need;
    e->graph_epoch++;
  }
  if (e->cap != cap || e->list_lanes != T) e->graph_epoch++;  // layout of captured kernels
  e->cap = cap;
  e->list_lanes = T;
  return SRE_OK;
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Everything a captured rebuild reads by value or address: a change of any
// of it means a new capture.
static std::vector<unsigned char> rebuild_signature(const sre_engine *e,
                                                    int n_parts) {
  std::vector<unsigned char> v;
  auto put = [&v](const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    v.insert(v.end(), c, c + n);
  };
  const int ints[] = {e->n_atoms, e->key_bits, e->n_cells, e->n_cols, e->n_rep,
                      e->cap, e->tune.build_mode, e->tune.build_lanes, e->tune.sort_mode,
                      e->tune.force_mode,
                      e->tune.force_lanes, e->n_types, e->pos_stride, n_parts,
                      (e->tail_bins > 0 && e->tail_mode == SRE_TAIL_EXACT) ? 1 : 0,
                      e->rc_explicit ? 1 : 0};
  put(ints, sizeof ints);
  put(&e->skin, sizeof e->skin);
  const void *ptrs[] = {e->d_pos, e->d_rep_of, e->d_reps, e->sb.key_a, e->sb.key_b,
                        e->sb.val_a, e->sb.val_b, e->sb.tkey, e->sb.tval,
                        e->sb.counts, e->sb.offsets, e->sb.scan_tmp, e->d_cell_count,
                        e->d_start,
                        e->d_types, e->d_rc, e->d_pos4, e->d_backup,
                        e->d_inv_perm, e->d_res4,
                        e->d_rcut, e->d_rc_atom, e->d_tail_idx, e->d_tail_atom,
                        e->d_paramsd,
                        e->d_rho, e->d_list, e->d_nnb, e->d_build_part,
                        e->d_tiles_per_col, e->d_tile_off, e->d_tile_tmp,
                        e->d_counters, e->d_skin_part, e->d_log, e->d_log_idx,
                        e->d_info_slices,
                        e->d_backup_s, e->pert_anchored ? e->d_anchor : nullptr,
                        e->d_anchor_s};
  put(ptrs, sizeof ptrs);
  return v;
}

static void enqueue_build(sre_engine *e, hipStream_t st) {
  const int N = e->n_atoms;
  // the cell-window force form reads the tile map of the tile build
  if (e->tune.force_mode == 1 && e->tune.build_mode != 1)
    launch_tile_map(e->n_cols, e->n_rep, e->d_reps, e->d_start, kCellTile,
                    e->d_tiles_per_col, e->d_tile_off, e->d_tile_tmp, st);
  if (e->tune.build_mode == 1)
    launch_build_tiles(N, e->n_cols, e->n_rep, e->tune.force_lanes, e->d_pos4,
                       e->d_rcut, e->d_reps, e->d_start, e->skin, e->cap,
                       e->d_list, e->d_nnb, e->d_build_part, e->d_tiles_per_col,
                       e->d_tile_off, e->d_tile_tmp, st);
  else
    launch_build_list(e->tune.build_lanes, e->tune.force_lanes, N, e->d_pos4,
                      e->d_rcut, e->d_reps, e->d_start, e->skin, e->cap,
                      e->d_list, e->d_nnb, e->d_build_part, st);
}

// The device part of a rebuild (capturable, no host decisions): cell keys,
// binning by cell (counting or radix sort), cell offsets, gather into cell
// order with rc assignment, exact tail, list build, and the reduction of the
// longest/total list length and of the segment's largest displacement (whose
// partials it clears).
// counting sort while one block scans all cell counts (one placement
// launch); beyond that the radix sort's passes are cheaper than the counting
// sort's atomic round trips over many cells
static bool counting_sort(const sre_engine *e) {
  return e->tune.sort_mode == 2 || (e->tune.sort_mode == 0 && e->n_cells <= kPlaceCells);
}

// counted: the cell histogram was made by the force launch just before
static void enqueue_rebuild(sre_engine *e, int n_parts, hipStream_t st,
                            bool counted = false) {
  const int N = e->n_atoms;
  const float *anchor = e->pert_anchored ? e->d_anchor : nullptr;
  const float *rc_atom = e->rc_explicit ? e->d_rc_atom : nullptr;
  if (counting_sort(e)) {
    launch_cell_sort_gather(N, e->n_cells, e->d_pos, e->pos_stride, e->d_types,
                            e->d_rep_of, e->d_reps, e->d_rc, e->n_types, e->sb,
                            e->d_cell_count, e->d_start, e->d_pos4, e->d_backup,
                            e->d_inv_perm, e->d_res4, e->d_rcut, e->d_tail_idx,
                            rc_atom, e->d_backup_s, anchor, e->d_anchor_s, st, counted);
  } else {
    launch_radix_sort(N, e->key_bits, e->d_pos, e->pos_stride, e->d_rep_of,
                      e->d_reps, e->sb, st);
    launch_gather_sorted(N, e->n_cells, e->sb.key_a, e->sb.val_a, e->d_pos,
                         e->pos_stride, e->d_types, e->d_rep_of, e->d_reps,
                         e->d_rc, e->n_types, e->d_start, e->d_pos4, e->d_backup,
                         e->d_inv_perm, e->d_res4, e->d_rcut, e->d_tail_idx,
                         rc_atom, st, e->d_backup_s, anchor, e->d_anchor_s);
  }
  if (e->tail_bins > 0 && e->tail_mode == SRE_TAIL_EXACT)
    launch_tail_exact(N, e->n_types, e->d_pos4, e->d_rcut, e->d_reps,
                      e->d_paramsd, e->d_rho, e->d_tail_atom, st);
  enqueue_build(e, st);
  launch_rebuild_info(n_parts, e->d_build_part, e->d_counters, skin_part_count(e->n_atoms),
                      e->d_skin_part, 1, reinterpret_cast<unsigned *>(e->d_counters + 2),
                      st, e->d_log, e->d_log_idx, kRebuildLog, e->d_info_slices);
}

// Queues one launch of the captured rebuild; returns its rebuild-log entry
// (valid once the stream has been synchronised).
static hipError_t launch_rebuild(sre_engine *e, const uint64_t **entry) {
  hipError_t r = hipGraphLaunch(e->rebuild_exec, e->stream);
  if (r != hipSuccess) return r;
  *entry = e->h_log + 3 * (size_t)(e->log_launches % kRebuildLog);
  ++e->log_launches;
  return hipSuccess;
}

static sre_status do_rebuild(sre_engine *e) {
  if (!e->system_set) return fail(SRE_ERR_INVALID, "rebuild: system not set");
  if (!e->types_set) return fail(SRE_ERR_INVALID, "rebuild: types not set");
  if (!e->params_set) return fail(SRE_ERR_INVALID, "rebuild: pair parameters not set");
  if (!e->d_pos) return fail(SRE_ERR_INVALID, "rebuild: positions not bound");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = prepare_tables(e);
  if (s != SRE_OK) return s;
  const int N = e->n_atoms;
  hipStream_t st = e->stream;
  if (e->key_bits > 32 || radix_passes(e->key_bits) > 4)
    return fail(SRE_ERR_CELL, "radix sort: key of %d bits unsupported", e->key_bits);
  if (e->tail_bins > 0 && e->tail_mode == SRE_TAIL_EXACT && !e->d_tail_atom)
    HIP_TRY(dalloc(&e->d_tail_atom, N));
  // list capacity: first guess from the densest possible shell, then exact
  int cap = e->cap;
  if (cap <= 0 || e->list_lanes != e->tune.force_lanes) {
    double guess = 0;
    for (const RepInfo &R : e->reps) {
      const double rl = R.rc_max + e->skin;
      const double rho = R.n_atoms / ((double)R.L[0] * R.L[1] * R.L[2]);
      guess = std::max(guess, 4.18879 * rl * rl * rl * rho * 2.0 + 32.0);
    }
    cap = round_up((int)guess, e->tune.force_lanes);
  }
  const int n_parts = build_part_count(e->tune.build_mode, N, e->tune.build_lanes,
                                       e->n_cols);
  if (e->build_part_cap < n_parts) {
    dfree(e->d_build_part);
    HIP_TRY(dalloc(&e->d_build_part, n_parts));
    e->build_part_cap = n_parts;
  }
  s = ensure_list(e, cap);
  if (s != SRE_OK) return s;
  // the whole device sequence as one captured graph (re-captured when any
  // argument changes), then one pinned readback
  std::vector<unsigned char> sig = rebuild_signature(e, n_parts);
  if (!e->rebuild_exec || sig != e->rebuild_sig) {
    if (e->rebuild_exec) {
      (void)hipGraphExecDestroy(e->rebuild_exec);
      e->rebuild_exec = nullptr;
    }
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueue_rebuild(e, n_parts, st);
    hipError_t r = hipStreamEndCapture(st, &g);
    if (r != hipSuccess)
      return fail(SRE_ERR_HIP, "rebuild capture: %s", hipGetErrorString(r));
    r = hipGraphInstantiate(&e->rebuild_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) {
      e->rebuild_exec = nullptr;
      return fail(SRE_ERR_HIP, "rebuild instantiate: %s", hipGetErrorString(r));
    }
    HIP_TRY(hipGraphUpload(e->rebuild_exec, st));
    e->rebuild_sig = std::move(sig);
  }
  const uint64_t *info = nullptr;
  HIP_TRY(launch_rebuild(e, &info));
  HIP_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < 3; ++k) e->h_info[k] = info[k];
  {
    const unsigned bits = (unsigned)(e->h_info[2] & 0xffffffffu);
    memcpy(&e->last_disp2, &bits, sizeof(float));
  }
  // overflow: grow the capacity and build again (eagerly; rare)
  for (int attempt = 0;; ++attempt) {
    const int mc = (int)e->h_info[0];
    e->max_count = mc;
    e->n_list = (int64_t)e->h_info[1];
    if (mc <= e->cap) break;
    if (attempt == 3)
      return fail(SRE_ERR_LIST_OVERFLOW, "neighbor list overflow (%d > %d)", mc,
                  e->cap);
    cap = round_up(std::max(mc + 16, (int)(mc * 1.2)), e->tune.force_lanes);
    s = ensure_list(e, cap);
    if (s != SRE_OK) return s;
    enqueue_build(e, st);
    launch_rebuild_info(n_parts, e->d_build_part, e->d_counters, 0, e->d_skin_part,
                        0, nullptr, st, nullptr, nullptr, 0, e->d_info_slices);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(e->h_info, e->d_counters, 2 * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  e->built = true;
  e->md_since_rebuild = 0;
  e->stats_valid = e->forces_valid = e->atom_stats_valid = false;
  return SRE_OK;
}

extern "C" sre_status sre_rebuild(sre_engine *e) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  return do_rebuild(e);
}

extern "C" sre_status sre_update_positions(sre_engine *e, int *need_rebuild) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "update before the first rebuild");
  HIP_TRY(hipSetDevice(e->device));
  launch_update_positions(e->n_atoms, e->d_inv_perm, const_cast<float *>(e->d_pos),
                          e->pos_stride, e->d_reps, rep0(e), e->d_pos4, e->d_backup,
                          e->d_res4, e->d_skin_part, nullptr, nullptr, e->stream);
  HIP_TRY(hipGetLastError());
  e->stats_valid = e->forces_valid = e->atom_stats_valid = false;
  if (need_rebuild) {
    launch_rebuild_info(0, nullptr, nullptr, skin_part_count(e->n_atoms),
                        e->d_skin_part, 0, e->d_flags, e->stream, nullptr, nullptr, 0,
                        e->d_info_slices);
    unsigned bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, e->d_flags, sizeof(unsigned),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    float d2;
    memcpy(&d2, &bits, sizeof d2);
    const float half = 0.5f * e->skin;
    *need_rebuild = d2 > half * half ? 1 : 0;
  }
  return SRE_OK;
}

// ---------------------------------------------------------------- compute --

static ForceArgs force_args(sre_engine *e, bool atom_stats) {
  ForceArgs a{};
  a.n = e->n_atoms;
  a.n_types = e->n_types;
  a.cap = e->cap;
  a.pos4 = e->d_pos4;
  a.res4 = e->d_res4;
  a.rcut = e->d_rcut;
  a.nnb = e->d_nnb;
  a.list = e->d_list;
  a.list_len = e->list_size;
  a.reps = e->d_reps;
  a.params = e->d_params;
  a.paramsd = e->d_paramsd;
  a.tail = e->d_tail;
  a.tail_idx = e->d_tail_idx;
  a.tail_atom = (e->tail_mode == SRE_TAIL_EXACT) ? e->d_tail_atom : nullptr;
  a.fx = e->d_fx; a.fy = e->d_fy; a.fz = e->d_fz;
  a.fxd = e->d_fxd; a.fyd = e->d_fyd; a.fzd = e->d_fzd;
  a.atom_stats = atom_stats ? e->d_atom_stats : nullptr;
  a.err_flag = e->d_flags + 1;
  a.ctr_in = nullptr;
  a.ctr_out = nullptr;
  // the per-atom builder writes one (longest list, total) pair per wave of
  // build_lanes lanes per atom: the force kernel's waves when the lane
  // counts agree
  a.wave_len = (e->tune.build_mode == 0 && e->tune.build_lanes == e->list_lanes)
                   ? e->d_build_part : nullptr;
  a.xcd_remap = e->tune.xcd_remap;
  a.has_shift = e->has_shift ? 1 : 0;
  // persistent launch: whole blocks of the tuned size
  const int per_block = e->tune.block / kWave;
  a.persist_waves = e->tune.persistent > 0
                        ? (e->tune.persistent * e->n_simd + per_block - 1) / per_block * per_block
                        : 0;
  return a;
}

// step_ctr: perturbation step counter the force kernel advances (graph
// replays with a perturbation), else NULL
static sre_status launch_compute(sre_engine *e, unsigned flags,
                                 uint64_t *step_ctr = nullptr) {
  const bool stats = flags & SRE_STATS;
  const bool tail = (flags & SRE_TAIL) && e->tail_bins > 0;
  const int T = e->list_lanes;
  refresh_tail_values(e);
  ForceArgs a = force_args(e, stats);
  a.ctr_in = step_ctr;
  a.ctr_out = step_ctr;
  if (e->tune.force_mode == 1) {
    if (e->precision != SRE_PREC_MIXED)
      return fail(SRE_ERR_INVALID, "the cell-window force form is mixed precision only");
    CellWindow c{e->n_cols, e->n_rep, e->d_tile_off, e->d_tiles_per_col, e->d_start,
                 e->d_backup_s, e->skin};
    launch_force_cells(stats, tail, a, c, e->stream);
  } else {
    launch_force(e->precision, T, e->tune.force_unroll, e->tune.block, stats, tail,
                 a, e->stream);
  }
  if ((flags & SRE_FCORR) && e->fc_cells > 0)
    launch_add_force_grid(e->n_atoms, e->d_pos, e->pos_stride, e->d_backup,
                          e->d_inv_perm, e->d_reps, e->d_fc, e->d_fx, e->d_fy,
                          e->d_fz, e->d_fxd, e->d_fyd, e->d_fzd, e->stream);
  if (stats) e->stats_pending = true;
  HIP_TRY(hipGetLastError());  // launch-configuration errors
  return SRE_OK;
}

// Reduces the per-atom statistics of the last SRE_STATS compute into
// d_stats (stream-ordered; fixed order, so the result is reproducible).
static sre_status reduce_stats(sre_engine *e) {
  if (e->stats_pending) {
    launch_stats_reduce(e->n_rep, e->d_reps, e->d_atom_stats, e->d_stat_part,
                        e->d_stats, e->stream);
    HIP_TRY(hipGetLastError());
    e->stats_pending = false;
  }
  return SRE_OK;
}

extern "C" sre_status sre_compute(sre_engine *e, unsigned flags) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "compute before the first rebuild");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = launch_compute(e, flags);
  if (s != SRE_OK) return s;
  e->forces_valid = true;
  e->stats_valid = (flags & SRE_STATS) != 0;
  e->atom_stats_valid = e->stats_valid;
  return SRE_OK;
}

extern "C" sre_status sre_step(sre_engine *e, unsigned flags, int *rebuilt) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  int need = 1;
  sre_status s;
  if (e->built) {
    s = sre_update_positions(e, &need);
    if (s != SRE_OK) return s;
  }
  if (need) {
    s = do_rebuild(e);
    if (s != SRE_OK) return s;
  }
  if (rebuilt) *rebuilt = need;
  return sre_compute(e, flags);
}

// ---------------------------------------------------------------- results --

static sre_status check_errors(sre_engine *e) {
  unsigned f = 0;
  HIP_TRY(hipMemcpyAsync(&f, e->d_flags + 1, sizeof(unsigned),
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (f & 1u) return report_force_range(e);
  return SRE_OK;
}

extern "C" sre_status sre_get_forces(sre_engine *e, double *dst,
                                     int dst_on_device) {
  if (!e || !dst) return fail(SRE_ERR_INVALID, "sre_get_forces: null argument");
  if (!e->forces_valid) return fail(SRE_ERR_NOT_READY, "no forces computed");
  HIP_TRY(hipSetDevice(e->device));
  const int N = e->n_atoms;
  double *out = dst_on_device ? dst : e->d_tmp_out;
  launch_export_forces(N, e->sb.val_a, e->d_fx, e->d_fy, e->d_fz, e->d_fxd,
                       e->d_fyd, e->d_fzd, out, nullptr, e->stream);
  if (!dst_on_device) {
    HIP_TRY(hipMemcpyAsync(dst, e->d_tmp_out, sizeof(double) * 3 * N,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_errors(e);
  }
  return SRE_OK;
}

extern "C" sre_status sre_add_forces_f32_soa(sre_engine *e, float *d_fx,
                                             float *d_fy, float *d_fz) {
  if (!e || !d_fx || !d_fy || !d_fz)
    return fail(SRE_ERR_INVALID, "sre_add_forces_f32_soa: null argument");
  if (!e->forces_valid) return fail(SRE_ERR_NOT_READY, "no forces computed");
  HIP_TRY(hipSetDevice(e->device));
  launch_add_forces_f32(e->n_atoms, e->sb.val_a, e->d_fx, e->d_fy, e->d_fz,
                        e->d_fxd, e->d_fyd, e->d_fzd, d_fx, d_fy, d_fz,
                        e->stream);
  return SRE_OK;
}

extern "C" sre_status sre_get_forces_fixed(sre_engine *e, int64_t *dst) {
  if (!e || !dst) return fail(SRE_ERR_INVALID, "sre_get_forces_fixed: null argument");
  if (!e->forces_valid) return fail(SRE_ERR_NOT_READY, "no forces computed");
  if (e->precision != SRE_PREC_MIXED)
    return fail(SRE_ERR_INVALID, "fixed-point forces exist only in mixed precision");
  HIP_TRY(hipSetDevice(e->device));
  const int N = e->n_atoms;
  int64_t *tmp = (int64_t *)e->d_tmp_out;
  launch_export_forces(N, e->sb.val_a, e->d_fx, e->d_fy, e->d_fz, nullptr,
                       nullptr, nullptr, nullptr, tmp, e->stream);
  HIP_TRY(hipMemcpyAsync(dst, tmp, sizeof(int64_t) * 3 * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return check_errors(e);
}

extern "C" sre_status sre_get_stats(sre_engine *e, double *out) {
  if (!e || !out) return fail(SRE_ERR_INVALID, "sre_get_stats: null argument");
  if (!e->stats_valid) return fail(SRE_ERR_NOT_READY, "no SRE_STATS compute yet");
  HIP_TRY(hipSetDevice(e->device));
  sre_status rs = reduce_stats(e);
  if (rs != SRE_OK) return rs;
  // the totals and the force-range flag into pinned memory, one synchronisation
  const size_t m = (size_t)4 * e->n_rep;
  if (e->h_stats_cap < m + 1) {
    if (e->h_stats) HIP_TRY(hipHostFree(e->h_stats));
    e->h_stats = nullptr;
    e->h_stats_cap = 0;
    HIP_TRY(hipHostMalloc((void **)&e->h_stats, sizeof(double) * (m + 1), hipHostMallocDefault));
    e->h_stats_cap = m + 1;
  }
  unsigned *flag = reinterpret_cast<unsigned *>(e->h_stats + m);
  HIP_TRY(hipMemcpyAsync(e->h_stats, e->d_stats, sizeof(double) * m,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(flag, e->d_flags + 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                         e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  memcpy(out, e->h_stats, sizeof(double) * m);
  if (*flag & 1u) return report_force_range(e);
  return SRE_OK;
}

extern "C" sre_status sre_get_atom_stats(sre_engine *e, double *out) {
  if (!e || !out) return fail(SRE_ERR_INVALID, "sre_get_atom_stats: null argument");
  if (!e->atom_stats_valid) return fail(SRE_ERR_NOT_READY, "no SRE_STATS compute yet");
  HIP_TRY(hipSetDevice(e->device));
  const int N = e->n_atoms;
  launch_export_atom_stats(N, e->sb.val_a, e->d_atom_stats, e->d_tmp_out,
                           e->stream);
  HIP_TRY(hipMemcpyAsync(out, e->d_tmp_out, sizeof(double) * 4 * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return SRE_OK;
}

extern "C" sre_status sre_neighbor_count(sre_engine *e, int64_t *n_list,
                                         int64_t *n_in_rc, int32_t *max_list) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "no neighbor list");
  HIP_TRY(hipSetDevice(e->device));
  if (n_list) {
    if (e->n_list < 0) {
      std::vector<int> nn(e->n_atoms);
      HIP_TRY(hipMemcpyAsync(nn.data(), e->d_nnb, sizeof(int) * e->n_atoms,
                             hipMemcpyDeviceToHost, e->stream));
      HIP_TRY(hipStreamSynchronize(e->stream));
      int64_t t = 0;
      for (int v : nn) t += v;
      e->n_list = t;
    }
    *n_list = e->n_list;
  }
  if (n_in_rc) {
    HIP_TRY(hipMemsetAsync(e->d_nin, 0, sizeof(unsigned long long) * e->n_rep,
                           e->stream));
    launch_count_in_rc(e->n_atoms, e->list_lanes, e->cap, e->d_pos4, e->d_rcut,
                       e->d_nnb, e->d_list, e->d_reps, e->d_nin, e->stream);
    HIP_TRY(hipGetLastError());
    std::vector<unsigned long long> c(e->n_rep);
    HIP_TRY(hipMemcpyAsync(c.data(), e->d_nin, sizeof(unsigned long long) * e->n_rep,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    int64_t t = 0;
    for (unsigned long long v : c) t += (int64_t)v;
    *n_in_rc = t;
  }
  if (max_list) *max_list = e->max_count;
  return SRE_OK;
}

extern "C" sre_status sre_export_pairs(sre_engine *e, int64_t *offsets,
                                       int32_t *j) {
  if (!e || !offsets) return fail(SRE_ERR_INVALID, "sre_export_pairs: null argument");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "no neighbor list");
  HIP_TRY(hipSetDevice(e->device));
  const int N = e->n_atoms;
  std::vector<int> nn(N), list(e->list_size);
  std::vector<uint32_t> perm(N);
  HIP_TRY(hipMemcpyAsync(nn.data(), e->d_nnb, sizeof(int) * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(perm.data(), e->sb.val_a, sizeof(uint32_t) * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(list.data(), e->d_list, sizeof(int) * e->list_size,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  std::vector<int64_t> cnt(N + 1, 0);
  for (int s = 0; s < N; ++s) cnt[perm[s] + 1] = nn[s];
  for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
  for (int i = 0; i <= N; ++i) offsets[i] = cnt[i];
  if (j) {
    for (int s = 0; s < N; ++s) {
      const int i = (int)perm[s];
      int32_t *row = j + cnt[i];
      for (int k = 0; k < nn[s]; ++k)
        row[k] = (int32_t)perm[list[list_slot(s, k, e->list_lanes, e->cap)]];
      std::sort(row, row + nn[s]);
    }
  }
  return SRE_OK;
}

extern "C" sre_status sre_get_rcut(sre_engine *e, float *rc) {
  if (!e || !rc) return fail(SRE_ERR_INVALID, "sre_get_rcut: null argument");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "no rc assigned yet");
  HIP_TRY(hipSetDevice(e->device));
  const int N = e->n_atoms;
  std::vector<float> r(N);
  std::vector<uint32_t> perm(N);
  HIP_TRY(hipMemcpyAsync(r.data(), e->d_rcut, sizeof(float) * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(perm.data(), e->sb.val_a, sizeof(uint32_t) * N,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (int s = 0; s < N; ++s) rc[perm[s]] = r[s];
  return SRE_OK;
}

// AssignRCut::assign (H/src/AssignRCut.cu:107-117), called by the adaptive
// driver every rcutAssignFeq steps (H/app/lj.nhc-xtc.cu:184-188): cutoffs
// from the current table at the current positions, without a list rebuild.
extern "C" sre_status sre_assign_rcut(sre_engine *e) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  if (!e->system_set || !e->d_pos)
    return fail(SRE_ERR_NOT_READY, "sre_assign_rcut: system or positions not set");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = prepare_tables(e);
  if (s != SRE_OK) return s;
  if (!e->rc_explicit) {
    e->rc_explicit = true;
    e->graph_epoch++;   // captured rebuilds now carry rc_atom
  }
  launch_assign_rc(e->n_atoms, e->d_pos, e->pos_stride, e->d_rep_of, e->d_reps,
                   e->d_rc, e->built ? e->d_inv_perm : nullptr, e->d_rc_atom,
                   e->d_rcut, e->stream);
  HIP_TRY(hipGetLastError());
  if (e->built && e->tail_bins > 0 && e->tail_mode == SRE_TAIL_EXACT)
    launch_tail_exact(e->n_atoms, e->n_types, e->d_pos4, e->d_rcut, e->d_reps,
                      e->d_paramsd, e->d_rho, e->d_tail_atom, e->stream);
  e->forces_valid = false;
  return SRE_OK;
}

// ---------------------------------------------------------------- adjoint --

extern "C" sre_status sre_adjoint(sre_engine *e, const double *dLdE,
                                  const double *d_dLdF, unsigned flags,
                                  double *g_eps, double *g_sigma,
                                  double *g_rc) {
  if (!e || !g_eps || !g_sigma)
    return fail(SRE_ERR_INVALID, "sre_adjoint: null argument");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "adjoint before the first rebuild");
  if (g_rc && e->rc_explicit)
    return fail(SRE_ERR_INVALID,
                "rc gradients need cutoffs assigned at rebuild (sre_assign_rcut was used)");
  const bool use_tail = (flags & SRE_TAIL) && e->tail_bins > 0;
  if (use_tail && e->tail_mode != SRE_TAIL_BIN)
    return fail(SRE_ERR_INVALID, "adjoint of the tail term needs SRE_TAIL_BIN");
  HIP_TRY(hipSetDevice(e->device));
  const int nt = e->n_types, npt = nt * nt;
  const int gw = adjoint_grid_waves(e->n_atoms, e->list_lanes);
  const int n_entries = std::max(1, e->tail_bins * nt);
  const int rc_n = std::max(1, e->rc_total);
  // one persistent workspace: [wave partials | dLdE | out | tout | g_rc |
  // tail contributions] doubles + tail weights (fixed point); the results
  // (out, tout, g_rc) are contiguous for one readback
  // pair part: 4 integer term totals per ordered type pair (k_adjoint.hip)
  const size_t n_wp = (size_t)gw * 4 * npt, n_res = (size_t)6 * npt + rc_n;
  const size_t n_ctb = (size_t)std::max(1, e->tail_bins) * 2 * npt;
  const size_t need = n_wp + e->n_rep + n_res + n_ctb;
  if (e->adj_cap < need) {
    dfree(e->d_adj);
    HIP_TRY(dalloc(&e->d_adj, need));
    e->adj_cap = need;
  }
  const size_t tw_need = (size_t)2 * n_entries + tail_w_scratch(n_entries);
  if (e->adj_tw_cap < tw_need) {
    dfree(e->d_adj_tw);
    HIP_TRY(dalloc(&e->d_adj_tw, tw_need));
    e->adj_tw_cap = tw_need;
  }
  double *d_wp = e->d_adj, *d_dLdE = d_wp + n_wp, *d_out = d_dLdE + e->n_rep;
  double *d_tout = d_out + 4 * npt, *d_grc = d_tout + 2 * npt;
  double *d_contrib = d_grc + rc_n;
  int64_t *d_tw = e->d_adj_tw;
  sre_status s = SRE_OK;
  std::vector<double> hE(e->n_rep, 0.0);
  if (dLdE) hE.assign(dLdE, dLdE + e->n_rep);
  HIP_TRY(hipMemcpyAsync(d_dLdE, hE.data(), sizeof(double) * e->n_rep,
                         hipMemcpyHostToDevice, e->stream));
  // zero tout and g_rc (all-zero bits)
  launch_fill_i64(reinterpret_cast<int64_t *>(d_tout), 0, (size_t)2 * npt + rc_n,
                  e->stream);
  AdjointArgs a{};
  a.n = e->n_atoms;
  a.n_types = nt;
  a.cap = e->cap;
  a.n_rep = e->n_rep;
  a.list_len = e->list_size;
  a.wave_len = (e->tune.build_mode == 0 && e->tune.build_lanes == e->list_lanes)
                   ? e->d_build_part : nullptr;
  a.pos4 = e->d_pos4;
  a.rcut = e->d_rcut;
  a.nnb = e->d_nnb;
  a.list = e->d_list;
  a.perm = e->sb.val_a;
  a.reps = e->d_reps;
  a.params = e->d_params;
  a.dLdE = d_dLdE;
  a.dLdF = d_dLdF;
  a.tail_idx = e->d_tail_idx;
  a.use_tail = use_tail ? 1 : 0;
  a.wave_part = reinterpret_cast<long long *>(d_wp);
  a.tail_w = d_tw;
  launch_adjoint_pairs(a, e->precision, e->list_lanes, reinterpret_cast<long long *>(d_out),
                       e->stream);
  if (use_tail) {
    launch_tail_weights(e->n_atoms, n_entries, e->d_tail_idx, e->sb.val_a, e->d_pos4,
                        d_dLdF, d_dLdE, reinterpret_cast<long long *>(d_tw + 2 * n_entries),
                        d_tw, e->stream);
    launch_adjoint_tail(e->tail_bins, e->d_tail_entry_rep, e->d_reps, nt,
                        e->d_paramsd, e->d_rc, e->d_rho, d_tw, d_contrib, d_grc,
                        d_tout, e->stream);
  }
  HIP_TRY(hipGetLastError());
  std::vector<double> res(n_res);
  HIP_TRY(hipMemcpyAsync(res.data(), d_out, sizeof(double) * n_res,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  const double *tout = res.data() + 4 * npt, *grc = tout + 2 * npt;
  if (s != SRE_OK) return s;
  // pair part from the exact term totals T1..T4 (scale 2^32):
  // d/d eps = 4 T1 - 24 T2,  d/d sigma = eps / sigma (24 T3 - 144 T4)
  std::vector<double> out(2 * npt);
  for (int p = 0; p < npt; ++p) {
    double T[4];
    for (int q = 0; q < 4; ++q) {
      long long v;
      memcpy(&v, &res[(size_t)q * npt + p], sizeof v);
      T[q] = (double)v * (1.0 / 4294967296.0);
    }
    out[p] = 4.0 * T[0] - 24.0 * T[1];
    out[npt + p] = (double)e->eps[p] / (double)e->sig[p] * (24.0 * T[2] - 144.0 * T[3]);
  }
  // Values are per ordered type pair; a symmetric parameter (a,b) = (b,a)
  // collects both orders.
  for (int pa = 0; pa < nt; ++pa)
    for (int pb = 0; pb < nt; ++pb) {
      const int ab = pa * nt + pb, ba = pb * nt + pa;
      double ge = out[ab] + tout[ab], gs = out[npt + ab] + tout[npt + ab];
      if (ab != ba) {
        ge += out[ba] + tout[ba];
        gs += out[npt + ba] + tout[npt + ba];
      }
      g_eps[ab] = ge;
      g_sigma[ab] = gs;
    }
  if (g_rc)
    for (int k = 0; k < e->rc_total; ++k) g_rc[k] = grc[k];
  return SRE_OK;
}

// ----------------------------------------------------------------- tuning --

extern "C" sre_status sre_get_tuning(sre_engine *e, sre_tuning *t) {
  if (!e || !t) return fail(SRE_ERR_INVALID, "null argument");
  *t = e->tune;
  return SRE_OK;
}

extern "C" sre_status sre_set_tuning(sre_engine *e, const sre_tuning *t) {
  if (!e || !t) return fail(SRE_ERR_INVALID, "null argument");
  if (!pow2_lanes(t->force_lanes) || !pow2_lanes(t->build_lanes))
    return fail(SRE_ERR_INVALID, "lanes must be 1, 2, 4, 8 or 16");
  // the force kernels are compiled with __launch_bounds__(256)
  if (t->block != 64 && t->block != 128 && t->block != 256)
    return fail(SRE_ERR_INVALID, "block must be 64, 128 or 256");
  if (!(t->cell_xy >= 0.05f && t->cell_xy <= 4.f && t->cell_z >= 0.05f &&
        t->cell_z <= 4.f))
    return fail(SRE_ERR_INVALID, "cell sizes must be in [0.05, 4]");
  if (t->build_mode != 0 && t->build_mode != 1)
    return fail(SRE_ERR_INVALID, "build_mode must be 0 or 1");
  if (t->xcd_remap != 0 && t->xcd_remap != 1)
    return fail(SRE_ERR_INVALID, "xcd_remap must be 0 or 1");
  if (t->force_unroll != 1 && t->force_unroll != 2 && t->force_unroll != 4)
    return fail(SRE_ERR_INVALID, "force_unroll must be 1, 2 or 4");
  if (t->persistent < 0 || t->persistent > 16)
    return fail(SRE_ERR_INVALID, "persistent must be in [0, 16]");
  if (t->sort_mode < 0 || t->sort_mode > 2)
    return fail(SRE_ERR_INVALID, "sort_mode must be 0, 1 or 2");
  if (t->force_mode != 0 && t->force_mode != 1)
    return fail(SRE_ERR_INVALID, "force_mode must be 0 or 1");
  if (t->force_mode == 1 && t->build_mode == 1)
    return fail(SRE_ERR_INVALID, "force_mode 1 needs build_mode 0 (both use the tile map)");
  // (the cell-window force form reads a tile map made at the rebuild)
  const bool layout = t->force_lanes != e->tune.force_lanes ||
                      t->build_lanes != e->tune.build_lanes ||
                      t->build_mode != e->tune.build_mode ||
                      t->force_mode != e->tune.force_mode ||
                      t->cell_xy != e->tune.cell_xy || t->cell_z != e->tune.cell_z;
  if (memcmp(&e->tune, t, sizeof *t) != 0) e->graph_epoch++;  // captured launches
  e->tune = *t;
  if (layout) {
    e->built = false;
    e->cap = 0;
  }
  return SRE_OK;
}

static sre_status time_region(sre_engine *e, int n_iter, double *ms,
                              sre_status (*body)(sre_engine *, unsigned),
                              unsigned flags) {
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  HIP_TRY(hipEventRecord(a, e->stream));
  sre_status s = SRE_OK;
  for (int i = 0; i < n_iter && s == SRE_OK; ++i) s = body(e, flags);
  HIP_TRY(hipEventRecord(b, e->stream));
  HIP_TRY(hipEventSynchronize(b));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = (double)t / n_iter;
  return s;
}

// one force launch (the statistic reduction is deferred to the read)
static sre_status body_compute(sre_engine *e, unsigned flags) {
  return launch_compute(e, flags & (SRE_STATS | SRE_TAIL | SRE_FCORR));
}

extern "C" sre_status sre_time_compute(sre_engine *e, unsigned flags,
                                       int n_iter, double *ms_per_iter) {
  if (!e || !ms_per_iter || n_iter < 1) return fail(SRE_ERR_INVALID, "bad arguments");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "compute before the first rebuild");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = time_region(e, n_iter, ms_per_iter, body_compute, flags);
  if (s == SRE_OK) {
    e->forces_valid = true;
    e->stats_valid = e->atom_stats_valid = (flags & SRE_STATS) != 0;
  }
  return s;
}

static sre_status body_rebuild(sre_engine *e, unsigned) { return do_rebuild(e); }

extern "C" sre_status sre_time_rebuild(sre_engine *e, int n_iter,
                                       double *ms_per_iter) {
  if (!e || !ms_per_iter || n_iter < 1) return fail(SRE_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(e->device));
  return time_region(e, n_iter, ms_per_iter, body_rebuild, 0);
}

extern "C" sre_status sre_synchronize(sre_engine *e) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return SRE_OK;
}

extern "C" sre_status sre_autotune(sre_engine *e, unsigned flags,
                                   const char *cache_path) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  HIP_TRY(hipSetDevice(e->device));
  char sig[512];
  {
    double lsum = 0, rcs = 0;
    for (int r = 0; r < e->n_rep; ++r) {
      lsum += (double)e->box[3 * r] + 3.0 * e->box[3 * r + 1] + 7.0 * e->box[3 * r + 2];
      for (float v : e->rc_tab[r]) rcs += v;
    }
    snprintf(sig, sizeof sig, "N=%d R=%d T=%d prec=%d flags=%u skin=%.4f L=%.6g rc=%.6g",
             e->n_atoms, e->n_rep, e->n_types, e->precision, flags,
             (double)e->skin, lsum, rcs);
  }
  if (cache_path) {
    FILE *f = fopen(cache_path, "r");
    if (f) {
      char line[1024];
      while (fgets(line, sizeof line, f)) {
        char *bar = strchr(line, '|');
        if (!bar) continue;
        *bar = 0;
        if (strcmp(line, sig) == 0) {
          sre_tuning t = e->tune;
          if (sscanf(bar + 1, "%d %d %d %f %f %d %d %d %d", &t.force_lanes,
                     &t.build_lanes, &t.block, &t.cell_xy, &t.cell_z,
                     &t.build_mode, &t.xcd_remap, &t.force_unroll, &t.persistent) == 9) {
            fclose(f);
            sre_status s = sre_set_tuning(e, &t);
            if (s != SRE_OK) return s;
            return do_rebuild(e);
          }
        }
      }
      fclose(f);
    }
  }
  // phase 1: binning + list build
  const float cells_xy[] = {0.35f, 0.5f, 0.7f, 1.0f};
  const float cells_z[] = {0.2f, 0.25f, 0.35f, 0.5f};
  // (build_mode, build_lanes): per-atom search with 4/8/16 lanes, or tiles
  const int builds[][2] = {{0, 4}, {0, 8}, {0, 16}, {1, 8}};
  sre_tuning best = e->tune;
  double best_ms = 1e30;
  for (float cxy : cells_xy)
    for (float cz : cells_z)
      for (const auto &bd : builds) {
        sre_tuning t = e->tune;
        t.cell_xy = cxy;
        t.cell_z = cz;
        t.build_mode = bd[0];
        t.build_lanes = bd[1];
        sre_status s = sre_set_tuning(e, &t);
        if (s != SRE_OK) return s;
        s = do_rebuild(e);
        if (s != SRE_OK) return s;
        double ms;
        s = sre_time_rebuild(e, 3, &ms);
        if (s != SRE_OK) return s;
        if (ms < best_ms) { best_ms = ms; best = t; }
      }
  // phase 2: force kernel lanes and block
  const int force_lanes[] = {1, 2, 4, 8, 16};
  const int blocks[] = {64, 128, 256};
  double best_f = 1e30;
  sre_tuning bestf = best;
  std::vector<int64_t> ref, cur;
  const int unrolls[] = {1, 2, 4};
  for (int fl : force_lanes)
    for (int bk : blocks)
      for (int fu : unrolls)
      for (int xr = 0; xr < 2; ++xr) {
      if (e->precision == SRE_PREC_FP64 && fu != 4) continue;  // unused there
      sre_tuning t = best;
      t.force_lanes = fl;
      t.block = bk;
      t.force_unroll = fu;
      t.xcd_remap = xr;
      sre_status s = sre_set_tuning(e, &t);
      if (s != SRE_OK) return s;
      s = do_rebuild(e);
      if (s != SRE_OK) return s;
      double ms;
      s = sre_time_compute(e, flags, 2, &ms);
      if (s != SRE_OK) return s;
      s = sre_time_compute(e, flags, 10, &ms);
      if (s != SRE_OK) return s;
      if (e->precision == SRE_PREC_MIXED) {
        cur.resize((size_t)3 * e->n_atoms);
        s = sre_get_forces_fixed(e, cur.data());
        if (s != SRE_OK) return s;
        if (ref.empty()) ref = cur;
        else if (ref != cur)
          return fail(SRE_ERR_INVALID,
                      "autotune: forces differ between lane counts (bug)");
      }
      if (ms < best_f) { best_f = ms; bestf = t; }
    }
  // phase 3 (mixed form): flat or persistent launch of the best kernel
  if (e->precision == SRE_PREC_MIXED) {
    const sre_tuning flat = bestf;
    const int persist[] = {1, 2, 4, 8};
    for (int pw : persist) {
      sre_tuning t = flat;
      t.persistent = pw;
      sre_status s = sre_set_tuning(e, &t);
      if (s != SRE_OK) return s;
      s = do_rebuild(e);
      if (s != SRE_OK) return s;
      double ms;
      s = sre_time_compute(e, flags, 2, &ms);
      if (s != SRE_OK) return s;
      s = sre_time_compute(e, flags, 10, &ms);
      if (s != SRE_OK) return s;
      cur.resize((size_t)3 * e->n_atoms);
      s = sre_get_forces_fixed(e, cur.data());
      if (s != SRE_OK) return s;
      if (ref != cur)
        return fail(SRE_ERR_INVALID, "autotune: persistent launch changed the forces (bug)");
      if (ms < best_f) { best_f = ms; bestf = t; }
    }
  }
  sre_status s = sre_set_tuning(e, &bestf);
  if (s != SRE_OK) return s;
  s = do_rebuild(e);
  if (s != SRE_OK) return s;
  if (cache_path) {
    FILE *f = fopen(cache_path, "a");
    if (f) {
      fprintf(f, "%s|%d %d %d %.4f %.4f %d %d %d %d\n", sig, bestf.force_lanes,
              bestf.build_lanes, bestf.block, (double)bestf.cell_xy,
              (double)bestf.cell_z, bestf.build_mode, bestf.xcd_remap,
              bestf.force_unroll, bestf.persistent);
      fclose(f);
    }
  }
  return SRE_OK;
}

// ----------------------------------------------------------------- graphs --

// MD step pieces (k_md.hip): the factors of the step from the last kick's
// sums, then kick + drift + refresh, and after the force the second kick
static void enqueue_md_drift(sre_engine *e) {
  launch_md_factors(e->n_atoms, e->d_md_part, e->d_md, e->stream);
  DriftArgs d{e->d_vel, e->d_fcall, e->d_mass, e->d_types, e->d_md,
              (float)e->md.dt};
  launch_update_positions(e->n_atoms, e->d_inv_perm, const_cast<float *>(e->d_pos),
                          e->pos_stride, e->d_reps, rep0(e), e->d_pos4, e->d_backup,
                          e->d_res4, e->d_skin_part, nullptr, &d, e->stream);
}

static void enqueue_md_kick(sre_engine *e, float hdt) {
  launch_md_kick(e->n_atoms, e->d_inv_perm, e->d_types, e->d_fx, e->d_fy,
                 e->d_fz, e->d_fxd, e->d_fyd, e->d_fzd, e->d_mass, hdt,
                 e->d_vel, e->d_fcall, e->d_md_part, e->stream);
}

static sre_status enqueue_md_step(sre_engine *e, unsigned flags) {
  enqueue_md_drift(e);
  sre_status s = launch_compute(e, flags & (SRE_STATS | SRE_TAIL | SRE_FCORR), nullptr);
  if (s != SRE_OK) return s;
  enqueue_md_kick(e, (float)(0.5 * e->md.dt));
  return SRE_OK;
}

// K perturbation steps (no integrator, no correction force) with the
// refresh of each next step folded into the force launch: one refresh of
// the first step, then K force launches, launch j reading coordinate buffer
// X_j and refreshing step j + 1 into X_{j+1} (j < K - 1).  The buffers
// alternate so that the last step reads d_pos4 (X_{K-1}), which therefore
// holds the current coordinates afterwards as after unfused steps.  The
// step counter alternates between its two slots in the same way (launch j
// reads slot j % 2, the refresh inside it draws step *slot + 1, one thread
// writes the next slot; the last launch writes slot 0, which the next
// graph's first refresh reads).  Same values as the unfused steps: the
// refresh arithmetic is the same and no launch reads a buffer another
// writes.
// count_next: the last launch also makes the cell histogram of the counting
// sort of a rebuild queued right after these steps (CountFold).
static sre_status enqueue_fused_steps(sre_engine *e, unsigned flags, int K,
                                      const PerturbArgs &pa, bool count_next = false) {
  const bool stats = flags & SRE_STATS;
  const bool tail = (flags & SRE_TAIL) && e->tail_bins > 0;
  auto pos4_of = [&](int j) { return ((K - 1 - j) % 2 == 0) ? e->d_pos4 : e->d_pos4_alt; };
  auto res4_of = [&](int j) { return ((K - 1 - j) % 2 == 0) ? e->d_res4 : e->d_res4_alt; };
  launch_update_positions(e->n_atoms, e->d_inv_perm, const_cast<float *>(e->d_pos),
                          e->pos_stride, e->d_reps, rep0(e), pos4_of(0), e->d_backup,
                          res4_of(0), e->d_skin_part, &pa, nullptr, e->stream);
  refresh_tail_values(e);
  NextRefresh nx{};
  nx.active = 1;
  // one thread per atom, blocks of the force launch's size, a multiple of 8
  nx.n_blocks = ((e->n_atoms + e->tune.block - 1) / e->tune.block + 7) / 8 * 8;
  nx.perm = e->sb.val_a;
  nx.pos = const_cast<float *>(e->d_pos);
  nx.stride = e->pos_stride;
  nx.anchor_s = pa.anchor ? e->d_anchor_s : nullptr;
  nx.backup_s = e->d_backup_s;
  nx.skin_part = e->d_skin_part;
  nx.sigma = pa.sigma;
  nx.seed = pa.seed;
  if (const RepInfo *r0 = rep0(e)) {
    nx.one_rep = 1;
    for (int d = 0; d < 3; ++d) {
      nx.g1.L[d] = r0->L[d];
      nx.g1.q[d] = r0->q[d];
      nx.g1.invq[d] = r0->invq[d];
      nx.g1.qd[d] = r0->qd[d];
    }
  }
  for (int j = 0; j < K; ++j) {
    ForceArgs a = force_args(e, stats);
    a.pos4 = pos4_of(j);
    a.res4 = res4_of(j);
    a.ctr_in = e->d_step_ctr + (j % 2);
    a.ctr_out = j == K - 1 ? e->d_step_ctr : e->d_step_ctr + ((j + 1) % 2);
    if (j < K - 1) {
      a.next = nx;
      a.next.pos4 = pos4_of(j + 1);
      a.next.res4 = res4_of(j + 1);
    } else if (count_next) {
      a.count.n_blocks = nx.n_blocks;
      a.count.pos = e->d_pos;
      a.count.stride = e->pos_stride;
      a.count.rep_of = e->d_rep_of;
      a.count.n_cells = e->n_cells;
      a.count.keys = e->sb.tkey;
      a.count.count = e->d_cell_count;
    }
    launch_force(e->precision, e->list_lanes, e->tune.force_unroll, e->tune.block, stats,
                 tail, a, e->stream);
  }
  if (stats) e->stats_pending = true;
  HIP_TRY(hipGetLastError());
  return SRE_OK;
}

// The kernels of K <= kMaxGraphSteps steps (the contents of a step graph),
// queued on the engine stream.  count_next asks the last force launch for the
// histogram of the counting sort of a rebuild that follows; *counted tells
// whether it was made (only the fused perturbation steps make it).
static sre_status enqueue_steps(sre_engine *e, unsigned flags, int K,
                                bool count_next = false, bool *counted = nullptr) {
  const bool pert = (flags & SRE_PERTURB) != 0;
  PerturbArgs pa{e->pert_sigma, e->pert_seed,
                 e->pert_anchored ? e->d_anchor : nullptr, e->d_step_ctr};
  const bool md = (flags & kGraphMD) != 0;
  if (counted) *counted = false;
  if (pert && !md && !(flags & SRE_FCORR) && K >= 2 && (!pa.anchor || e->d_anchor_s) &&
      e->tune.force_mode == 0) {
    if (counted) *counted = count_next;
    return enqueue_fused_steps(e, flags & (SRE_STATS | SRE_TAIL), K, pa, count_next);
  }
  sre_status s = SRE_OK;
  for (int k = 0; k < K && s == SRE_OK; ++k) {
    if (md) {
      s = enqueue_md_step(e, k == K - 1 ? flags : flags & ~SRE_STATS);
      continue;
    }
    launch_update_positions(e->n_atoms, e->d_inv_perm, const_cast<float *>(e->d_pos),
                            e->pos_stride, e->d_reps, rep0(e), e->d_pos4, e->d_backup,
                            e->d_res4, e->d_skin_part, pert ? &pa : nullptr, nullptr,
                            e->stream);
    s = launch_compute(e, flags & (SRE_STATS | SRE_TAIL | SRE_FCORR),
                       pert ? e->d_step_ctr : nullptr);
  }
  return s;
}

// Captures K steps of [refresh (+ perturbation) + force] into one graph, or
// returns the cached one; the graph reads the engine's buffers by address, so
// every reallocation or layout change bumps graph_epoch and invalidates it.
static sre_status step_graph(sre_engine *e, unsigned flags, int K,
                             hipGraphExec_t *out) {
  sre_engine::StepGraph *slot = nullptr, *lru = &e->step_graphs[0];
  for (sre_engine::StepGraph &g : e->step_graphs) {
    if (g.exec && g.flags == flags && g.steps == K && g.epoch == e->graph_epoch)
      slot = &g;
    if (g.last_use < lru->last_use) lru = &g;
  }
  if (!slot) {
    slot = lru;
    if (slot->exec) {
      (void)hipGraphExecDestroy(slot->exec);
      slot->exec = nullptr;
    }
    const bool pert = (flags & SRE_PERTURB) != 0;
    if (pert && !e->d_step_ctr)
      return fail(SRE_ERR_INVALID, "SRE_PERTURB needs sre_set_perturbation first");
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    sre_status s = enqueue_steps(e, flags, K);
    hipError_t r = hipStreamEndCapture(e->stream, &g);
    if (s != SRE_OK) {
      if (r == hipSuccess) (void)hipGraphDestroy(g);
      return s;
    }
    if (r != hipSuccess)
      return fail(SRE_ERR_HIP, "graph capture: %s", hipGetErrorString(r));
    r = hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) {
      slot->exec = nullptr;
      return fail(SRE_ERR_HIP, "graph instantiate: %s", hipGetErrorString(r));
    }
    HIP_TRY(hipGraphUpload(slot->exec, e->stream));
    slot->flags = flags;
    slot->steps = K;
    slot->epoch = e->graph_epoch;
  }
  slot->last_use = ++e->graph_uses;
  *out = slot->exec;
  return SRE_OK;
}

// Checks before queued steps; uploads changed tail values.
static sre_status steps_ready(sre_engine *e, unsigned flags) {
  refresh_tail_values(e);
  if ((flags & SRE_PERTURB) && !e->d_step_ctr)
    return fail(SRE_ERR_INVALID, "SRE_PERTURB needs sre_set_perturbation first");
  if ((flags & SRE_PERTURB) && e->pert_anchored &&
      (e->anchor_stride != e->pos_stride ||
       e->anchor_count != (size_t)e->n_atoms * e->pos_stride))
    return fail(SRE_ERR_INVALID,
                "perturbation anchor does not match the bound positions "
                "(call sre_set_perturbation again)");
  return SRE_OK;
}

extern "C" sre_status sre_graph_run(sre_engine *e, unsigned flags, int n_steps) {
  if (!e || n_steps < 0) return fail(SRE_ERR_INVALID, "bad arguments");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "graph before the first rebuild");
  HIP_TRY(hipSetDevice(e->device));
  {
    sre_status s = steps_ready(e, flags);
    if (s != SRE_OK) return s;
  }
  // whole graphs of kMaxGraphSteps steps, then one graph for the remainder
  int left = n_steps;
  while (left > 0) {
    const int K = std::min(left, kMaxGraphSteps);
    hipGraphExec_t ex;
    sre_status s = step_graph(e, flags, K, &ex);
    if (s != SRE_OK) return s;
    const int reps = left / K;
    for (int i = 0; i < reps; ++i) HIP_TRY(hipGraphLaunch(ex, e->stream));
    left -= reps * K;
  }
  if (n_steps > 0) {
    e->forces_valid = true;
    e->stats_valid = e->atom_stats_valid = (flags & SRE_STATS) != 0;
    e->stats_pending = e->stats_valid;
  }
  return SRE_OK;
}

// Segments of `steps` graph steps each followed by a list rebuild, queued
// back to back with one synchronisation per kRebuildLog segments: each
// rebuild logs its longest list, total length and the largest displacement
// of the segment before it (the rebuild log, host memory), checked once the
// queued segments have run.
// Segments of `steps` steps, each followed by the device part of a
// rebuild, captured m at a time as one graph: between graph launches the GPU
// idles ~8 us (measured at 64k atoms), between the kernels of one graph it
// does not.  The graphs for every m = 2^k (k < kSegGraphs) are captured
// together, so that any later count of segments is replayed from them
// (largest first) without a capture.
static sre_status segment_graphs(sre_engine *e, unsigned flags, int steps, int n_parts) {
  if (e->seg_exec[0] && e->seg_flags == flags && e->seg_steps == steps &&
      e->seg_epoch == e->graph_epoch && e->seg_sig == e->rebuild_sig)
    return SRE_OK;
  for (hipGraphExec_t &g : e->seg_exec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  e->seg_epoch = -1;   // valid again only once every size is captured
  for (int k = 0; k < sre_engine::kSegGraphs; ++k) {
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    sre_status s = SRE_OK;
    for (int seg = 0; seg < (1 << k) && s == SRE_OK; ++seg) {
      // the last force launch of the segment makes the rebuild's histogram
      bool counted = false;
      for (int left = steps; left > 0 && s == SRE_OK; left -= kMaxGraphSteps)
        s = enqueue_steps(e, flags, std::min(left, kMaxGraphSteps),
                          left <= kMaxGraphSteps && counting_sort(e), &counted);
      if (s == SRE_OK) enqueue_rebuild(e, n_parts, e->stream, counted);
    }
    hipError_t r = hipStreamEndCapture(e->stream, &g);
    if (s != SRE_OK) {
      if (r == hipSuccess) (void)hipGraphDestroy(g);
      return s;
    }
    if (r != hipSuccess)
      return fail(SRE_ERR_HIP, "segment graph capture: %s", hipGetErrorString(r));
    r = hipGraphInstantiate(&e->seg_exec[k], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) {
      e->seg_exec[k] = nullptr;
      return fail(SRE_ERR_HIP, "segment graph instantiate: %s", hipGetErrorString(r));
    }
    // upload now: the first launch of a graph not uploaded carries the upload
    HIP_TRY(hipGraphUpload(e->seg_exec[k], e->stream));
  }
  e->seg_flags = flags;
  e->seg_steps = steps;
  e->seg_epoch = e->graph_epoch;
  e->seg_sig = e->rebuild_sig;
  return SRE_OK;
}

static sre_status run_segment_queue(sre_engine *e, unsigned flags, int g0, int n_seg,
                                    int steps, int n_parts, int64_t *n_list,
                                    float *d2_out) {
  std::vector<const uint64_t *> info(n_seg);
  {
    sre_status s = steps_ready(e, flags);
    if (s != SRE_OK) return s;
  }
  {
    sre_status s = segment_graphs(e, flags, steps, n_parts);
    if (s != SRE_OK) return s;
  }
  for (int g = 0; g < n_seg;) {
    int k = sre_engine::kSegGraphs - 1;
    while ((1 << k) > n_seg - g) --k;
    HIP_TRY(hipGraphLaunch(e->seg_exec[k], e->stream));
    // each rebuild of the graph writes the next rebuild-log entry
    for (int q = 0; q < (1 << k); ++q, ++g) {
      info[g] = e->h_log + 3 * (size_t)(e->log_launches % kRebuildLog);
      ++e->log_launches;
    }
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->built = true;
  e->md_since_rebuild = 0;
  // the per-atom statistics of the last step are still those of the atoms'
  // slots before the final rebuild, and every replica keeps its slot range,
  // so their totals stay readable; the forces and per-atom values are in
  // the old order and are not
  e->stats_valid = e->stats_pending = (flags & SRE_STATS) != 0;
  e->forces_valid = e->atom_stats_valid = false;
  float d2max = 0.f;
  int worst = -1;
  std::vector<float> d2(n_seg);
  for (int g = 0; g < n_seg; ++g) {
    const uint64_t *h = info[g];
    const unsigned bits = (unsigned)(h[2] & 0xffffffffu);
    memcpy(&d2[g], &bits, sizeof(float));
    d2max = std::max(d2max, d2[g]);
    if (n_list) n_list[g] = (int64_t)h[1];
    if (worst < 0 && (int)h[0] > e->cap) worst = g;
  }
  const uint64_t *last = info[n_seg - 1];
  e->max_count = (int)last[0];
  e->n_list = (int64_t)last[1];
  e->last_disp2 = d2max;
  *d2_out = d2max;
  if (worst >= 0) {
    // lists were cut at the capacity from rebuild `worst` on: regrow now so
    // that the next call is exact, and report the segments as invalid
    int mc = 0;
    for (int g = worst; g < n_seg; ++g) mc = std::max(mc, (int)info[g][0]);
    const int old_cap = e->cap;
    sre_status s = ensure_list(
        e, round_up(std::max(mc + 16, (int)(mc * 1.2)), e->tune.force_lanes));
    if (s != SRE_OK) return s;
    s = do_rebuild(e);
    if (s != SRE_OK) return s;
    return fail(SRE_ERR_LIST_OVERFLOW,
                "neighbor list overflow at the rebuild ending segment %d (%d > %d); "
                "the capacity has been regrown to %d, the later segments are invalid",
                g0 + worst, mc, old_cap, e->cap);
  }
  // a segment in which an atom moved more than skin/2 from its rebuild
  // position may have missed pairs: its forces are invalid (the lists are
  // rebuilt, so the next call starts from a consistent state)
  const float half = 0.5f * e->skin;
  for (int g = 0; g < n_seg; ++g)
    if (d2[g] > half * half)
      return fail(SRE_ERR_SKIN,
                  "segment %d: max displacement %g exceeds skin/2 = %g; its forces "
                  "may have missed pairs", g0 + g, sqrt((double)d2[g]), (double)half);
  return SRE_OK;
}

extern "C" sre_status sre_run_segments(sre_engine *e, unsigned flags, int n_segments,
                                       int steps, int64_t *n_list, float *max_disp) {
  if (!e || n_segments < 0 || steps < 1)
    return fail(SRE_ERR_INVALID, "sre_run_segments: bad arguments");
  if (!e->built) return fail(SRE_ERR_NOT_READY, "segments before the first rebuild");
  if (n_segments == 0) {
    if (max_disp) *max_disp = 0.f;
    return SRE_OK;
  }
  HIP_TRY(hipSetDevice(e->device));
  // the captured rebuild must match the current layout; otherwise one
  // ordinary rebuild captures it again
  const int n_parts = build_part_count(e->tune.build_mode, e->n_atoms,
                                       e->tune.build_lanes, e->n_cols);
  if (!e->rebuild_exec || rebuild_signature(e, n_parts) != e->rebuild_sig) {
    sre_status s = do_rebuild(e);
    if (s != SRE_OK) return s;
  }
  float d2max = 0.f;
  for (int g0 = 0; g0 < n_segments; g0 += kRebuildLog) {
    const int m = std::min(kRebuildLog, n_segments - g0);
    float d2 = 0.f;
    sre_status s = run_segment_queue(e, flags, g0, m, steps, n_parts,
                                     n_list ? n_list + g0 : nullptr, &d2);
    if (s != SRE_OK) return s;
    d2max = std::max(d2max, d2);
  }
  e->last_disp2 = d2max;
  if (max_disp) *max_disp = sqrtf(d2max);
  return SRE_OK;
}

extern "C" sre_status sre_last_rebuild_displacement(sre_engine *e,
                                                    float *max_disp) {
  if (!e || !max_disp) return fail(SRE_ERR_INVALID, "null argument");
  *max_disp = sqrtf(e->last_disp2);
  return SRE_OK;
}

extern "C" sre_status sre_check_skin_flag(sre_engine *e) {
  if (!e) return fail(SRE_ERR_INVALID, "null engine");
  HIP_TRY(hipSetDevice(e->device));
  if (!e->d_skin_part) return fail(SRE_ERR_INVALID, "system not set");
  launch_rebuild_info(0, nullptr, nullptr, skin_part_count(e->n_atoms),
                      e->d_skin_part, 0, e->d_flags, e->stream, nullptr, nullptr, 0,
                      e->d_info_slices);
  unsigned bits = 0;
  HIP_TRY(hipMemcpyAsync(&bits, e->d_flags, sizeof(unsigned),
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  float d2;
  memcpy(&d2, &bits, sizeof d2);
  const float half = 0.5f * e->skin;
  if (d2 > half * half)
    return fail(SRE_ERR_SKIN, "max displacement %g exceeds skin/2 = %g",
                sqrt((double)d2), (double)half);
  return SRE_OK;
}

extern "C" sre_status sre_perturb_positions(sre_engine *e, float sigma,
                                            uint64_t seed, uint64_t step) {
  if (!e || !e->d_pos) return fail(SRE_ERR_INVALID, "no positions bound");
  HIP_TRY(hipSetDevice(e->device));
  launch_perturb(e->n_atoms, const_cast<float *>(e->d_pos), e->pos_stride,
                 sigma, seed, step, e->stream);
  return SRE_OK;
}

extern "C" sre_status sre_set_perturbation(sre_engine *e, float sigma,
                                           uint64_t seed, uint64_t first_step,
                                           int anchored) {
  if (!e || !(sigma >= 0.f) || (anchored != 0 && anchored != 1))
    return fail(SRE_ERR_INVALID, "bad arguments");
  if (anchored && !e->d_pos) return fail(SRE_ERR_INVALID, "no positions bound");
  HIP_TRY(hipSetDevice(e->device));
  if (anchored) {
    const size_t cnt = (size_t)e->n_atoms * e->pos_stride;
    if (e->anchor_cap < cnt) {
      dfree(e->d_anchor);
      HIP_TRY(dalloc(&e->d_anchor, cnt));
      e->anchor_cap = cnt;
      e->graph_epoch++;
    }
    HIP_TRY(hipMemcpyAsync(e->d_anchor, e->d_pos, sizeof(float) * cnt,
                           hipMemcpyDeviceToDevice, e->stream));
    e->anchor_count = cnt;
    e->anchor_stride = e->pos_stride;
    if (!e->d_anchor_s) {
      HIP_TRY(dalloc(&e->d_anchor_s, e->n_atoms));
      e->graph_epoch++;
    }
    // the anchor in the current cell order (later rebuilds write it too)
    if (e->built)
      launch_gather_anchor(e->n_atoms, e->sb.val_a, e->d_anchor, e->pos_stride,
                           e->d_anchor_s, e->stream);
    HIP_TRY(hipGetLastError());
  }
  if (e->pert_anchored != (anchored != 0)) e->graph_epoch++;
  e->pert_anchored = anchored != 0;
  if (!e->d_step_ctr) {
    HIP_TRY(dalloc(&e->d_step_ctr, 2));   // slot 0 current; slot 1 for fused graphs
    e->graph_epoch++;
  }
  const uint64_t ctr2[2] = {first_step, first_step};
  HIP_TRY(hipMemcpyAsync(e->d_step_ctr, ctr2, sizeof ctr2, hipMemcpyHostToDevice,
                         e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->pert_sigma != sigma || e->pert_seed != seed) e->graph_epoch++;
  e->pert_sigma = sigma;
  e->pert_seed = seed;
  return SRE_OK;
}

// ------------------------------------------------------------- integrator --

static sre_status md_buffers(sre_engine *e) {
  const int N = e->n_atoms;
  if (!e->d_vel) {
    HIP_TRY(dalloc(&e->d_vel, (size_t)3 * N));
    HIP_TRY(dalloc(&e->d_fcall, (size_t)3 * N));
    HIP_TRY(dalloc(&e->d_md_part, (size_t)4 * md_part_count(N)));
    e->graph_epoch++;
  }
  if (!e->d_md) {
    HIP_TRY(dalloc(&e->d_md, 1));
    HIP_TRY(dalloc(&e->d_mass, kMaxTypes));
    e->graph_epoch++;
  }
  if (e->mass.empty()) {
    e->mass.assign(kMaxTypes, 1.0f);
    HIP_TRY(hipMemcpyAsync(e->d_mass, e->mass.data(), sizeof(float) * kMaxTypes,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  }
  return SRE_OK;
}

// read-modify-write of the device MD state (stream-ordered)
template <typename Fn>
static sre_status md_patch(sre_engine *e, Fn fn) {
  MDState m;
  HIP_TRY(hipMemcpyAsync(&m, e->d_md, sizeof m, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  fn(m);
  HIP_TRY(hipMemcpyAsync(e->d_md, &m, sizeof m, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return SRE_OK;
}

static double md_total_mass(const sre_engine *e) {
  double M = 0.0;
  for (int t : e->h_types) M += e->mass[t];
  return M;
}

static sre_status md_update_total_mass(sre_engine *e) {
  const double M = md_total_mass(e);
  e->md.total_mass = M;
  return md_patch(e, [M](MDState &m) { m.total_mass = M; });
}

extern "C" sre_status sre_set_masses(sre_engine *e, const float *mass) {
  if (!e || !mass || !e->system_set)
    return fail(SRE_ERR_INVALID, "sre_set_masses: system not set");
  for (int t = 0; t < e->n_types; ++t)
    if (!(mass[t] > 0.f) || !std::isfinite(mass[t]))
      return fail(SRE_ERR_INVALID, "mass of type %d is %g", t, (double)mass[t]);
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = md_buffers(e);
  if (s != SRE_OK) return s;
  std::copy(mass, mass + e->n_types, e->mass.begin());
  HIP_TRY(hipMemcpyAsync(e->d_mass, e->mass.data(), sizeof(float) * kMaxTypes,
                         hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->md_set) return md_update_total_mass(e);
  return SRE_OK;
}

extern "C" sre_status sre_set_velocities(sre_engine *e, const float *v) {
  if (!e || !v || !e->system_set)
    return fail(SRE_ERR_INVALID, "sre_set_velocities: system not set");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = md_buffers(e);
  if (s != SRE_OK) return s;
  HIP_TRY(hipMemcpyAsync(e->d_vel, v, sizeof(float) * 3 * e->n_atoms,
                         hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->vel_set = true;
  return SRE_OK;
}

extern "C" sre_status sre_get_velocities(sre_engine *e, float *v) {
  if (!e || !v) return fail(SRE_ERR_INVALID, "sre_get_velocities: null argument");
  if (!e->vel_set) return fail(SRE_ERR_NOT_READY, "velocities not set");
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipMemcpyAsync(v, e->d_vel, sizeof(float) * 3 * e->n_atoms,
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return SRE_OK;
}

extern "C" sre_status sre_set_integrator(sre_engine *e, double dt,
                                         int thermostat, double ref_T,
                                         double tau_T, int remove_com) {
  if (!e || !e->system_set || !e->types_set)
    return fail(SRE_ERR_INVALID, "sre_set_integrator: system or types not set");
  if (e->n_rep != 1)
    return fail(SRE_ERR_INVALID, "the integrator runs one replica (have %d)",
                e->n_rep);
  if (!(dt > 0.0) || !std::isfinite(dt) ||
      (thermostat != SRE_THERMO_NONE && thermostat != SRE_THERMO_NHC2) ||
      (remove_com != 0 && remove_com != 1))
    return fail(SRE_ERR_INVALID, "sre_set_integrator: bad arguments");
  if (thermostat == SRE_THERMO_NHC2 &&
      !(ref_T > 0.0 && tau_T > 0.0 && std::isfinite(ref_T) && std::isfinite(tau_T)))
    return fail(SRE_ERR_INVALID, "NHC needs ref_T > 0 and tau_T > 0");
  HIP_TRY(hipSetDevice(e->device));
  sre_status s = md_buffers(e);
  if (s != SRE_OK) return s;
  MDState m{};
  // NoseHoover_Chains2::reinit (H/src/Thermostat.cu:23-53)
  m.LL = 3.0 * e->n_atoms;
  m.refT = thermostat == SRE_THERMO_NHC2 ? ref_T : 0.0;
  m.Q1 = thermostat == SRE_THERMO_NHC2
             ? tau_T * tau_T * ref_T / (4.0 * M_PI * M_PI) * m.LL : 1.0;
  m.Q2 = m.Q1 / m.LL;
  m.xi1 = m.xi2 = 1.0;
  m.vxi1 = m.vxi2 = 0.0;
  m.a1 = m.a2 = 1.0;
  m.dt = dt;
  m.total_mass = md_total_mass(e);
  m.thermostat = thermostat;
  m.remove_com = remove_com;
  m.have_end = 0;
  if (e->md_set && e->md.dt != dt) e->graph_epoch++;  // dt is a kernel argument
  e->md = m;
  HIP_TRY(hipMemcpyAsync(e->d_md, &m, sizeof(MDState), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->md_set = true;
  return SRE_OK;
}

extern "C" sre_status sre_get_md_state(sre_engine *e, sre_md_state *out) {
  if (!e || !out) return fail(SRE_ERR_INVALID, "sre_get_md_state: null argument");
  if (!e->md_set) return fail(SRE_ERR_NOT_READY, "integrator not set");
  HIP_TRY(hipSetDevice(e->device));
  MDState m;
  HIP_TRY(hipMemcpyAsync(&m, e->d_md, sizeof m, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  out->kinetic = m.K;
  out->xi1 = m.xi1;
  out->xi2 = m.xi2;
  out->vxi1 = m.vxi1;
  out->vxi2 = m.vxi2;
  // NoseHoover_Chains2::HamiltonianContribution (H/src/Thermostat.cu:246-253)
  out->nhc_energy = m.thermostat
                        ? 0.5 * m.Q1 * m.vxi1 * m.vxi1 +
                              0.5 * m.Q2 * m.vxi2 * m.vxi2 +
                              m.LL * m.refT * m.xi1 + m.refT * m.xi2
                        : 0.0;
  out->Q1 = m.Q1;
  out->Q2 = m.Q2;
  out->adapt_steps = e->md_adapt_k;
  out->rollbacks = e->md_rollbacks;
  return SRE_OK;
}

extern "C" sre_status sre_set_md_state(sre_engine *e, const sre_md_state *in) {
  if (!e || !in) return fail(SRE_ERR_INVALID, "sre_set_md_state: null argument");
  if (!e->md_set) return fail(SRE_ERR_NOT_READY, "integrator not set");
  HIP_TRY(hipSetDevice(e->device));
  const sre_md_state c = *in;
  return md_patch(e, [c](MDState &m) {
    m.xi1 = c.xi1;
    m.xi2 = c.xi2;
    m.vxi1 = c.vxi1;
    m.vxi2 = c.vxi2;
  });
}

// largest displacement^2 since the last rebuild against skin/2 (*ok), and
// the force-range flag, in one readback.  A skin violation is reported first:
// missed pairs are the likelier cause of a range error that follows it.
static sre_status md_skin_ok(sre_engine *e, bool *ok, float *disp2 = nullptr) {
  launch_rebuild_info(0, nullptr, nullptr, skin_part_count(e->n_atoms),
                      e->d_skin_part, 0, e->d_flags, e->stream, nullptr, nullptr, 0,
                      e->d_info_slices);
  unsigned bits[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(bits, e->d_flags, 2 * sizeof(unsigned),
                         hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  float d2;
  memcpy(&d2, &bits[0], sizeof d2);
  const float half = 0.5f * e->skin;
  *ok = !(d2 > half * half);
  if (disp2) *disp2 = d2;
  if (*ok && (bits[1] & 1u)) return report_force_range(e);
  return SRE_OK;
}

// Saves (restore = false) or restores the state an MD stretch starts from:
// bound positions, velocities, saved forces, kick sums and the chain.
static sre_status md_snapshot(sre_engine *e, bool restore) {
  const int N = e->n_atoms;
  const size_t npos = (size_t)N * e->pos_stride;
  if (!e->d_snap_vel || e->snap_pos_cap < npos) {
    dfree(e->d_snap_pos);
    HIP_TRY(dalloc(&e->d_snap_pos, npos));
    e->snap_pos_cap = npos;
    if (!e->d_snap_vel) {
      HIP_TRY(dalloc(&e->d_snap_vel, (size_t)3 * N));
      HIP_TRY(dalloc(&e->d_snap_fcall, (size_t)3 * N));
      HIP_TRY(dalloc(&e->d_snap_part, (size_t)4 * md_part_count(N)));
    }
    if (!e->d_snap_md) HIP_TRY(dalloc(&e->d_snap_md, 1));
  }
  struct Pair { void *live; void *snap; size_t bytes; };
  const Pair pairs[] = {
      {const_cast<float *>(e->d_pos), e->d_snap_pos, sizeof(float) * npos},
      {e->d_vel, e->d_snap_vel, sizeof(float) * 3 * N},
      {e->d_fcall, e->d_snap_fcall, sizeof(float) * 3 * N},
      {e->d_md_part, e->d_snap_part, sizeof(double) * 4 * md_part_count(N)},
      {e->d_md, e->d_snap_md, sizeof(MDState)}};
  static_assert(sizeof(MDState) % 4 == 0, "MDState copies as 4-byte words");
  CopySegs c{};
  for (const Pair &p : pairs) {
    c.src[c.count] = static_cast<const uint32_t *>(restore ? p.snap : p.live);
    c.dst[c.count] = static_cast<uint32_t *>(restore ? p.live : p.snap);
    c.words[c.count] = p.bytes / 4;
    ++c.count;
  }
  launch_md_copy(c, e->stream);     // one launch for all five buffers
  if (restore) clear_force_flag(e);
  HIP_TRY(hipGetLastError());
  return SRE_OK;
}

// H/app/lj.nhc-xtc.cu:158-196: n_steps of [COM removal, NHC(dt/2), kick +
// drift, (rebuild), force, kick, NHC(dt/2)].
extern "C" sre_status sre_md_run(sre_engine *e, int n_steps, int rebuild_every,
                                 unsigned flags, int *n_rebuilds) {
  if (!e || n_steps < 0 || rebuild_every < SRE_MD_ADAPTIVE ||
      (flags & ~(unsigned)(SRE_STATS | SRE_TAIL | SRE_FCORR)))
    return fail(SRE_ERR_INVALID, "sre_md_run: bad arguments");
  if (!e->md_set) return fail(SRE_ERR_NOT_READY, "integrator not set");
  if (!e->vel_set) return fail(SRE_ERR_NOT_READY, "velocities not set");
  if (n_rebuilds) *n_rebuilds = 0;
  HIP_TRY(hipSetDevice(e->device));
  sre_status s;
  bool fresh = !e->built;
  if (e->built) {
    // the bound positions may have been changed since the last step:
    // refresh the lattice copy and rebuild if they moved beyond skin/2
    launch_update_positions(e->n_atoms, e->d_inv_perm,
                            const_cast<float *>(e->d_pos), e->pos_stride,
                            e->d_reps, rep0(e), e->d_pos4, e->d_backup, e->d_res4,
                            e->d_skin_part, nullptr, nullptr, e->stream);
    HIP_TRY(hipGetLastError());
    bool ok;
    s = md_skin_ok(e, &ok);
    if (s != SRE_OK) return s;
    fresh = !ok;
  }
  if (fresh) {
    s = do_rebuild(e);
    if (s != SRE_OK) return s;
    if (n_rebuilds) ++*n_rebuilds;
  }
  refresh_tail_values(e);
  const unsigned fl = flags & (SRE_TAIL | SRE_FCORR);
  // forces at the current positions, then a kick of length 0: saves them in
  // caller order and leaves the kinetic energy and momentum sums
  s = launch_compute(e, fl);
  if (s != SRE_OK) return s;
  enqueue_md_kick(e, 0.f);
  HIP_TRY(hipGetLastError());
  if (n_steps == 0) {
    launch_md_finish(e->n_atoms, e->d_md_part, e->d_md, e->d_vel, e->stream);
    HIP_TRY(hipGetLastError());
    e->forces_valid = true;
    e->stats_valid = e->atom_stats_valid = false;
    e->stats_pending = false;
    bool ok;
    s = md_skin_ok(e, &ok);
    if (s != SRE_OK) return s;
    if (!ok)
      return fail(SRE_ERR_SKIN,
                  "an atom moved more than skin/2 since the last rebuild");
    return SRE_OK;
  }
  const float half = 0.5f * e->skin;
  bool verified = false;
  if (rebuild_every == SRE_MD_ADAPTIVE && e->md_since_rebuild > 0) {
    // every speculative stretch starts from a list built at its first step
    s = do_rebuild(e);
    if (s != SRE_OK) return s;
    if (n_rebuilds) ++*n_rebuilds;
  }
  int left = n_steps;
  while (left > 0) {
    if (rebuild_every == SRE_MD_ADAPTIVE) {
      // a stretch of K steps as one graph from a snapshot; the displacement
      // read back at its end decides: kept (next K aims at 80 % of skin/2,
      // displacements growing about linearly) or rolled back and redone
      // with K/2.  Each kept stretch used complete lists throughout.
      const int K = std::min({left, e->md_adapt_k, kMaxGraphSteps});
      const bool last = K == left;
      s = md_snapshot(e, false);
      if (s != SRE_OK) return s;
      hipGraphExec_t ex;
      s = step_graph(e, kGraphMD | (last ? flags : fl), K, &ex);
      if (s != SRE_OK) return s;
      HIP_TRY(hipGraphLaunch(ex, e->stream));
      e->md_since_rebuild += K;
      bool ok;
      float d2;
      if (!last) {
        s = do_rebuild(e);   // reads the stretch's displacement back
        if (s != SRE_OK) return s;
        if (n_rebuilds) ++*n_rebuilds;
        d2 = e->last_disp2;
        ok = !(d2 > half * half);
      } else {
        s = md_skin_ok(e, &ok, &d2);
        if (s != SRE_OK) return s;
      }
      if (!ok) {
        if (K == 1)
          return fail(SRE_ERR_SKIN, "an atom moved %g > skin/2 in one step",
                      sqrt((double)d2));
        s = md_snapshot(e, true);
        if (s != SRE_OK) return s;
        s = do_rebuild(e);
        if (s != SRE_OK) return s;
        if (n_rebuilds) ++*n_rebuilds;
        e->md_adapt_k = std::max(1, K / 2);
        ++e->md_rollbacks;
        continue;
      }
      left -= K;
      const double d = sqrt((double)d2);
      const double k_next = d > 0.0 ? K * 0.8 * half / d : (double)kMaxGraphSteps;
      e->md_adapt_k = (int)std::max(1.0, std::min((double)kMaxGraphSteps, floor(k_next)));
      verified = last;
      continue;
    }
    if (rebuild_every == 0) {
      // the reference's rule: rebuild when the largest displacement since
      // the last rebuild exceeds skin/2 (one readback per step)
      enqueue_md_drift(e);
      HIP_TRY(hipGetLastError());
      bool ok;
      s = md_skin_ok(e, &ok);
      if (s != SRE_OK) return s;
      if (!ok) {
        // (a range flag raised before this point is seen by the next check)
        s = do_rebuild(e);
        if (s != SRE_OK) return s;
        if (n_rebuilds) ++*n_rebuilds;
      }
      s = launch_compute(e, left == 1 ? flags : fl);
      if (s != SRE_OK) return s;
      enqueue_md_kick(e, (float)(0.5 * e->md.dt));
      HIP_TRY(hipGetLastError());
      ++e->md_since_rebuild;
      --left;
      continue;
    }
    if (e->md_since_rebuild >= rebuild_every) {
      // fixed interval: every step since the last rebuild must have stayed
      // within skin/2 (the rebuild reads the largest displacement back),
      // else the lists of that stretch were incomplete
      s = do_rebuild(e);
      if (s != SRE_OK) return s;
      if (n_rebuilds) ++*n_rebuilds;
      if (e->last_disp2 > half * half)
        return fail(SRE_ERR_SKIN,
                    "an atom moved %g > skin/2 within %d steps: lower "
                    "rebuild_every or raise the skin",
                    sqrt((double)e->last_disp2), rebuild_every);
    }
    const int K = std::min({left, rebuild_every - e->md_since_rebuild,
                            kMaxGraphSteps});
    const unsigned gflags = kGraphMD | (K == left ? flags : fl);
    hipGraphExec_t ex;
    s = step_graph(e, gflags, K, &ex);
    if (s != SRE_OK) return s;
    HIP_TRY(hipGraphLaunch(ex, e->stream));
    e->md_since_rebuild += K;
    left -= K;
  }
  launch_md_finish(e->n_atoms, e->d_md_part, e->d_md, e->d_vel, e->stream);
  HIP_TRY(hipGetLastError());
  e->forces_valid = true;
  e->stats_valid = e->atom_stats_valid = (flags & SRE_STATS) != 0;
  e->stats_pending = e->stats_valid;
  if (verified) return SRE_OK;
  bool ok;
  s = md_skin_ok(e, &ok);
  if (s != SRE_OK) return s;
  if (!ok)
    return fail(SRE_ERR_SKIN,
                "an atom moved more than skin/2 since the last rebuild: lower "
                "rebuild_every or raise the skin");
  return SRE_OK;
}

hipStream_t sre::engine_stream(sre_engine *e) { return e->stream; }

sre_status sre::engine_replica_positions(sre_engine *e, int replica,
                                         const float **pos, int *n, int *stride,
                                         float box[3]) {
  if (!e || !e->system_set || replica < 0 || replica >= e->n_rep)
    return fail(SRE_ERR_INVALID, "no replica %d", replica);
  if (!e->d_pos) return fail(SRE_ERR_INVALID, "positions not bound");
  const RepInfo &R = e->reps[replica];
  *pos = e->d_pos + (size_t)R.atom_off * e->pos_stride;
  *n = R.n_atoms;
  *stride = e->pos_stride;
  for (int d = 0; d < 3; ++d) box[d] = R.L[d];
  return SRE_OK;
}
