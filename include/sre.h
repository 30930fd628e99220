/*
 * sre.h -- C ABI of the MI355X short-range engine (SRE).
 *
 * The engine replaces the cell-list -> Verlet-list -> Lennard-Jones force,
 * energy and virial path of the reference's adaptive-cutoff GPU MD engine
 * (shortRange.inhomo/template/tools/simulation.hetero.cutoff, "H/" below),
 * including its per-atom cutoff rc(r) taken from a tabulated lattice, and the
 * per-bin mean error force of the force-correction engine
 * (shortRange.inhomo/template/tools/simulation.fcorr, "FC/" below).
 *
 * The reference has no plugin ABI; it is a set of C++ objects that own device
 * memory and run kernels on the default stream.  Every entry point below names
 * the reference member function it stands in for.  All functions return an
 * sre_status (0 = success); on failure sre_last_error() describes the cause.
 * No exception crosses this boundary.
 *
 * Conventions
 *  - One engine holds R >= 1 independent periodic systems ("replicas") that
 *    share one force field.  Atoms of replica r occupy the caller-order index
 *    range [atom_offset(r), atom_offset(r) + n_atoms[r]).  R = 1 is the
 *    reference's single MDSystem.
 *  - Positions are FP32 (x, y, z[, w]) like the reference's CoordType float4
 *    (H/include/systemDefines.h:48).  Any position is accepted: periodic
 *    wrapping is exact inside the engine.
 *  - Forces are produced as 64-bit fixed point with scale 2^32 (resolution
 *    2.3e-10 in LJ units), so they are bitwise independent of summation order.
 *  - Energy and virial follow H/src/InteractionEngine_interface.cu:1276-1279:
 *    per atom 1/2 sum_j u_ij and 1/2 sum_j F_ij,a * (x_j - x_i)_a (diagonal a).
 *  - Every launch goes on the engine's HIP stream (sre_options.stream), so a
 *    caller may capture the per-step calls into a hipGraph.
 */
#ifndef SRE_H
#define SRE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sre_engine sre_engine;

/* Status codes.  The first three mirror the reference's device error flags
 * (H/include/error.h:19-26). */
typedef enum {
  SRE_OK = 0,
  SRE_ERR_LIST_OVERFLOW = 1,   /* mdErrorShortNeighborList: only reported if regrowing failed */
  SRE_ERR_CELL = 2,            /* mdErrorShortCellList: cell grid too large for the sort key */
  SRE_ERR_OUT_OF_BOX = 3,      /* mdErrorOverFlowCellIdx: never raised (wrapping is exact); kept for ABI parity */
  SRE_ERR_INVALID = 4,         /* bad argument or call order */
  SRE_ERR_HIP = 5,             /* a HIP runtime call failed (cf. checkCUDAError, H/include/error.h:8-16) */
  SRE_ERR_BOX_TOO_SMALL = 6,   /* rc_max + skin >= L/2 in some direction: minimum image is ambiguous */
  SRE_ERR_FORCE_RANGE = 7,     /* pair forces beyond the exact fixed-point range (overlap);
                                  reported once by the read that finds it, which also marks the
                                  forces and statistics invalid; the flag is cleared then and
                                  by sre_set_system, sre_upload_positions, sre_bind_positions
                                  and an MD rollback, so the next compute starts clean */
  SRE_ERR_SKIN = 8,            /* an atom moved more than skin/2 since the last rebuild (fixed-period mode) */
  SRE_ERR_NOT_READY = 9,       /* compute requested before the first rebuild */
  SRE_ERR_IO = 10,             /* a file could not be opened, read or written
                                  (MDExcptCannotOpenFile, H/include/MDException.h) */
  SRE_END_OF_FILE = 11         /* sre_xtc_read: no further frame (not an error) */
} sre_status;

typedef enum {
  SRE_PREC_MIXED = 0,  /* FP32 pair math, int64 fixed-point forces, FP64 energy/virial reduction */
  SRE_PREC_FP64 = 1    /* FP64 pair math and accumulation ("plain FP64 GPU form") */
} sre_precision;

/* compute flags */
#define SRE_STATS 1u        /* also energy and virial (the statistic kernel variant) */
#define SRE_TAIL 2u         /* add the per-atom inhomogeneous tail correction */
#define SRE_FCORR 8u        /* add the grid force correction (sre_set_force_grid) */

/* tail-correction evaluation modes (sre_set_tail_mode) */
#define SRE_TAIL_BIN 0      /* value at the rc-bin centre, assigned per atom by bin (FC/src/AssignForceCorr.cu:95-128) */
#define SRE_TAIL_EXACT 1    /* value evaluated at each atom's own z */

typedef struct {
  int n_types;          /* atom types (1..8) */
  float skin;           /* Verlet skin ("rlistExten", H/app/lj.nhc-xtc.cu:44) */
  int precision;        /* sre_precision */
  void *stream;         /* hipStream_t to launch on; NULL = engine creates its own */
  int device;           /* HIP device ordinal; -1 = current */
} sre_options;

/* ---- lifetime ------------------------------------------------------------ */

/* Creates an engine.  Replaces the constructors of MDSystem device data,
 * CellList, NeighborList and InteractionEngine (H/src/MDSystem_interface.cu:146-171,
 * H/include/CellList_interface.h:47-66, H/include/NeighborList_interface.h:61-80,
 * H/include/InteractionEngine_interface.h:67-111). */
sre_status sre_create(sre_engine **out, const sre_options *opt);
void sre_destroy(sre_engine *e);
const char *sre_last_error(void);
const char *sre_version(void);

/* ---- system description -------------------------------------------------- */

/* Declares R replicas with n_atoms[r] atoms and box edge lengths box[3r..3r+2]
 * (fewer than 2^27 atoms in total; batched replicas need >= 64 atoms each).
 * Replaces MDSystem::initConfig/initDeviceData and RectangularBox setup
 * (H/src/MDSystem_interface.cu:39-171, H/include/BoxGeometry.h:10-14). */
sre_status sre_set_system(sre_engine *e, int n_replicas, const int *n_atoms,
                          const float *box);

/* Per-atom types (host array, caller order, values in [0, n_types)).
 * Replaces MDSystem::initTopology (H/src/MDSystem_interface.cu:115-133). */
sre_status sre_set_types(sre_engine *e, const int32_t *types);

/* Type-pair LJ parameters, full symmetric n_types x n_types matrices (host).
 * u(r) = 4 eps [(sig/r)^12 - (sig/r)^6] + shift.  Replaces
 * SystemNonBondedInteraction::add + InteractionEngine::registNonBondedInteraction
 * (H/src/SystemNonBondedInteraction.cu:235-330, H/src/InteractionEngine_interface.cu:92-166,
 *  H/include/NonBondedInteraction.h:866-889).  shift may be NULL (0). */
sre_status sre_set_pair_params(sre_engine *e, const float *eps,
                               const float *sigma, const float *shift);

/* Cutoff lattice for replica r: rc[iz + nz*(iy + ny*ix)], FP32.  Replaces
 * AssignRCut::reinit/getRCut/uniform (H/src/AssignRCut.cu:17-63).  The lattice
 * is assigned to atoms at every rebuild with the reference's bin rule
 * (H/src/AssignRCut.cu:77-104), or at sre_assign_rcut calls once that is
 * used. */
sre_status sre_set_rc_table(sre_engine *e, int replica, int nx, int ny, int nz,
                            const float *rc);

/* Density profile rho_t(z) for replica r, t = 0..n_types-1, nz bins over
 * [0, Lz) (host, FP64, layout rho[t*nz + iz]).  Input of the tail correction,
 * the analogue of ForceCorr::calError's density input (FC/src/ForceCorr.cu:131-162). */
sre_status sre_set_density_profile(sre_engine *e, int replica, int nz,
                                   const double *rho);
sre_status sre_set_tail_mode(sre_engine *e, int mode);

/* ---- positions -------------------------------------------------------------- */

/* Copies host positions (stride floats per atom, >= 3) into an engine-owned
 * device array and binds it.  Replaces sys.ddata.coord uploads
 * (H/src/MDSystem_interface.cu:146-171). */
sre_status sre_upload_positions(sre_engine *e, const float *xyz, int stride);
/* Binds a caller-owned device array (stride floats per atom).  The engine
 * reads it on every sre_update_positions / sre_rebuild. */
sre_status sre_bind_positions(sre_engine *e, const void *d_xyz, int stride);

/* ---- the hot path ------------------------------------------------------------ */

/* Assigns rc to every atom, bins atoms into cells with an atomic-free counting
 * sort and builds the Verlet list j in L(i) <=> j != i and |d_ij|^2 < (rc_i + skin)^2.
 * Replaces AssignRCut::assign + normalizeDeviceData + CellList::rebuild +
 * NeighborList::rebuild + Displacement_max::recordCoord
 * (H/app/lj.nhc-xtc.cu:170-181, H/src/NeighborList.cu:327-420,
 *  U/src/NeighborList_interface.cu:361-491).  Regrows the list on overflow. */
sre_status sre_rebuild(sre_engine *e);

/* Refreshes device positions from the bound array and tests the skin (max
 * displacement since the last rebuild vs skin/2, reduced per wavefront).
 * If need_rebuild != NULL the call synchronises and reports the decision.
 * Replaces Displacement_max::calMaxDisplacemant (H/src/Displacement_interface.cu:79-126). */
sre_status sre_update_positions(sre_engine *e, int *need_rebuild);

/* Forces (and with SRE_STATS energy/virial; with SRE_TAIL the tail term).
 * Replaces InteractionEngine::clearInteraction + applyNonBondedInteraction
 * (H/src/InteractionEngine_interface.cu:218-225, 240-382). */
sre_status sre_compute(sre_engine *e, unsigned flags);

/* update -> (rebuild if the skin test fires) -> compute; the reference loop
 * body H/app/lj.nhc-xtc.cu:169-182.  *rebuilt (optional) reports a rebuild. */
sre_status sre_step(sre_engine *e, unsigned flags, int *rebuilt);

/* ---- results -------------------------------------------------------------- */

/* Forces in caller order as double[N][3].  dst_on_device != 0: dst is a
 * device pointer (no host sync).  Otherwise the call synchronises. */
sre_status sre_get_forces(sre_engine *e, double *dst, int dst_on_device);
/* Reference-compatible FP32 SoA accumulation forc{x,y,z}[i] += F_i (device
 * pointers), as calNonBondedInteraction_neighbor writes them
 * (H/src/InteractionEngine_interface.cu:1270-1273). */
sre_status sre_add_forces_f32_soa(sre_engine *e, float *d_fx, float *d_fy,
                                  float *d_fz);
/* Raw fixed-point forces in caller order, int64[N][3] (scale 2^32), host. */
sre_status sre_get_forces_fixed(sre_engine *e, int64_t *dst);

/* Per replica: out[4r+0] = energy, out[4r+1..3] = virial xx, yy, zz (FP64),
 * from the last SRE_STATS compute.  Replaces MDStatistic::updateHost +
 * nonBondedEnergy (H/include/Statistic_interface.h:73-117). */
sre_status sre_get_stats(sre_engine *e, double *out);
/* Per-atom energy and virial diagonal (caller order, double[N][4]); the
 * reference's per-atom statistic buffers (H/src/InteractionEngine_interface.cu:1276-1279). */
sre_status sre_get_atom_stats(sre_engine *e, double *out);

/* Sum of neighbor-list lengths (the reference's "tc" counter,
 * NeighborList::calSumNeighbor, H/src/NeighborList_interface.cu:652-662),
 * pairs of the list inside rc_i at the current positions (counted on demand
 * with the force kernel's mask), and the largest list. */
sre_status sre_neighbor_count(sre_engine *e, int64_t *n_list, int64_t *n_in_rc,
                              int32_t *max_list);
/* Neighbor lists as CSR in caller order with caller indices, rows sorted
 * ascending: offsets[N+1], j[n_list].  Host arrays. */
sre_status sre_export_pairs(sre_engine *e, int64_t *offsets, int32_t *j);
/* Per-atom cutoff rc_i (caller order, host float[N]) as last assigned. */
sre_status sre_get_rcut(sre_engine *e, float *rc);
/* Replaces AssignRCut::assign (H/src/AssignRCut.cu:107-117; the adaptive
 * driver calls it every rcutAssignFeq steps, H/app/lj.nhc-xtc.cu:184-188):
 * every atom's rc_i from the current rc tables at its current position,
 * without rebuilding the lists.  By default the engine assigns rc_i at each
 * rebuild instead, so that lists always hold every pair within rc_i + skin;
 * after the first call it keeps each atom's assigned rc_i across rebuilds
 * until the next call, as the reference does (a list then misses the pairs
 * of an atom whose rc_i grew beyond its rc_i + skin of the last rebuild,
 * until the next rebuild).  rc gradients (sre_adjoint g_rc) are refused in
 * this mode.  sre_set_system returns to assignment at rebuild. */
sre_status sre_assign_rcut(sre_engine *e);

/* ---- adjoint (no reference counterpart) ----------------------------------- */

/* Gradient of  L = sum_r dLdE[r] * E_r + sum_i dLdF_i . F_i  with respect to
 * eps[a][b], sigma[a][b] (n_types^2 each, symmetric pairs summed into both
 * entries) and every rc lattice value (concatenated over replicas).  Pair
 * terms are recomputed from the neighbor list.  The cutoff mask is held fixed,
 * so d/d rc acts through the tail correction only.  dLdF is a device pointer
 * (double[N][3], caller order) or NULL; outputs are host arrays. */
sre_status sre_adjoint(sre_engine *e, const double *dLdE, const double *d_dLdF,
                       unsigned flags, double *g_eps, double *g_sigma,
                       double *g_rc);

/* ---- tuning, graphs, utilities ------------------------------------------- */

typedef struct {
  int force_lanes;      /* lanes per atom in the force kernel: 1,2,4,8,16 */
  int build_lanes;      /* lanes per atom in the list build */
  int block;            /* threads per block of the force kernel: 64..1024 */
  float cell_xy;        /* cell edge in x,y as a multiple of (rc_min + skin) */
  float cell_z;         /* cell edge in z as a multiple of (rc_min + skin) */
  int build_mode;       /* 0: per-atom column search (build_lanes lanes per atom);
                           1: one wavefront per 64-atom tile of a cell column */
  int xcd_remap;        /* 1: map contiguous atom ranges to each XCD (L2 locality) */
  int force_unroll;     /* list rows per group in the mixed force kernel: 1, 2, 4;
                           FP64 form: 1 = the mixed kernel's pipelined loop,
                           2 or 4 = the plain per-lane loop (same sums) */
  int persistent;       /* mixed force kernel: 0 = one wavefront per tile of 64/lanes
                           atoms (flat launch over all replicas); k = 1..16: a
                           persistent launch of k wavefronts per SIMD, each taking a
                           contiguous chunk of tiles (same sums) */
  int sort_mode;        /* binning at a rebuild (same cell order in every mode):
                           0 = counting sort by cell up to 15360 cells, radix sort
                           above; 1 = LSD radix sort of the cell key; 2 = counting
                           sort at any cell count */
  int force_mode;       /* mixed form: 0 = neighbor-list kernel; 1 = cell-window
                           kernel without a list (one wavefront per 16-atom tile of
                           a cell column, four lanes per atom, candidates staged in
                           LDS; bitwise the same forces, measured slower; needs
                           build_mode 0) */
} sre_tuning;
sre_status sre_get_tuning(sre_engine *e, sre_tuning *t);
/* Changing cell sizes or lanes takes effect at the next rebuild. */
sre_status sre_set_tuning(sre_engine *e, const sre_tuning *t);
/* Times candidate tunings on the current system and keeps the fastest
 * (forces are bitwise identical across candidates).  cache_path (optional)
 * stores/reads results keyed by the system's signature. */
sre_status sre_autotune(sre_engine *e, unsigned flags, const char *cache_path);

/* Captures [perturb +] update_positions + compute(flags) on the engine stream
 * into a hipGraph and replays it n times; the skin test only records
 * violations (SRE_ERR_SKIN is returned by sre_check_skin_flag).  With
 * SRE_PERTURB in flags every replay first applies the perturbation set by
 * sre_set_perturbation, advancing its step counter on the device. */
#define SRE_PERTURB 4u
sre_status sre_graph_run(sre_engine *e, unsigned flags, int n_steps);
/* n_segments times: sre_graph_run(e, flags, steps) then a list rebuild, all
 * queued without a synchronisation in between (the rebuild of the reference
 * driver's loop, H/app/lj.nhc-xtc.cu:169-181, at a fixed period).  The checks
 * sre_rebuild makes after each rebuild are made once at the end: n_list[g]
 * (optional) receives the list length built at the end of segment g,
 * max_disp (optional) the largest displacement of any segment from its
 * rebuild positions (compare with skin/2).  If a rebuild's longest list
 * exceeded the capacity, the lists from it on were cut: the capacity is
 * regrown and the lists rebuilt, and SRE_ERR_LIST_OVERFLOW reports those
 * segments as invalid.  If an atom moved more than skin/2 within a segment,
 * SRE_ERR_SKIN names the first such segment (its forces may have missed
 * pairs); the outputs are filled in either case and the lists are current.
 * With SRE_STATS, sre_get_stats afterwards returns the energy and virial of
 * the last step (the final rebuild does not change them); forces and
 * per-atom statistics need a new compute. */
sre_status sre_run_segments(sre_engine *e, unsigned flags, int n_segments, int steps,
                            int64_t *n_list, float *max_disp);
/* Perturbation of the bound positions applied by SRE_PERTURB replays, one
 * Gaussian draw (std sigma per component, counter-based on seed and step) per
 * replay.  anchored = 0: a random walk, x += sigma N(0,1).  anchored = 1:
 * vibration about the bound positions as they are at this call (copied to
 * the device), x = x0 + sigma N(0,1), so displacements stay bounded however
 * many steps run. */
sre_status sre_set_perturbation(sre_engine *e, float sigma, uint64_t seed,
                                uint64_t first_step, int anchored);
sre_status sre_check_skin_flag(sre_engine *e);
/* Largest displacement of any atom between the previous rebuild and the last
 * one, as recorded by the skin test (read back by sre_rebuild itself, no
 * extra synchronisation).  A value above skin/2 means forces computed before
 * the last rebuild may have missed pairs. */
sre_status sre_last_rebuild_displacement(sre_engine *e, float *max_disp);

/* Gaussian displacement of the bound positions (device), std sigma, seeded,
 * counter-based: x += sigma * N(0,1) (the same deviates as a graph replay of
 * that step).  Stand-in for the integrator step outside the force path. */
sre_status sre_perturb_positions(sre_engine *e, float sigma, uint64_t seed,
                                 uint64_t step);

/* Average duration (ms) of n_iter back-to-back launches of the force kernel
 * alone (no statistic reduction), from HIP events on the engine stream. */
sre_status sre_time_compute(sre_engine *e, unsigned flags, int n_iter,
                            double *ms_per_iter);
/* Average duration (ms) of n_iter complete rebuilds (includes their host
 * synchronisation for the list-capacity check). */
sre_status sre_time_rebuild(sre_engine *e, int n_iter, double *ms_per_iter);
sre_status sre_synchronize(sre_engine *e);

/* Grid force correction of replica r: f[3*(iz + nz*(iy + ny*ix)) + a] (host,
 * FP32).  With SRE_FCORR every force evaluation adds to each atom the value
 * of the cell its position falls in, by the reference's FP32 rule
 * (AssignForceCorr::getForceCorr + assign, FC/src/AssignForceCorr.cu:44-145;
 * the add FC/src/InteractionEngine_interface.cu:1118).  Energy and virial are
 * not changed.  f = NULL removes the grid. */
sre_status sre_set_force_grid(sre_engine *e, int replica, int nx, int ny, int nz,
                              const float *f);

/* ---- MD integrator ------------------------------------------------------- */
/* The MD loop around the force path (H/app/lj.nhc-xtc.cu:158-196), device-
 * resident: per step the centre-of-mass velocity removal
 * (TranslationalFreedomRemover::remove, H/src/Integrator_interface.cu:40-68),
 * a half step of the two-link Nose-Hoover chain (NoseHoover_Chains2::
 * operator_L(dt/2), H/src/Thermostat.cu:255-276), velocity Verlet part 1
 * (H/src/Integrator.cu:273-324), the force, velocity Verlet part 2
 * (H/src/Integrator.cu:327-347) and the second chain half step.  Positions
 * are the bound ones (wrapped into [0, L) as they move), velocities and
 * masses are held by the engine, in caller order, FP32 as in the reference;
 * the chain variables and the kinetic-energy sums are FP64.  One replica. */
#define SRE_THERMO_NONE 0   /* NVE */
#define SRE_THERMO_NHC2 1   /* two-link Nose-Hoover chain */
typedef struct sre_md_state {
  double kinetic;       /* sum m v^2 / 2 at the end of the last step */
  double xi1, xi2;      /* chain positions (start at 1) */
  double vxi1, vxi2;    /* chain velocities (start at 0) */
  double nhc_energy;    /* HamiltonianContribution: Q1 vxi1^2/2 + Q2 vxi2^2/2
                           + 3N T xi1 + T xi2 (0 without thermostat) */
  double Q1, Q2;        /* chain masses, tau^2 T 3N / (4 pi^2) and Q1 / 3N */
  int adapt_steps;      /* length of the next SRE_MD_ADAPTIVE stretch (read only) */
  int rollbacks;        /* adaptive stretches redone so far (read only) */
} sre_md_state;
/* Per-type masses, n_types values > 0 (default 1). */
sre_status sre_set_masses(sre_engine *e, const float *mass);
/* Velocities v[3N] (host, caller order; xyz per atom). */
sre_status sre_set_velocities(sre_engine *e, const float *v);
sre_status sre_get_velocities(sre_engine *e, float *v);
/* Time step, thermostat (SRE_THERMO_*; NoseHoover_Chains2::reinit with
 * reference temperature ref_T and period tau_T), remove_com 1 = remove the
 * centre-of-mass velocity at the start of every step.  Resets the chain.
 * sre_set_system discards the integrator (its 3N degrees of freedom and
 * chain masses belong to the old system): call this again after it. */
sre_status sre_set_integrator(sre_engine *e, double dt, int thermostat,
                              double ref_T, double tau_T, int remove_com);
/* n_steps MD steps.  rebuild_every > 0: the lists are rebuilt every that many
 * steps and the steps in between run as captured graphs; SRE_ERR_SKIN if an
 * atom moved more than skin/2 between two rebuilds (the forces of that
 * stretch may have missed pairs).  rebuild_every = 0: the reference's rule,
 * a rebuild whenever the largest displacement since the last one exceeds
 * skin/2 (one host synchronisation per step).  rebuild_every =
 * SRE_MD_ADAPTIVE: stretches of K steps run as graphs from a snapshot and are
 * kept only if no atom moved more than skin/2 (else restored and redone with
 * K/2); K adapts to the measured displacement.  Forces are exact fixed-point
 * sums, so all three give the same trajectory.  flags: SRE_TAIL adds the tail
 * force every step; SRE_STATS computes energy/virial on the last step (read
 * by sre_get_stats).  Forces are recomputed at the start of every call. */
#define SRE_MD_ADAPTIVE (-1)
sre_status sre_md_run(sre_engine *e, int n_steps, int rebuild_every,
                      unsigned flags, int *n_rebuilds);
sre_status sre_get_md_state(sre_engine *e, sre_md_state *out);
/* Restores the chain positions and velocities (checkpoint restart). */
sre_status sre_set_md_state(sre_engine *e, const sre_md_state *in);

/* ---- adaptive-cutoff producer -------------------------------------------- */
/* The rc lattice producer of the adaptive method on the GPU: a density grid
 * fed by deposits of positions (DensityProfile_PiecewiseConst,
 * H/src/DensityProfile.cu:27-99), the error estimate of every rc candidate
 * (AdaptRCut::reinit/makeS1k/makeS2k/calError, H/src/AdaptRCut.cu:96-416),
 * the per-cell choice of rc (calRCut, :418-468) and its widening (refineRCut,
 * :470-508).  One grid per box; LJ units with eps = sigma = 1 as in the
 * reference (makeS1k(1., 1.)). */
typedef struct sre_arc sre_arc;
/* Grid of n_d = int(L_d / refh) cells per edge, made odd (reinit).  device -1:
 * the current device; stream NULL: a stream of its own. */
sre_status sre_arc_create(sre_arc **out, const double box[3], double refh,
                          int device, void *stream);
void sre_arc_destroy(sre_arc *a);
sre_status sre_arc_dims(sre_arc *a, int *dims /* [3] */, int *n_rc);
/* Candidates rcmin, rcmin + rcstep, ... <= rcmax (rcmax < 30) and their error
 * kernels (AdaptRCut::reinit; computed on the device). */
sre_status sre_arc_set_candidates(sre_arc *a, double rcmin, double rcmax,
                                  double rcstep);
sre_status sre_arc_rc_list(sre_arc *a, double *rcs /* [n_rc] */);
/* Density samples: clearData, deposite (device positions, n atoms, stride >= 3
 * floats), the same from an engine replica's bound positions, a host density
 * in place of samples, and calculate() read back.  sre_arc_deposit reads
 * d_pos on the arc's stream (the caller orders it after the writes);
 * sre_arc_deposit_engine orders itself after the engine's queued work and
 * the engine's later work after it. */
sre_status sre_arc_clear(sre_arc *a);
sre_status sre_arc_deposit(sre_arc *a, const float *d_pos, int n, int stride);
sre_status sre_arc_deposit_engine(sre_arc *a, sre_engine *e, int replica);
sre_status sre_arc_set_density(sre_arc *a, const double *rho /* [nele] */);
sre_status sre_arc_density(sre_arc *a, double *rho /* [nele] */);
/* calError; keep_parts = 1 also keeps (e1, e2x, e2y, e2z) per candidate. */
sre_status sre_arc_cal_error(sre_arc *a, int keep_parts);
/* calRCut(prec): per cell the candidate index by the reference's bisection. */
sre_status sre_arc_cal_rcut(sre_arc *a, double prec);
/* refineRCut: index = max over the periodic 3x3x3 neighbourhood. */
sre_status sre_arc_refine(sre_arc *a);
/* Result per cell rc[iz + nz*(iy + ny*ix)] (FP32 as AssignRCut::getRCut
 * copies it), its error and candidate index; any pointer may be NULL. */
sre_status sre_arc_result(sre_arc *a, float *rc, double *err, int *idx);
/* load_rc: installs a saved candidate index per cell (values in [0, n_rc)). */
sre_status sre_arc_set_index(sre_arc *a, const int *idx /* [nele] */);
/* error[c][cell]; parts[c][cell][4] (needs keep_parts). */
sre_status sre_arc_errors(sre_arc *a, double *err, double *parts);
/* Kernel tables S1[c][cell] and G[c][cell] (S2_a = i k_a G). */
sre_status sre_arc_kernels(sre_arc *a, double *S1, double *G);
/* PressureCorrection::correction (H/src/PressureCorrection.cu:176-211): the
 * diagonal pressure correction p[3] = (pxx, pyy, pzz) of the pairs beyond each
 * cell's chosen rc, from the current density and candidates. */
sre_status sre_arc_pressure(sre_arc *a, double *p /* [3] */);
/* Its kernel integrals i1, i5[c][cell] = int_rc^30 -24 x^-4.5 J_{1/2, 5/2}(2 pi k x) dx. */
sre_status sre_arc_pressure_integrals(sre_arc *a, double *I1, double *I5);
/* ForceCorr (FC/src/ForceCorr.cu:131-162): the mean error force of each cell
 * at its chosen candidate, f[3*cell + a] (needs sre_arc_cal_error(a, 1)); a
 * single candidate (rcmin = rcmax) gives the uniform-rc force correction. */
sre_status sre_arc_force_grid(sre_arc *a, float *f);
/* AssignForceCorr::getForceCorr: sre_set_force_grid of that grid. */
sre_status sre_arc_apply_force_correction(sre_arc *a, sre_engine *e, int replica);
/* AssignRCut::getRCut: installs the result as the rc lattice of a replica. */
sre_status sre_arc_apply(sre_arc *a, sre_engine *e, int replica);

/* ---- one force evaluation without device allocation ------------------------ */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc): the allocation-free variant of sre_step (rebuild + compute)
 * followed by sre_get_forces(dst_on_device) and sre_get_stats, for one
 * periodic system in mixed precision with an rc(z) table and no tail term.
 * Every buffer comes from a 40 MiB device array held by each GPU the library
 * runs on until the process ends; no call allocates device memory.  Calls
 * that overlap on one GPU race (they share that array).  Limits:
 * n_atoms <= SRE_STATIC_MAX_ATOMS, nz <= SRE_STATIC_MAX_RC_BINS, at most
 * SRE_STATIC_MAX_LIST neighbor-list entries per atom (else
 * SRE_ERR_LIST_OVERFLOW; no regrowing).  Inputs: d_xyz device positions
 * (stride floats per atom), d_types device int32 types in [0, n_types) or
 * NULL (all 0), box[3], eps/sigma symmetric n_types^2 (host), the rc(z) table
 * rc_z[nz] (host), skin.  Outputs: d_forces device double[N][3] (caller
 * order; the same values as the engine path), stats host double[4] (energy,
 * virial xx, yy, zz) or NULL.  Synchronises the stream (NULL: the null
 * stream) before returning. */
#define SRE_STATIC_MAX_ATOMS 16384
#define SRE_STATIC_MAX_RC_BINS 4096
#define SRE_STATIC_MAX_LIST 512
sre_status sre_forces_static(int n_atoms, const float *d_xyz, int stride,
                             const int32_t *d_types, const float box[3], int n_types,
                             const float *eps, const float *sigma, int nz,
                             const float *rc_z, float skin, void *stream,
                             double *d_forces, double *stats);

/* ---- synthetic configurations (host only, no GPU use) --------------------- */

/* FCC lattice of ncell^3 unit cells with lattice constant a, every coordinate
 * jittered by U(-jitter, jitter).  xyz[4N] (w = 0). N = 4 ncell^3. */
sre_status sre_gen_fcc(int ncell, double a, double jitter, uint64_t seed,
                       float *xyz);
/* Liquid-vapour slab: n atoms, box (lx, ly, lz), liquid of density rho_l in
 * |z - lz/2| < h, vapour rho_v elsewhere, liquid count round(rho_l * lx*ly*2h),
 * uniform random positions with a minimum pair distance dmin (random
 * sequential addition), seeded.  xyz[4n]. */
sre_status sre_gen_slab(int n, double lx, double ly, double lz, double h,
                        double rho_l, double dmin, uint64_t seed, float *xyz);

/* ---- xtc trajectories (host only, no GPU use) ----------------------------- */

/* GROMACS xtc frames (compressed coordinates), the format the reference's
 * driver writes every few steps through libxdrfile: MDSystem::initWriteXtc
 * (xdrfile_open(name, "w"), H/src/MDSystem_interface.cu:243-263),
 * writeHostDataXtc (write_xtc(xd, natoms, step, time, box, x, prec),
 * :265-284) and endWriteXtc (xdrfile_close, :286-290).  Frames carry a
 * diagonal or general 3x3 box (row-major), natoms x 3 FP32 coordinates,
 * rounded to 1/precision (1000 when precision <= 0; frames of at most 9
 * atoms are stored uncompressed).  mode: "w" (truncate), "a" (append) or
 * "r".  sre_xtc_read returns SRE_END_OF_FILE after the last frame and
 * SRE_ERR_INVALID when the frame's atom count differs from natoms;
 * *precision is 0 for an uncompressed frame. */
typedef struct sre_xtc sre_xtc;
sre_status sre_xtc_open(const char *path, const char *mode, sre_xtc **out);
sre_status sre_xtc_write(sre_xtc *f, int natoms, int step, float time,
                         const float box[9], const float *xyz, float precision);
sre_status sre_xtc_read(sre_xtc *f, int natoms, int *step, float *time,
                        float box[9], float *xyz, float *precision);
/* atom count of the first frame (read_xtc_natoms) */
sre_status sre_xtc_natoms(const char *path, int *natoms);
sre_status sre_xtc_close(sre_xtc *f);

#ifdef __cplusplus
}
#endif
#endif /* SRE_H */
