/*
 * xnb_amd.h -- C ABI of the MI355X (gfx950) short-range particle hot path:
 * cell grid with periodic ghost layers, chunk_neighbors CS=1 neighbour stream,
 * Lennard-Jones pair traversal (FP64 plain form and mixed FP32-pair/FP64-accumulate),
 * velocity-Verlet push operators and move_particles re-binning.
 *
 * Conventions
 *  - Every entry point returns 0 (XNB_OK) or a negative XNB_E_* code; no C++
 *    exception crosses this boundary.  xnb_last_error() gives a message.
 *  - Host arrays are caller-owned and copied; the context owns all device memory
 *    and one HIP stream (default priority).  Calls are synchronous on return
 *    unless stated otherwise.  One host thread per context.
 *  - r, v, f host arrays are n x 3 interleaved doubles.
 *
 * Each declaration names the reference operator or function it replaces
 * (paths relative to the exaNBody source tree).
 */
#ifndef XNB_AMD_H
#define XNB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XNB_OK 0
#define XNB_E_ARG (-1)
#define XNB_E_DOMAIN (-2)
#define XNB_E_OUTSIDE (-3)
#define XNB_E_STATE (-4)
#define XNB_E_HIP (-5)
#define XNB_E_NOMEM (-6)
#define XNB_E_OVERFLOW (-7)
#define XNB_E_CAPACITY (-9)

#define XNB_PRECISION_FP64 0  /* reference arithmetic, FP64 throughout */
#define XNB_PRECISION_MIXED 1 /* FP32 pair arithmetic, FP64 positions, cutoff test and accumulation */

typedef struct xnb_ctx xnb_ctx;

/* Domain (src/core/include/exanb/core/domain.h:36-129): bounds, grid_dims, cell_size,
 * periodic flags.  Identity xform only. */
typedef struct {
  double bounds_min[3];
  double bounds_max[3];
  int64_t grid_dims[3];
  double cell_size;
  int periodic[3];
} xnb_domain;

/* LennardJonesParms (contribs/md/lennard_jones/lennard_jones.cu:40-44) */
typedef struct {
  double epsilon;
  double sigma;
} xnb_lj_params;

/* ChunkNeighborsConfig (src/particle_neighbors/include/exanb/particle_neighbors/
 * chunk_neighbors_config.h:27-39); chunk_size is always 1 (chunk_neighbors.cpp:62). */
typedef struct {
  int build_particle_offset;
  int dual_particle_offset;
  int random_access;
  int half_symmetric;
  int skip_ghosts;
} xnb_chunk_neighbors_config;

/* ---- library / context ---------------------------------------------------- */
int xnb_version(void);
const char* xnb_error_string(int code);
const char* xnb_last_error(const xnb_ctx* ctx);
int xnb_device_count(int* count);

/* domain + init_rcb_grid on one rank (src/grid_cell_particles/init_rcb_grid.cpp:32-84),
 * check_domain (src/core/lib/domain.cpp:78-85). */
int xnb_create(xnb_ctx** ctx, const xnb_domain* domain, int device);
int xnb_destroy(xnb_ctx* ctx);

/* nbh_dist (src/particle_neighbors/nbh_dist.cpp:45-72): sets nbh_dist_lab = rcut_max+rcut_inc,
 * max_displ = rcut_inc/2, ghost_dist = rcut_max+rcut_inc, and the grid's ghost layers
 * ceil(ghost_dist/cell_size) (src/core/include/exanb/core/grid.h:110). */
int xnb_nbh_dist(xnb_ctx* ctx, double rcut_max, double rcut_inc, double* nbh_dist_lab,
                 double* max_displ, double* ghost_dist);

/* Particle input (replaces a reader or lattice operator followed by the first
 * move_particles binning): particles are appended, in input order, to the cell given by
 * domain_periodic_location (src/core/include/exanb/core/domain.h:160-191).
 * v, id, type may be NULL (zero velocity, id = index, type 0). */
int xnb_set_particles(xnb_ctx* ctx, size_t n, const double* r, const double* v,
                      const uint64_t* id, const uint8_t* type);
/* A grid whose field set holds extra per-particle FP64 fields (a field set beyond
 * rx..type, as grid variants of make_grid_variant_operator.h:104-164 / grid_fields.h:114-142
 * carry them): xnb_set_particles plus n_extra (0..4) host arrays of n values in input order.
 * Every operator that moves, re-orders or images particles carries them (move_particles,
 * rebuild_amr, ghost_update_all, the batch engine, snapshots); no force operator reads them.
 * xnb_set_particles is this call with n_extra = 0 (r, v, id, type as there: type NULL is a
 * grid without the type field, every particle type 0; id NULL numbers them 0..n-1). */
int xnb_set_particles_with_fields(xnb_ctx* ctx, size_t n, const double* r, const double* v, const uint64_t* id,
                                  const uint8_t* type, int n_extra, const double* const* extra);
/* The first n_extra extra fields, in the order of xnb_get_particles (out[k] may be NULL). */
int xnb_get_fields(xnb_ctx* ctx, int include_ghosts, int n_extra, double* const* out);
/* Extra fields the grid holds (also after xnb_snapshot_load). */
int xnb_field_count(const xnb_ctx* ctx, int* n_extra);
int xnb_num_particles(const xnb_ctx* ctx, size_t* n_inner, size_t* n_ghost);
int xnb_grid_info(const xnb_ctx* ctx, int64_t dims[3], int* ghost_layers, size_t* n_cells);
/* Particles in cell order (ghost cells too if include_ghosts); any pointer may be NULL. */
int xnb_get_particles(xnb_ctx* ctx, int include_ghosts, uint32_t* cell, double* r, double* v,
                      double* f, double* ep, uint64_t* id, uint8_t* type);
int xnb_get_cell_counts(xnb_ctx* ctx, uint32_t* counts);

/* ---- operators (YAML names of the reference) ------------------------------ */
/* move_particles (src/grid_cell_particles/move_particles.cpp:39-47,
 * include/exanb/grid_cell_particles/move_particles_across_cells.h:78-235) */
int xnb_move_particles(xnb_ctx* ctx);
/* backup_r (src/io/backup_r.cpp:35-116) */
int xnb_backup_r(xnb_ctx* ctx);
/* rebuild_amr (src/amr/rebuild_amr.cpp:34-69; project_particles_in_sub_grids,
 * include/exanb/amr/amr_grid_algorithm.h:66-434): particles of each cell whose sub-grid side
 * sub_grid_size(n, sub_grid_density) exceeds 1 are stably reordered by sub-cell.  Only
 * enforced_ordering 1 (values <= 1); runs between move_particles and ghost_update_all, as in
 * the reference's update deck.  init, runs and batches apply it with density 6.5
 * (update-particles.msp:8-11). */
int xnb_rebuild_amr(xnb_ctx* ctx, double sub_grid_density, int enforced_ordering);
/* ghost_comm_scheme + ghost_update_all (src/mpi/update_ghosts_comm_scheme.cpp:84-489,
 * src/mpi/update_ghosts.cu:45-64) on one rank: periodic self images */
int xnb_ghost_update_all(xnb_ctx* ctx);
/* ghost_update_r (src/mpi/update_ghosts.cu:58) */
int xnb_ghost_update_r(xnb_ctx* ctx);
/* particle_displ_over (src/mpi/particle_displ_over.cu:38-218): count of particles moved
 * at least threshold since backup_r */
int xnb_particle_displ_over(xnb_ctx* ctx, double threshold, uint64_t* count);
/* chunk_neighbors (src/particle_neighbors/chunk_neighbors.cpp:47-90,
 * include/exanb/particle_neighbors/chunk_neighbors_execute.h:40-423), chunk size 1.
 * all_cells != 0 also builds lists of ghost cells (the reference builds every cell). */
int xnb_chunk_neighbors(xnb_ctx* ctx, double nbh_dist_lab, const xnb_chunk_neighbors_config* cfg,
                        int all_cells);
/* GridChunkNeighbors stream export (chunk_neighbors.h:42-120; Appendix-A format): per-cell
 * uint16 streams concatenated; cell_offsets[n_cells+1] in uint16 words.  out may be NULL to
 * query *total. */
int xnb_chunk_neighbors_stream(xnb_ctx* ctx, uint16_t* out, size_t cap, uint64_t* cell_offsets,
                               size_t* total, uint32_t* max_neighbors);
/* zero_particle_force (src/compute/zero_particle_force.cu:19-55) */
int xnb_zero_particle_force(xnb_ctx* ctx, int ghost);
/* divide_force_by_type_scalar (src/compute/vec3_typescalar_op.cu:52-120, 161): every inner
 * particle's force (ghosts too if ghost) divided by values[type], e.g. the type masses to turn
 * forces into accelerations.  1..256 values; a particle type without a value is XNB_E_ARG. */
int xnb_divide_force_by_type_scalar(xnb_ctx* ctx, int ntypes, const double* values, int ghost);
/* lennard_jones_force (contribs/md/lennard_jones/lennard_jones.cu:171-223) over the
 * compute_cell_particle_pairs traversal (src/compute/include/exanb/compute/
 * compute_cell_particle_pairs.h:119-243); adds to f and to per-atom energy ep. */
int xnb_lennard_jones_force(xnb_ctx* ctx, const xnb_lj_params* params, double rcut, int precision,
                            int ghost);
/* Multi-material LJ over compute_pair_multimat (src/compute/include/exanb/compute/
 * compute_pair_multimat.h:256-316): per unique type pair (particle_type_pair.h:39-55)
 * epsilon, sigma, rcut arrays of ntypes*(ntypes+1)/2 entries. */
int xnb_lennard_jones_force_multimat(xnb_ctx* ctx, int ntypes, const double* epsilon,
                                     const double* sigma, const double* rcut, int precision);
/* gravitational_force (contribs/pi/gravitational_force.cu:163-225, functor :44-82): over the
 * pairs within rcut, f += G m_a m_b (r_b - r_a) / r^3 with the masses given per particle type
 * (the "mass" type property; ntypes 1..16, every particle's type below ntypes, else
 * XNB_E_ARG), each pair added into the stored force in stream order (zero_particle_force
 * first for the force alone); ghost: the ghost cells too (lists over all cells).  Per-atom
 * energies are not touched.  Runs the library's pair traversal with the reference's
 * buffer-less functor form (include/xnb_compute_pair.h). */
int xnb_gravitational_force(xnb_ctx* ctx, double G, int ntypes, const double* type_mass, double rcut, int ghost);
/* push_f_v_r (src/defbox/include/exanb/defbox/push_vec3_2nd_order.h:29-119) */
int xnb_push_f_v_r(xnb_ctx* ctx, double dt, double dt_scale);
/* push_f_v (src/defbox/include/exanb/defbox/push_vec3_1st_order.h:29-108) */
int xnb_push_f_v(xnb_ctx* ctx, double dt, double dt_scale);
/* kinetic energy (mass 1, src/compute/simulation_stats.cpp:82-101) and PE = 1/2 sum ep */
int xnb_simulation_energies(xnb_ctx* ctx, double* ke, double* pe);
/* No counterpart in the reference.  xnb_simulation_energies for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * the same kernels and sums, bitwise, with the reduction partials in a fixed per-GPU device
 * array of 1024 (ke, pe) pairs (16 KiB) instead of a library allocation; every particle count
 * fits.  ke_pe is caller-owned DEVICE memory for 2 doubles (ke, pe), written asynchronously on
 * the context's stream (xnb_synchronize before reading it from the host).  Calls on one GPU
 * share the array: calls that overlap on one GPU (two contexts' streams) race. */
int xnb_simulation_energies_noalloc(xnb_ctx* ctx, double* ke_pe);

/* ---- composite step loop --------------------------------------------------- */
/* first_iteration (data/config/compute-loop.msp:1-7): move_particles, backup_r, ghosts,
 * chunk_neighbors, zero_particle_force + lennard_jones_force. */
int xnb_init(xnb_ctx* ctx, const xnb_lj_params* params, double rcut, double rcut_inc, int precision);

typedef struct {
  double dt;
  xnb_lj_params lj;
  double rcut;
  double rcut_inc;
  int precision;
  int rebuild_every; /* <= 0: displacement trigger (reference); > 0: fixed interval */
  int64_t steps;
  /* outputs */
  int64_t n_rebuilds;
  int64_t n_violations; /* interval mode: steps with an atom past max_displ and no rebuild */
} xnb_run_params;
/* numerical_scheme loop (data/config/numerical-scheme.msp:8-25, update-particles.msp:42-79):
 * push_f_v_r, push_f_v, trigger, full update or ghost_update_r, forces, push_f_v. */
int xnb_run(xnb_ctx* ctx, xnb_run_params* params);
/* xnb_run with compute_pair_multimat LJ forces (compute_pair_multimat.h:256-316; parameters
 * per unique type pair as for xnb_lennard_jones_force_multimat; params->lj and ->rcut are not
 * used; nbh_dist must cover the largest pair cutoff). */
int xnb_run_multimat(xnb_ctx* ctx, xnb_run_params* params, int ntypes, const double* eps_pair, const double* sigma_pair,
                     const double* rcut_pair);
/* particle_displ_over counts (particle_displ_over.cu:38-77) of every step of the last xnb_run /
 * xnb_run_multimat: counts[k - 1] = atoms past max_displ after step k's first half (before its
 * rebuild, if any).  n <= the steps of that run. */
int xnb_run_step_counts(xnb_ctx* ctx, uint64_t* counts, size_t n);

/* ---- measurement ----------------------------------------------------------- */
#define XNB_TIMER_FORCE 0
#define XNB_TIMER_NBH 1
#define XNB_TIMER_PUSH 2
#define XNB_TIMER_GHOST 3
#define XNB_TIMER_REBIN 4
#define XNB_TIMER_TRIGGER 5
#define XNB_TIMER_COUNT 6
/* HIP-event timing of each kernel class on the context's stream. */
int xnb_timer_enable(xnb_ctx* ctx, int enable);
int xnb_timer_reset(xnb_ctx* ctx);
int xnb_timer_read(xnb_ctx* ctx, int timer, double* total_ms, int64_t* launches);
/* Pairs seen by the last force call: list entries and in-cutoff pairs. */
int xnb_pair_counts(xnb_ctx* ctx, uint64_t* list_entries, uint64_t* cutoff_pairs);
int xnb_synchronize(xnb_ctx* ctx);
/* Device memory the context holds now and the most it has held at once (its own arrays). */
int xnb_device_bytes(const xnb_ctx* ctx, size_t* current, size_t* peak);
/* Device copy of `bytes` (read + write), iters times, timed with HIP events: achieved GB/s
 * counting read and write bytes.  Also the known-size kernel that calibrates HBM counters. */
int xnb_copy_probe(xnb_ctx* ctx, size_t bytes, int iters, double* gbps);
/* Diagnostic with no counterpart in the reference: evaluates on the device the FP64 square
 * root and reciprocal sequences the plain-form LJ kernels use for r = sqrt(d2) and 1.0 / r,
 * so tests can check them against IEEE results.  sqrt_out[i] is defined for x[i] in
 * [2^-767, 2^1000], recip_out[i] for x[i] in [2^-383, 2^383]. */
int xnb_fp64_core(xnb_ctx* ctx, const double* x, size_t n, double* sqrt_out, double* recip_out);

/* ---- autotuning ------------------------------------------------------------------ */
/* No counterpart in the reference.  Picks the brick shape (cells per workgroup tile; the tile's
 * atoms are the workgroup's lanes) of an initialised box by timing its LJ force kernel for each
 * candidate shape, and rebuilds its lists with the fastest; results are cached per process by
 * (cells per side up to 16, mean atoms per cell, precision), so boxes of a seen class are not
 * timed again.  Forces and energies are recomputed (bitwise the same: the shape does not enter
 * the arithmetic).  dims_out (optional) receives the shape, 0s off the brick path. */
int xnb_autotune(xnb_ctx* ctx, const xnb_lj_params* params, double rcut, int precision, int* dims_out);

/* ---- snapshots ------------------------------------------------------------------- */
/* Device-resident box -> host bytes and back, for preemption and restart (the fields of
 * write_dump / read_dump, write_dump.cpp:38,43: positions, velocities, ids, types), plus what
 * the step loop carries between steps (forces, per-atom energies, slot order, backup_r), so a
 * box restored into a fresh context (xnb_create with the same domain) continues bitwise.
 * Little-endian layout: header, per-cell counts, then r (x, y, z, type), v, f, ep, id, backup. */
int xnb_snapshot_size(xnb_ctx* ctx, size_t* bytes);
int xnb_snapshot_save(xnb_ctx* ctx, void* buf, size_t bytes);
int xnb_snapshot_load(xnb_ctx* ctx, const void* buf, size_t bytes);

/* ---- batches of boxes ---------------------------------------------------------- */
/* No counterpart in the reference, which advances one domain per process: many independent
 * periodic boxes advanced together on one GPU, each kernel of a step being one launch over
 * all boxes.  A box is an xnb_ctx after xnb_init (brick-path geometry: cell edge >=
 * nbh_dist); attached with the parameters xnb_run would take (fixed-interval or
 * displacement-triggered rebuilds), it follows xnb_run's operator sequence and ends bitwise
 * equal to the same box advanced by xnb_run alone.  Boxes may be attached and detached
 * between steps.  While a displacement-triggered box runs, each step reads the boxes'
 * trigger flags back to the host once. */
typedef struct xnb_batch xnb_batch;
int xnb_batch_create(xnb_batch** out, int device);
int xnb_batch_destroy(xnb_batch* batch);
const char* xnb_batch_error(const xnb_batch* batch);
/* Attach an initialised box for p->steps steps; *slot identifies it in the batch. */
int xnb_batch_attach(xnb_batch* batch, xnb_ctx* box, const xnb_run_params* p, int* slot);
/* Advance every unfinished box by up to nsteps steps; *n_active = unfinished boxes left. */
int xnb_batch_step(xnb_batch* batch, int nsteps, int* n_active);
/* Steps done and state (0 running, 1 finished, -1 failed) of the box in a slot. */
int xnb_batch_box_status(xnb_batch* batch, int slot, int64_t* done, int* state);
/* Rebuilds of the box that outgrew the sizes it was attached with and were redone by the
 * one-box operators (which re-size its arrays). */
int xnb_batch_box_recoveries(xnb_batch* batch, int slot, int64_t* n);
/* Steps replayed from a captured HIP graph and graphs captured so far.  A step with no box
 * starting or rebuilt on displacement is one launch sequence (rebuilds at fixed intervals,
 * ghost refresh, force, clock); quiet steps and rebuild steps each replay their own graph once
 * the same sequence has come twice.  XNB_BATCH_GRAPH=0 in the environment at
 * xnb_batch_create launches every step eagerly. */
int xnb_batch_graph_stats(xnb_batch* batch, int64_t* graph_steps, int64_t* captures);
/* Autotuner, before the first step of every attached box: the LDS a brick workgroup may take
 * (as a fraction of the CU's: 1/3, 1/2, 3/4 or all of it) decides each box's brick shape
 * (cells and lanes per workgroup) and how many of its workgroups share a CU.  Every box is
 * laid out under each candidate and the batched neighbour build and force are timed; the
 * fraction with the least force time plus a tenth of the build time is kept (*lds_frac) for
 * the boxes and for later attaches and recoveries.  Results do not depend on it.  Default
 * without a tuning: 1/2 (XNB_BATCH_LDS). */
int xnb_batch_autotune(xnb_batch* batch, double* lds_frac);
/* Detach the box of a slot; its xnb_ctx then holds the state xnb_run would have produced
 * (for a finished box) and p->n_rebuilds / n_violations are filled as by xnb_run.  A box
 * detached before its last step holds part of its next step: its lists and backup are marked
 * invalid, so xnb_run and the force operators return XNB_E_STATE until xnb_init runs again. */
int xnb_batch_detach(xnb_batch* batch, int slot, xnb_run_params* p);

/* ---- serving loop ---------------------------------------------------------------- */
/* No counterpart in the reference.  Continuous batching of MD jobs: jobs become visible at
 * their arrival time, are admitted in arrival order while the estimated device bytes of the
 * admitted jobs stay within the budget (and at most max_active run), are set up on a worker
 * thread (context, binning, first forces on the job's own stream) while the batch keeps
 * stepping, join the one xnb_batch when ready, and leave when their steps are done.
 * max_active = 1 serves one job at a time on the same path. */
typedef struct {
  double arrival_s;                    /* seconds after the loop starts */
  size_t n;                            /* atoms */
  const double* r;                     /* n x 3 */
  const double* v;                     /* n x 3, or NULL */
  const uint64_t* id;                  /* n, or NULL (0..n-1) */
  xnb_domain domain;
  xnb_run_params run;                  /* dt, lj, rcut, rcut_inc, precision, rebuild_every, steps */
  double* r_out;                       /* optional n x 3: final positions, cell order */
  double* v_out;                       /* optional n x 3 */
  uint64_t* id_out;                    /* optional n */
} xnb_job;

typedef struct {
  double start_s, finish_s, latency_s; /* admission, completion, completion - arrival */
  int status;                          /* XNB_OK or the error of the job */
  int64_t n_rebuilds, n_violations;
  size_t peak_device_bytes;            /* most device memory the job's context held at once */
} xnb_job_result;

typedef struct {
  int device;
  size_t hbm_budget_bytes;             /* 0: no limit */
  int max_active;                      /* <= 0: no limit */
  int steps_per_iter;                  /* batch steps between admission checks (>= 1) */
  int autotune;                        /* 1: xnb_autotune each admitted job (cached per class) */
} xnb_serve_config;

typedef struct {
  double wall_s, atom_steps, atom_steps_per_s;
  double p50_latency_s, p99_latency_s, max_latency_s, mean_active;
  int completed, failed, peak_active;
  size_t peak_bytes;                   /* most budgeted bytes (xnb_job_bytes_for) of admitted jobs at once */
  size_t peak_device_bytes;            /* most device bytes held by the contexts in the batch, summed
                                        * (sampled after each batch step) */
} xnb_serve_stats;

/* Device bytes the loop budgets for a job: an upper bound of what the job's context
 * allocates while it runs in the loop with the default batch margins (checked by the tests
 * against the contexts' counted allocations, xnb_job_result.peak_device_bytes).  It grows
 * with the atom and periodic-ghost slots, with the neighbours an atom has within
 * rcut + rcut_inc at the job's mean density, and with the cells. */
size_t xnb_job_bytes_for(const xnb_job* job);
/* The same for n atoms with ghosts at most 0.4 n, at most 110 neighbours each within
 * rcut + rcut_inc (density 1.2 at rcut + rcut_inc = 2.8) and at most n cells. */
size_t xnb_job_bytes(size_t n);
int xnb_serve_run(const xnb_job* jobs, int njobs, const xnb_serve_config* cfg, xnb_job_result* results,
                  xnb_serve_stats* stats);

/* ---- operator decks ---------------------------------------------------------- */
/* YAML-subset decks with the reference's operator names and slots (the .msp structure of
 * data/config/numerical-scheme.msp, update-particles.msp, compute-loop.msp): named
 * definitions, batches with condition / loop / rebind, "+name" appends, a `global` slot
 * block.  Each registered operator forwards to the entry point of this header that
 * replaces it.  Loading and slot queries make no GPU call. */
typedef struct xnb_deck xnb_deck;
int xnb_deck_create(xnb_deck** deck);
int xnb_deck_destroy(xnb_deck* deck);
int xnb_deck_load(xnb_deck* deck, const char* yaml_text); /* merge a deck text */
/* particles consumed by the deck's `particle_input` operator (copied) */
int xnb_deck_set_particles(xnb_deck* deck, size_t n, const double* r, const double* v,
                           const uint64_t* id, const uint8_t* type);
int xnb_deck_run(xnb_deck* deck, const char* entry); /* entry NULL: "simulation" */
xnb_ctx* xnb_deck_context(xnb_deck* deck);          /* the box built by the `domain` operator */
const char* xnb_deck_error(const xnb_deck* deck);
const char* xnb_deck_log(const xnb_deck* deck);
int xnb_deck_slot(xnb_deck* deck, const char* name, char* out, size_t cap);

/* ---- input generation (host only, no GPU call) ------------------------------ */
/* FCC lattice of ncell^3 unit cells at reduced density rho, first n_atoms sites in lattice
 * order (n_atoms <= 4 ncell^3; pass 4 ncell^3 for a full lattice), each coordinate jittered
 * uniformly in [-jitter, jitter], Maxwell-Boltzmann velocities at temperature T (mass 1,
 * zero net momentum, exact rescale with 3N-3 degrees of freedom), from a counter-based
 * SplitMix64 stream keyed by seed.  Box edge L = ncell * a, a = (4/rho)^(1/3). */
int xnb_generate_fcc(int64_t ncell, double rho, int64_t n_atoms, double jitter, double temperature,
                     uint64_t seed, double* r, double* v, uint64_t* id, double* box_length);

#ifdef __cplusplus
}
#endif
#endif
