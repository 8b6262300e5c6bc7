/*
 * xnb_oracle.h -- CPU FP64 restatement of exaNBody's short-range Lennard-Jones
 * hot path (cell grid with periodic ghost layers, chunk_neighbors CS=1 stream,
 * compute_cell_particle_pairs + LJ functor, velocity-Verlet push operators,
 * move_particles re-binning, backup_r / particle_displ_over trigger).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (exanbody_amd, libxnb_amd.so) never links or calls it.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef XNB_ORACLE_H
#define XNB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct op_sim op_sim;

/* error codes (negative) */
#define OP_OK 0
#define OP_E_ARG (-1)
#define OP_E_DOMAIN (-2)
#define OP_E_OUTSIDE (-3)
#define OP_E_STATE (-4)
#define OP_E_OVERFLOW (-5)

/* sub_grid_density of rebuild_amr in the reference's update deck (update-particles.msp:8-11) */
#define OP_SUB_GRID_DENSITY 6.5

/* chunk_neighbors config, subset of ChunkNeighborsConfig
 * (src/particle_neighbors/include/exanb/particle_neighbors/chunk_neighbors_config.h:27-39) */
typedef struct {
  int build_particle_offset;
  int dual_particle_offset;
  int random_access;
  int half_symmetric;
  int skip_ghosts;
} op_nbh_config;

/* Domain + grid construction.  Mirrors Domain (src/core/include/exanb/core/domain.h:36-129),
 * check_domain (src/core/lib/domain.cpp:78-85) and init_rcb_grid on one rank
 * (src/grid_cell_particles/init_rcb_grid.cpp:64-77) followed by the ghost-layer
 * resize done by migrate_cell_particles (src/mpi/include/exanb/mpi/migrate_cell_particles.h:630-635):
 * ghost_layers = ceil(ghost_dist / cell_size) (src/core/include/exanb/core/grid.h:110). */
int op_create(op_sim** out, const double origin[3], const double extent[3],
              const int64_t grid_dims[3], double cell_size, const int periodic[3],
              double ghost_dist);
void op_destroy(op_sim* s);

int op_ghost_layers(const op_sim* s);
void op_grid_dims(const op_sim* s, int64_t dims[3]);
size_t op_num_cells(const op_sim* s);
size_t op_num_particles(const op_sim* s);        /* inner particles */
size_t op_num_ghost_particles(const op_sim* s);
uint32_t op_cell_count(const op_sim* s, size_t cell);

/* Insert particles in input order, each appended to the cell given by
 * domain_periodic_location (domain.h:160-191) after periodic wrap.
 * r, v: n x 3 interleaved.  type may be NULL (all 0). */
int op_set_particles(op_sim* s, size_t n, const double* r, const double* v,
                     const uint64_t* id, const uint8_t* type);

/* Export particles in cell order.  include_ghosts selects ghost cells too.
 * Any output pointer may be NULL.  Returns number written. */
size_t op_export(const op_sim* s, int include_ghosts, uint32_t* cell, double* r,
                 double* v, double* f, double* ep, uint64_t* id, uint8_t* type);

/* Operators, by reference YAML name. */
int op_move_particles(op_sim* s);                       /* move_particles */
int op_ghost_update_all(op_sim* s);                     /* ghost_comm_scheme + ghost_update_all */
int op_ghost_update_r(op_sim* s);                       /* ghost_update_r */
int op_backup_r(op_sim* s);                             /* backup_r */
int op_rebuild_amr(op_sim* s, double sub_grid_density); /* rebuild_amr (enforced_ordering 1) */
int op_sub_grid_size(uint32_t n, double density);       /* sub_grid_size of a cell of n particles */
int64_t op_particle_displ_over(op_sim* s, double threshold); /* particle_displ_over: count */
int op_chunk_neighbors(op_sim* s, double nbh_dist_lab, const op_nbh_config* cfg); /* chunk_neighbors */
int op_zero_force(op_sim* s, int ghost);                /* zero_particle_force */
int op_lj_force(op_sim* s, double epsilon, double sigma, double rcut, int ghost); /* lennard_jones_force */
int op_push_f_v_r(op_sim* s, double dt, double dt_scale); /* push_f_v_r */
int op_push_f_v(op_sim* s, double dt, double dt_scale);   /* push_f_v */

/* Multi-material LJ (compute_pair_multimat.h:100-316 restated): ntypes types,
 * per unique pair (unique_pair_id, particle_type_pair.h:39-55) epsilon, sigma, rcut. */
/* gravitational_force (contribs/pi/gravitational_force.cu:163-225): buffer-less functor
 * (:56-82), masses per type; pairs added into the stored force in stream order. */
int op_gravitational_force(op_sim* s, double G, int ntypes, const double* type_mass, double rcut, int ghost);
int op_lj_force_multimat(op_sim* s, int ntypes, const double* eps_pair,
                         const double* sigma_pair, const double* rcut_pair);

/* Neighbour stream export: per-cell uint16 streams. */
size_t op_stream_size(const op_sim* s, size_t cell);
const uint16_t* op_stream_data(const op_sim* s, size_t cell);
uint32_t op_max_neighbors(const op_sim* s);

/* Composite: first_iteration + numerical_scheme loop (data/config/compute-loop.msp:1-32,
 * numerical-scheme.msp:8-25, update-particles.msp:42-79).
 * rebuild_every <= 0: displacement trigger (reference); > 0: fixed interval, with the
 * displacement check still evaluated and violations counted. */
typedef struct {
  double dt, epsilon, sigma, rcut, rcut_inc;
  int rebuild_every;
  int64_t steps;
  /* outputs */
  int64_t n_rebuilds;
  int64_t n_violations;
} op_run_params;
int op_init(op_sim* s, double epsilon, double sigma, double rcut, double rcut_inc); /* init_particles + first forces */
int op_run(op_sim* s, op_run_params* p);                    /* steps of numerical_scheme */

/* Kinetic (mass 1) and potential energy, PE = 1/2 sum_i e_i (needs op_lj_force first). */
void op_energies(const op_sim* s, double* ke, double* pe);

/* backup_r u32 encoding, exposed for closed-form tests (backup_r.h:31-51) */
uint32_t op_encode_double_u32(double x, double o, double r);
double op_restore_u32_double(uint32_t x, double o, double r);

int op_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
