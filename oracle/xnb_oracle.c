/*
 * xnb_oracle.c -- CPU FP64 restatement of exaNBody's short-range LJ hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  Never part of the product path.
 *
 * Data model: Grid<FieldSet> with per-cell SoA particle storage
 * (src/core/include/exanb/core/grid.h:57-135), k-major cell index
 * ((k*dj)+j)*di+i (src/core/include/exanb/core/grid_algorithm.h:84-91),
 * ghost layers ceil(ghost_dist/cell_size) (grid.h:110), grid offset -gl.
 *
 * Floating point: every expression is written in the evaluation order of the
 * reference source and the file is compiled with -ffp-contract=off, so no FMA
 * contraction happens (the reference's default x86-64 build has no FMA).
 * norm2(v) is taken as (x*x + y*y) + z*z (onika's Vec3d, un-vendored).
 *
 * Canonical particle order.  The reference's move_particles appends movers under
 * "omp critical" in thread-timing order (move_particles_across_cells.h:144-148), so
 * its within-cell order is only deterministic with one OpenMP thread.  This oracle
 * fixes the order to what one thread produces: movers are collected in cell-index
 * order, then particle order, and each cell is packed by the reference's
 * swap-with-last scheme (move_particles_across_cells.h:186-214).
 */
#include "xnb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ cells */

typedef struct {
  uint32_t n, cap;
  double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz, *ep;
  uint64_t* id;
  uint8_t* type;
} op_cell;

typedef struct {
  double r[3], v[3], f[3], ep;
  uint64_t id;
  uint8_t type;
} op_tuple;

typedef struct {
  /* ghost cell receive list: source inner cell + particle indices + shift flags
   * (GhostCellSendScheme, src/mpi/include/exanb/mpi/ghosts_comm_scheme.h:116-125) */
  int64_t src_cell;
  int shift[3];
  uint32_t n;
  uint32_t* pidx;
} op_ghost;

struct op_sim {
  /* domain */
  double origin[3], extent[3], cell_size;
  int64_t dom_dims[3];
  int periodic[3];
  /* grid */
  int gl;
  int64_t dims[3];
  int64_t offset[3];
  double max_nbh_dist; /* grid.max_neighbor_distance() == ghost_dist */
  size_t ncells;
  op_cell* cells;
  /* ghosts */
  op_ghost* ghosts; /* indexed by cell (only ghost cells used) */
  int ghosts_valid;
  /* backup_r */
  uint32_t** backup;
  /* neighbour streams */
  uint16_t** stream;
  size_t* stream_len;
  uint32_t max_neighbors;
  int streams_valid;
  /* run state */
  double nbh_dist_lab, max_displ, ghost_dist;
};

static int cell_reserve(op_cell* c, uint32_t cap) {
  if (cap <= c->cap) return 0;
  uint32_t nc = c->cap ? c->cap : 8;
  while (nc < cap) nc *= 2;
#define RE(p, T)                                          \
  do {                                                    \
    void* q = realloc(c->p, (size_t)nc * sizeof(T));      \
    if (!q) return OP_E_ARG;                              \
    c->p = (T*)q;                                         \
  } while (0)
  RE(rx, double); RE(ry, double); RE(rz, double);
  RE(vx, double); RE(vy, double); RE(vz, double);
  RE(fx, double); RE(fy, double); RE(fz, double);
  RE(ep, double); RE(id, uint64_t); RE(type, uint8_t);
#undef RE
  c->cap = nc;
  return 0;
}

static void cell_free(op_cell* c) {
  free(c->rx); free(c->ry); free(c->rz); free(c->vx); free(c->vy); free(c->vz);
  free(c->fx); free(c->fy); free(c->fz); free(c->ep); free(c->id); free(c->type);
  memset(c, 0, sizeof(*c));
}

static void cell_get(const op_cell* c, uint32_t j, op_tuple* t) {
  t->r[0] = c->rx[j]; t->r[1] = c->ry[j]; t->r[2] = c->rz[j];
  t->v[0] = c->vx[j]; t->v[1] = c->vy[j]; t->v[2] = c->vz[j];
  t->f[0] = c->fx[j]; t->f[1] = c->fy[j]; t->f[2] = c->fz[j];
  t->ep = c->ep[j]; t->id = c->id[j]; t->type = c->type[j];
}

static void cell_put(op_cell* c, uint32_t j, const op_tuple* t) {
  c->rx[j] = t->r[0]; c->ry[j] = t->r[1]; c->rz[j] = t->r[2];
  c->vx[j] = t->v[0]; c->vy[j] = t->v[1]; c->vz[j] = t->v[2];
  c->fx[j] = t->f[0]; c->fy[j] = t->f[1]; c->fz[j] = t->f[2];
  c->ep[j] = t->ep; c->id[j] = t->id; c->type[j] = t->type;
}

static int cell_push(op_cell* c, const op_tuple* t) {
  if (cell_reserve(c, c->n + 1)) return OP_E_ARG;
  cell_put(c, c->n, t);
  c->n++;
  return 0;
}

/* ------------------------------------------------------------- geometry */

/* grid_ijk_to_index (grid_algorithm.h:84-91) */
static inline int64_t ijk_index(const int64_t d[3], int64_t i, int64_t j, int64_t k) {
  return (k * d[1] + j) * d[0] + i;
}

static inline void index_ijk(const int64_t d[3], int64_t idx, int64_t* i, int64_t* j, int64_t* k) {
  *i = idx % d[0];
  idx /= d[0];
  *j = idx % d[1];
  *k = idx / d[1];
}

/* details::pmod (domain.h:140-146) */
static inline int64_t pmod(int64_t i, int64_t n) { return ((i % n) + n) % n; }

/* domain_periodic_location (domain.h:160-191): floor((r-o)/cs), wrap on periodic axes,
 * r += (new-old)*cs.  Returns domain cell location. */
static void domain_periodic_location(const op_sim* s, double r[3], int64_t loc[3]) {
  for (int a = 0; a < 3; a++) {
    double rel = r[a] - s->origin[a];
    loc[a] = (int64_t)floor(rel / s->cell_size);
  }
  for (int a = 0; a < 3; a++) {
    if ((loc[a] < 0 || loc[a] >= s->dom_dims[a]) && s->periodic[a]) {
      int64_t old = loc[a];
      loc[a] = pmod(loc[a], s->dom_dims[a]);
      r[a] += (double)(loc[a] - old) * s->cell_size;
    }
  }
}

/* Grid::cell_position (grid.h:113-117): origin + (offset+loc)*cell_size */
static inline double cell_pos(const op_sim* s, int a, int64_t loc) {
  return s->origin[a] + (double)(s->offset[a] + loc) * s->cell_size;
}

/* min_distance2_between(Vec3d, AABB) (geometry.h:140-155) */
static double min_dist2_point_box(const double p[3], const double bmin[3], const double bmax[3]) {
  double d[3];
  for (int a = 0; a < 3; a++) {
    d[a] = 0.;
    if (p[a] < bmin[a]) d[a] = bmin[a] - p[a];
    else if (p[a] > bmax[a]) d[a] = p[a] - bmax[a];
  }
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

static inline int is_ghost_loc(const op_sim* s, int64_t i, int64_t j, int64_t k) {
  int64_t g = s->gl;
  return i < g || j < g || k < g || i >= s->dims[0] - g || j >= s->dims[1] - g || k >= s->dims[2] - g;
}

static inline int is_ghost_cell(const op_sim* s, size_t c) {
  int64_t i, j, k;
  index_ijk(s->dims, (int64_t)c, &i, &j, &k);
  return is_ghost_loc(s, i, j, k);
}

/* --------------------------------------------------------------- create */

int op_create(op_sim** out, const double origin[3], const double extent[3],
              const int64_t grid_dims[3], double cell_size, const int periodic[3],
              double ghost_dist) {
  if (!out || !(cell_size > 0.) || !(ghost_dist >= 0.)) return OP_E_ARG;
  for (int a = 0; a < 3; a++) {
    if (grid_dims[a] <= 0) return OP_E_ARG;
    /* check_domain (domain.cpp:78-85) */
    double gsz = (double)grid_dims[a] * cell_size;
    double dsz = extent[a] - origin[a];
    if (!(fabs(1. - gsz / dsz) < 1.e-15)) return OP_E_DOMAIN;
  }
  op_sim* s = (op_sim*)calloc(1, sizeof(op_sim));
  if (!s) return OP_E_ARG;
  for (int a = 0; a < 3; a++) {
    s->origin[a] = origin[a];
    s->extent[a] = extent[a];
    s->dom_dims[a] = grid_dims[a];
    s->periodic[a] = periodic[a] ? 1 : 0;
  }
  s->cell_size = cell_size;
  s->max_nbh_dist = ghost_dist;
  s->gl = (int)ceil(ghost_dist / cell_size);
  for (int a = 0; a < 3; a++) {
    /* one periodic image shift per side only (update_ghosts_comm_scheme.cpp:125-127) */
    if (s->periodic[a] && s->gl > grid_dims[a]) { free(s); return OP_E_DOMAIN; }
    s->dims[a] = grid_dims[a] + 2 * s->gl;
    s->offset[a] = -s->gl;
  }
  s->ncells = (size_t)(s->dims[0] * s->dims[1] * s->dims[2]);
  s->cells = (op_cell*)calloc(s->ncells, sizeof(op_cell));
  s->ghosts = (op_ghost*)calloc(s->ncells, sizeof(op_ghost));
  s->backup = (uint32_t**)calloc(s->ncells, sizeof(uint32_t*));
  s->stream = (uint16_t**)calloc(s->ncells, sizeof(uint16_t*));
  s->stream_len = (size_t*)calloc(s->ncells, sizeof(size_t));
  if (!s->cells || !s->ghosts || !s->backup || !s->stream || !s->stream_len) {
    op_destroy(s);
    return OP_E_ARG;
  }
  *out = s;
  return OP_OK;
}

static void free_streams(op_sim* s) {
  for (size_t c = 0; c < s->ncells; c++) {
    free(s->stream[c]);
    s->stream[c] = NULL;
    s->stream_len[c] = 0;
  }
  s->streams_valid = 0;
}

void op_destroy(op_sim* s) {
  if (!s) return;
  for (size_t c = 0; c < s->ncells; c++) {
    if (s->cells) cell_free(&s->cells[c]);
    if (s->ghosts) free(s->ghosts[c].pidx);
    if (s->backup) free(s->backup[c]);
  }
  if (s->stream) free_streams(s);
  free(s->cells); free(s->ghosts); free(s->backup); free(s->stream); free(s->stream_len);
  free(s);
}

int op_ghost_layers(const op_sim* s) { return s->gl; }
void op_grid_dims(const op_sim* s, int64_t d[3]) { d[0] = s->dims[0]; d[1] = s->dims[1]; d[2] = s->dims[2]; }
size_t op_num_cells(const op_sim* s) { return s->ncells; }
uint32_t op_cell_count(const op_sim* s, size_t c) { return c < s->ncells ? s->cells[c].n : 0; }

size_t op_num_particles(const op_sim* s) {
  size_t n = 0;
  for (size_t c = 0; c < s->ncells; c++) if (!is_ghost_cell(s, c)) n += s->cells[c].n;
  return n;
}

size_t op_num_ghost_particles(const op_sim* s) {
  size_t n = 0;
  for (size_t c = 0; c < s->ncells; c++) if (is_ghost_cell(s, c)) n += s->cells[c].n;
  return n;
}

/* --------------------------------------------------------- particles io */

int op_set_particles(op_sim* s, size_t n, const double* r, const double* v,
                     const uint64_t* id, const uint8_t* type) {
  for (size_t c = 0; c < s->ncells; c++) s->cells[c].n = 0;
  s->ghosts_valid = 0;
  free_streams(s);
  for (size_t p = 0; p < n; p++) {
    op_tuple t;
    memset(&t, 0, sizeof(t));
    for (int a = 0; a < 3; a++) {
      t.r[a] = r[3 * p + a];
      t.v[a] = v ? v[3 * p + a] : 0.;
    }
    t.id = id ? id[p] : (uint64_t)p;
    t.type = type ? type[p] : 0;
    int64_t loc[3];
    domain_periodic_location(s, t.r, loc);
    for (int a = 0; a < 3; a++)
      if (loc[a] < 0 || loc[a] >= s->dom_dims[a]) return OP_E_OUTSIDE;
    int64_t c = ijk_index(s->dims, loc[0] + s->gl, loc[1] + s->gl, loc[2] + s->gl);
    if (cell_push(&s->cells[c], &t)) return OP_E_ARG;
  }
  return OP_OK;
}

size_t op_export(const op_sim* s, int include_ghosts, uint32_t* cell, double* r,
                 double* v, double* f, double* ep, uint64_t* id, uint8_t* type) {
  size_t k = 0;
  for (size_t c = 0; c < s->ncells; c++) {
    if (!include_ghosts && is_ghost_cell(s, c)) continue;
    const op_cell* C = &s->cells[c];
    for (uint32_t j = 0; j < C->n; j++, k++) {
      if (cell) cell[k] = (uint32_t)c;
      if (r) { r[3 * k] = C->rx[j]; r[3 * k + 1] = C->ry[j]; r[3 * k + 2] = C->rz[j]; }
      if (v) { v[3 * k] = C->vx[j]; v[3 * k + 1] = C->vy[j]; v[3 * k + 2] = C->vz[j]; }
      if (f) { f[3 * k] = C->fx[j]; f[3 * k + 1] = C->fy[j]; f[3 * k + 2] = C->fz[j]; }
      if (ep) ep[k] = C->ep[j];
      if (id) id[k] = C->id[j];
      if (type) type[k] = C->type[j];
    }
  }
  return k;
}

/* ------------------------------------------------------- move_particles */

typedef struct {
  op_tuple t;
  int64_t dst;
} op_mover;

/* move_particles_across_cells (move_particles_across_cells.h:78-235), single-thread order. */
int op_move_particles(op_sim* s) {
  const size_t nc = s->ncells;
  uint32_t** removed = (uint32_t**)calloc(nc, sizeof(uint32_t*));
  uint32_t* n_removed = (uint32_t*)calloc(nc, sizeof(uint32_t));
  uint32_t* cap_removed = (uint32_t*)calloc(nc, sizeof(uint32_t));
  op_mover** incoming = (op_mover**)calloc(nc, sizeof(op_mover*));
  uint32_t* n_in = (uint32_t*)calloc(nc, sizeof(uint32_t));
  uint32_t* cap_in = (uint32_t*)calloc(nc, sizeof(uint32_t));
  int err = OP_OK;
  if (!removed || !n_removed || !cap_removed || !incoming || !n_in || !cap_in) { err = OP_E_ARG; goto done; }

  const double eps = s->cell_size * (1.0 / (double)(1ull << 48)); /* grid.h:60,89-90 */
  const double eps2 = eps * eps;
  /* first pass: GRID_OMP_FOR over inner cells in index order (parallel_grid_algorithm.h:29-48) */
  for (int64_t k = 0; k < s->dom_dims[2]; k++)
    for (int64_t j = 0; j < s->dom_dims[1]; j++)
      for (int64_t i = 0; i < s->dom_dims[0]; i++) {
        int64_t src_loc[3] = {i + s->gl, j + s->gl, k + s->gl};
        int64_t ci = ijk_index(s->dims, src_loc[0], src_loc[1], src_loc[2]);
        op_cell* C = &s->cells[ci];
        for (uint32_t p = 0; p < C->n; p++) {
          double r[3] = {C->rx[p], C->ry[p], C->rz[p]};
          double ro[3] = {r[0], r[1], r[2]};
          int64_t dl[3];
          domain_periodic_location(s, r, dl);
          int64_t dst_loc[3];
          for (int a = 0; a < 3; a++) dst_loc[a] = dl[a] - s->offset[a];
          int moved = (dst_loc[0] != src_loc[0] || dst_loc[1] != src_loc[1] || dst_loc[2] != src_loc[2]);
          if (!moved) {
            double bmin[3], bmax[3];
            for (int a = 0; a < 3; a++) {
              bmin[a] = cell_pos(s, a, dst_loc[a]);
              bmax[a] = cell_pos(s, a, dst_loc[a] + 1);
            }
            moved = min_dist2_point_box(ro, bmin, bmax) >= eps2;
          }
          if (!moved) continue;
          int inside = 1;
          for (int a = 0; a < 3; a++)
            if (dst_loc[a] < s->gl || dst_loc[a] >= s->dims[a] - s->gl) inside = 0;
          if (!inside) { err = OP_E_OUTSIDE; goto done; } /* would go to otb_particles */
          int64_t cj = ijk_index(s->dims, dst_loc[0], dst_loc[1], dst_loc[2]);
          if (n_in[cj] == cap_in[cj]) {
            cap_in[cj] = cap_in[cj] ? 2 * cap_in[cj] : 4;
            incoming[cj] = (op_mover*)realloc(incoming[cj], cap_in[cj] * sizeof(op_mover));
          }
          op_mover* m = &incoming[cj][n_in[cj]++];
          cell_get(C, p, &m->t);
          m->t.r[0] = r[0]; m->t.r[1] = r[1]; m->t.r[2] = r[2];
          m->dst = cj;
          if (n_removed[ci] == cap_removed[ci]) {
            cap_removed[ci] = cap_removed[ci] ? 2 * cap_removed[ci] : 4;
            removed[ci] = (uint32_t*)realloc(removed[ci], cap_removed[ci] * sizeof(uint32_t));
          }
          removed[ci][n_removed[ci]++] = p;
        }
      }

  /* second pass: pack (swap-with-last) then append incoming (:186-214) */
#pragma omp parallel for schedule(dynamic)
  for (size_t ci = 0; ci < nc; ci++) {
    op_cell* C = &s->cells[ci];
    int32_t n = (int32_t)C->n;
    int32_t n_out = (int32_t)n_removed[ci];
    if (n_out > 0) {
      int32_t* pack = (int32_t*)malloc((size_t)n * sizeof(int32_t));
      for (int32_t j = 0; j < n; j++) pack[j] = j;
      for (int32_t j = 0; j < n_out; j++) pack[removed[ci][j]] = -1;
#ifdef OP_MUTANT_STABLE_PACK
      /* planted mistake for the order-fixture tests: stable compaction of the stayers */
      {
        int32_t w = 0;
        for (int32_t j = 0; j < n; j++)
          if (pack[j] != -1) pack[w++] = pack[j];
        n = w;
      }
#else
      for (int32_t j = n - 1; j >= 0; j--) {
        if (pack[j] == -1) {
          if (j < n - 1) pack[j] = pack[n - 1];
          --n;
        }
      }
#endif
      for (int32_t j = 0; j < n; j++) {
        if (j != pack[j]) {
          op_tuple t;
          cell_get(C, (uint32_t)pack[j], &t);
          cell_put(C, (uint32_t)j, &t);
        }
      }
      C->n = (uint32_t)n;
      free(pack);
    }
    for (uint32_t j = 0; j < n_in[ci]; j++) cell_push(C, &incoming[ci][j].t);
  }
  s->ghosts_valid = 0;
  free_streams(s);

done:
  if (removed) for (size_t c = 0; c < nc; c++) free(removed[c]);
  if (incoming) for (size_t c = 0; c < nc; c++) free(incoming[c]);
  free(removed); free(n_removed); free(cap_removed); free(incoming); free(n_in); free(cap_in);
  return err;
}

/* -------------------------------------------------------------- ghosts */

/* GhostBoundaryModifier::apply_coord_modifier for a periodic SHIFT
 * (ghosts_comm_scheme.h:76-81): x + (side ? 1.0 : -1.0) * (rmax - rmin) */
static inline double shift_coord(double x, int shift, double rmin, double rmax) {
  if (shift == 0) return x;
  return x + (shift > 0 ? 1.0 : -1.0) * (rmax - rmin);
}

/* ghost_comm_scheme on one rank (update_ghosts_comm_scheme.cpp:84-337, partner_comm_scheme
 * :400-470) followed by ghost_update_all (update_ghosts.cu:45-64, grid_update_ghosts.h:63-202).
 * Each ghost cell g at domain location gd receives from the inner cell gd - shift*dims,
 * shift in {-1,0,1}^3; kept particles are the source particles, in order, whose image lies
 * inside enlarge(block_to_bounds(domain block), max_nbh_dist) (geometry.h:184-228, inclusive). */
int op_ghost_update_all(op_sim* s) {
  double bmin[3], bmax[3];
  for (int a = 0; a < 3; a++) {
    bmin[a] = ((double)0 * s->cell_size + s->origin[a]) - s->max_nbh_dist;
    bmax[a] = ((double)s->dom_dims[a] * s->cell_size + s->origin[a]) + s->max_nbh_dist;
  }
  int err = OP_OK;
#pragma omp parallel for schedule(dynamic)
  for (size_t c = 0; c < s->ncells; c++) {
    int64_t i, j, k;
    index_ijk(s->dims, (int64_t)c, &i, &j, &k);
    if (!is_ghost_loc(s, i, j, k)) continue;
    op_ghost* G = &s->ghosts[c];
    op_cell* C = &s->cells[c];
    C->n = 0;
    G->n = 0;
    G->src_cell = -1;
    int64_t gd[3] = {i + s->offset[0], j + s->offset[1], k + s->offset[2]};
    int64_t src[3];
    int ok = 1;
    for (int a = 0; a < 3; a++) {
      int sh = gd[a] < 0 ? -1 : (gd[a] >= s->dom_dims[a] ? 1 : 0);
      if (sh != 0 && !s->periodic[a]) ok = 0;
      /* shifted_cell_loc = shift*dims + domain_cell_loc  =>  source = gd - shift*dims */
      G->shift[a] = sh;
      src[a] = gd[a] - (int64_t)sh * s->dom_dims[a];
    }
    if (!ok) continue;
    int64_t sc = ijk_index(s->dims, src[0] - s->offset[0], src[1] - s->offset[1], src[2] - s->offset[2]);
    G->src_cell = sc;
    const op_cell* S = &s->cells[sc];
    free(G->pidx);
    G->pidx = (uint32_t*)malloc((S->n ? S->n : 1) * sizeof(uint32_t));
    if (cell_reserve(C, S->n)) { err = OP_E_ARG; continue; }
#ifdef OP_MUTANT_REVERSE_GHOSTS
    /* planted mistake for the order-fixture tests: images in reversed source order */
    for (uint32_t pr = S->n; pr-- > 0;) {
      const uint32_t p = pr;
#else
    for (uint32_t p = 0; p < S->n; p++) {
#endif
      double gr[3] = {shift_coord(S->rx[p], G->shift[0], s->origin[0], s->extent[0]),
                      shift_coord(S->ry[p], G->shift[1], s->origin[1], s->extent[1]),
                      shift_coord(S->rz[p], G->shift[2], s->origin[2], s->extent[2])};
      if (gr[0] >= bmin[0] && gr[0] <= bmax[0] && gr[1] >= bmin[1] && gr[1] <= bmax[1] &&
          gr[2] >= bmin[2] && gr[2] <= bmax[2]) {
        op_tuple t;
        cell_get(S, p, &t);
        t.r[0] = gr[0]; t.r[1] = gr[1]; t.r[2] = gr[2];
        cell_put(C, C->n, &t);
        C->n++;
        G->pidx[G->n++] = p;
      }
    }
  }
  s->ghosts_valid = (err == OP_OK);
  return err;
}

/* ghost_update_r (update_ghosts.cu:58): re-copy positions along the stored scheme */
int op_ghost_update_r(op_sim* s) {
  if (!s->ghosts_valid) return OP_E_STATE;
#pragma omp parallel for schedule(dynamic)
  for (size_t c = 0; c < s->ncells; c++) {
    if (!is_ghost_cell(s, c)) continue;
    const op_ghost* G = &s->ghosts[c];
    if (G->src_cell < 0) continue;
    const op_cell* S = &s->cells[G->src_cell];
    op_cell* C = &s->cells[c];
    for (uint32_t q = 0; q < G->n; q++) {
      uint32_t p = G->pidx[q];
      C->rx[q] = shift_coord(S->rx[p], G->shift[0], s->origin[0], s->extent[0]);
      C->ry[q] = shift_coord(S->ry[p], G->shift[1], s->origin[1], s->extent[1]);
      C->rz[q] = shift_coord(S->rz[p], G->shift[2], s->origin[2], s->extent[2]);
    }
  }
  return OP_OK;
}

/* ------------------------------------------------------ backup / trigger */

/* restore_u32_double / encode_double_u32 (src/core/include/exanb/core/backup_r.h:31-51) */
double op_restore_u32_double(uint32_t x, double o, double r) {
  return o + ((double)x * r) / (double)(1ull << 32);
}

uint32_t op_encode_double_u32(double x, double o, double r) {
  double xo = x - o;
  int64_t v = (int64_t)((xo * (double)(1ull << 32)) / r);
  if (v < 0) v = 0;
  if (v > (int64_t)UINT32_MAX) v = (int64_t)UINT32_MAX;
  uint32_t a = (uint32_t)v;
  uint32_t b = a - 1u;
  uint32_t c = a + 1u;
  double ea = fabs(op_restore_u32_double(a, o, r) - x);
  double eb = fabs(op_restore_u32_double(b, o, r) - x);
  double ec = fabs(op_restore_u32_double(c, o, r) - x);
  if (eb < ea) return b;
  if (ec < ea) return c;
  return a;
}

/* backup_r (src/io/backup_r.cpp:42-107): inner cells, coordinates relative to the cell origin */
int op_backup_r(op_sim* s) {
#pragma omp parallel for schedule(dynamic)
  for (size_t c = 0; c < s->ncells; c++) {
    free(s->backup[c]);
    s->backup[c] = NULL;
    int64_t i, j, k;
    index_ijk(s->dims, (int64_t)c, &i, &j, &k);
    if (is_ghost_loc(s, i, j, k)) continue;
    const op_cell* C = &s->cells[c];
    s->backup[c] = (uint32_t*)malloc((C->n ? C->n : 1) * 3 * sizeof(uint32_t));
    double o[3] = {cell_pos(s, 0, i), cell_pos(s, 1, j), cell_pos(s, 2, k)};
    for (uint32_t p = 0; p < C->n; p++) {
      s->backup[c][3 * p + 0] = op_encode_double_u32(C->rx[p], o[0], s->cell_size);
      s->backup[c][3 * p + 1] = op_encode_double_u32(C->ry[p], o[1], s->cell_size);
      s->backup[c][3 * p + 2] = op_encode_double_u32(C->rz[p], o[2], s->cell_size);
    }
  }
  return OP_OK;
}

/* particle_displ_over (src/mpi/particle_displ_over.cu:38-77,115-179): number of inner
 * particles with norm2(r - restore(backup)) >= threshold^2 */
int64_t op_particle_displ_over(op_sim* s, double threshold) {
  const double thr2 = threshold * threshold;
  int64_t count = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : count)
  for (size_t c = 0; c < s->ncells; c++) {
    int64_t i, j, k;
    index_ijk(s->dims, (int64_t)c, &i, &j, &k);
    if (is_ghost_loc(s, i, j, k)) continue;
    const op_cell* C = &s->cells[c];
    const uint32_t* rb = s->backup[c];
    if (!rb) continue;
    double o[3] = {cell_pos(s, 0, i), cell_pos(s, 1, j), cell_pos(s, 2, k)};
    for (uint32_t p = 0; p < C->n; p++) {
      double dx = C->rx[p] - op_restore_u32_double(rb[3 * p + 0], o[0], s->cell_size);
      double dy = C->ry[p] - op_restore_u32_double(rb[3 * p + 1], o[1], s->cell_size);
      double dz = C->rz[p] - op_restore_u32_double(rb[3 * p + 2], o[2], s->cell_size);
      if ((dx * dx + dy * dy) + dz * dz >= thr2) ++count;
    }
  }
  return count;
}

/* ------------------------------------------------------ chunk_neighbors */

/* encode_cell_index (chunk_neighbors.h:137-150) */
static inline uint16_t encode_cell_index(int64_t ri, int64_t rj, int64_t rk) {
  uint16_t c = (uint16_t)(rk + 16);
  c = (uint16_t)(c << 5);
  c = (uint16_t)(c + (rj + 16));
  c = (uint16_t)(c << 5);
  c = (uint16_t)(c + (ri + 16));
  return c;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

/* sub_grid_size (amr_grid_algorithm.h:66-77): cbrt(n / density), 1 below 2, at most
 * MAX_SUB_GRID_RESOLUTION = 16 (amr_grid_algorithm.h:46), no power-of-two rounding (:47) */
static int sub_grid_size(uint32_t n, double density) {
  if (n == 0) return 0;
  double side = cbrt((double)n / density);
  if (side < 2.0) return 1;
  int s = (int)floor(side);
  return s < 16 ? s : 16;
}

int op_sub_grid_size(uint32_t n, double density) { return sub_grid_size(n, density); }

/* rebuild_amr (rebuild_amr.cpp:34-69; enforced_ordering 1: one sub-resolution level, no
 * Z order) = rebuild_sub_grids + project_particles_in_sub_grids (amr_grid_algorithm.h:
 * 112-434, 377-434).  In every cell whose side sg = sub_grid_size(n, density) exceeds 1 the
 * particles are reordered by sub-cell, stably: sub-cell (i, j, k) with each index
 * clamp((ssize_t)(pcoord * sg), 0, sg - 1), pcoord = (r - cell_low) / cell_size
 * (Grid::particle_pcoord, grid.h:184-190), index (k * sg + j) * sg + i
 * (sg_cell_index, amr_grid_algorithm.h:80-93); particles of one sub-cell keep their order
 * (the counting placement of amr_grid_algorithm.h:225-233 applied by swaps, :246-265). */
int op_rebuild_amr(op_sim* s, double density) {
  if (!(density > 0.)) return OP_E_ARG;
  int err = OP_OK;
#pragma omp parallel for schedule(dynamic) reduction(| : err)
  for (size_t c = 0; c < s->ncells; c++) {
    op_cell* C = &s->cells[c];
    const uint32_t n = C->n;
    const int sg = sub_grid_size(n, density);
    if (sg <= 1) continue;
    int64_t ci, cj, ck;
    index_ijk(s->dims, (int64_t)c, &ci, &cj, &ck);
    const double lo[3] = {cell_pos(s, 0, ci), cell_pos(s, 1, cj), cell_pos(s, 2, ck)};
    const uint32_t nsub = (uint32_t)(sg * sg * sg);
    uint32_t* key = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint32_t* start = (uint32_t*)calloc((size_t)nsub + 1, sizeof(uint32_t));
    op_tuple* tmp = (op_tuple*)malloc((size_t)n * sizeof(op_tuple));
    if (!key || !start || !tmp) {
      free(key); free(start); free(tmp);
      err |= 1;
      continue;
    }
    const double* rr[3] = {C->rx, C->ry, C->rz};
    for (uint32_t p = 0; p < n; p++) {
      int64_t q[3];
      for (int a = 0; a < 3; a++) {
        const double pc = (rr[a][p] - lo[a]) / s->cell_size;
        int64_t v = (int64_t)(pc * (double)sg);
        q[a] = v < 0 ? 0 : (v > sg - 1 ? sg - 1 : v);
      }
      key[p] = (uint32_t)((q[2] * sg + q[1]) * sg + q[0]);
      start[key[p] + 1]++;
    }
    for (uint32_t t = 0; t < nsub; t++) start[t + 1] += start[t];
    for (uint32_t p = 0; p < n; p++) cell_get(C, p, &tmp[start[key[p]]++]);
    for (uint32_t p = 0; p < n; p++) cell_put(C, p, &tmp[p]);
    free(key); free(start); free(tmp);
  }
  s->ghosts_valid = 0;
  s->streams_valid = 0;
  return err ? OP_E_ARG : OP_OK;
}

/* chunk_neighbors_execute (chunk_neighbors_execute.h:40-423), CS = 1, NeighborFilterHalfSymGhost
 * (neighbor_filter_func.h:35-52).  Streams are built for every cell, ghosts included.  Every
 * particle pair of two neighbouring cells is tested: the reference's sub-cell pairs
 * (amr_grid_algorithm.cpp:102-218) only skip pairs of sub-cells farther apart than nbh_dist,
 * and each particle's entries are sorted by (cell, index) (chunk_neighbors_execute.h:305-324),
 * so the streams are the same for any sub-grid side. */
int op_chunk_neighbors(op_sim* s, double nbh_dist_lab, const op_nbh_config* cfg) {
  op_nbh_config dflt = {1, 0, 0, 0, 0};
  if (!cfg) cfg = &dflt;
  const int build_offset = cfg->build_particle_offset || cfg->dual_particle_offset;
  const double max_dist2 = nbh_dist_lab * nbh_dist_lab;
  /* cell_layers = ceil(max_dist / cell_size) (amr_grid_algorithm.h:451) */
  const int64_t lay = (int64_t)ceil(nbh_dist_lab / s->cell_size);
  if (lay > 15) return OP_E_ARG;
  free_streams(s);
  for (size_t c = 0; c < s->ncells; c++)
    if (s->cells[c].n >= 65535u) return OP_E_OVERFLOW;
  uint32_t max_nbh = 0;
  const int64_t* d = s->dims;
#pragma omp parallel reduction(max : max_nbh)
  {
    uint32_t** plist = NULL;
    uint32_t* pn = NULL;
    uint32_t* pcap = NULL;
    uint32_t palloc = 0;
    uint16_t* cc = NULL;
    size_t cc_cap = 0;
#pragma omp for schedule(dynamic)
    for (size_t ca = 0; ca < s->ncells; ca++) {
      int64_t ai, aj, ak;
      index_ijk(d, (int64_t)ca, &ai, &aj, &ak);
      const op_cell* A = &s->cells[ca];
      const uint32_t na = A->n;
      if (na > palloc) {
        plist = (uint32_t**)realloc(plist, na * sizeof(uint32_t*));
        pn = (uint32_t*)realloc(pn, na * sizeof(uint32_t));
        pcap = (uint32_t*)realloc(pcap, na * sizeof(uint32_t));
        for (uint32_t q = palloc; q < na; q++) { plist[q] = NULL; pcap[q] = 0; }
        palloc = na;
      }
      for (uint32_t q = 0; q < na; q++) pn[q] = 0;
      /* candidate cells [a - lay, a + lay] clamped to the grid (chunk_neighbors_execute.h:134-139) */
      const int64_t bi0 = ai - lay > 0 ? ai - lay : 0, bi1 = ai + lay < d[0] - 1 ? ai + lay : d[0] - 1;
      const int64_t bj0 = aj - lay > 0 ? aj - lay : 0, bj1 = aj + lay < d[1] - 1 ? aj + lay : d[1] - 1;
      const int64_t bk0 = ak - lay > 0 ? ak - lay : 0, bk1 = ak + lay < d[2] - 1 ? ak + lay : d[2] - 1;
      for (int64_t bk = bk0; bk <= bk1; bk++)
        for (int64_t bj = bj0; bj <= bj1; bj++)
          for (int64_t bi = bi0; bi <= bi1; bi++) {
            const int64_t cb = ijk_index(d, bi, bj, bk);
            const op_cell* B = &s->cells[cb];
            const uint16_t enc = encode_cell_index(bi - ai, bj - aj, bk - ak);
            const int b_ghost = is_ghost_loc(s, bi, bj, bk);
            for (uint32_t pa = 0; pa < na; pa++) {
              for (uint32_t pb = 0; pb < B->n; pb++) {
                double dx = A->rx[pa] - B->rx[pb];
                double dy = A->ry[pa] - B->ry[pb];
                double dz = A->rz[pa] - B->rz[pb];
                double d2 = (dx * dx + dy * dy) + dz * dz;
                if (((int64_t)ca != cb || pa != pb) && d2 > 0.0 && d2 <= max_dist2) {
                  if (cfg->half_symmetric && ((int64_t)ca < cb || ((int64_t)ca == cb && pa < pb))) continue;
                  if (cfg->skip_ghosts && b_ghost) continue;
                  if (pn[pa] == pcap[pa]) {
                    pcap[pa] = pcap[pa] ? 2 * pcap[pa] : 64;
                    plist[pa] = (uint32_t*)realloc(plist[pa], pcap[pa] * sizeof(uint32_t));
                  }
                  plist[pa][pn[pa]++] = ((uint32_t)enc << 16) | pb;
                }
              }
            }
          }
      /* encode (:265-398) */
      size_t table = 0;
      unsigned num_tables = 0;
      if (build_offset) {
        table = ((size_t)na + 1) * 2;
        num_tables = 1;
        if (cfg->dual_particle_offset) { table += (size_t)na * 2; num_tables = 2; }
      }
      size_t need = table;
      for (uint32_t pa = 0; pa < na; pa++) need += 1 + 3 * (size_t)pn[pa] + 2;
      if (need > cc_cap) { cc_cap = need; cc = (uint16_t*)realloc(cc, cc_cap * sizeof(uint16_t)); }
      size_t len = 0;
      if (na > 0) { memset(cc, 0, table * sizeof(uint16_t)); len = table; }
      uint32_t cell_max = 0;
      for (uint32_t pa = 0; pa < na; pa++) {
        if (build_offset) {
          uint32_t off = (uint32_t)(len - table + num_tables);
          cc[pa * 2 + 0] = (uint16_t)off;
          cc[pa * 2 + 1] = (uint16_t)(off >> 16);
        }
        qsort(plist[pa], pn[pa], sizeof(uint32_t), cmp_u32);
        uint32_t cnt = pn[pa], nodup = 1;
        for (uint32_t q = 1; q < cnt; q++)
          if (plist[pa][q] != plist[pa][nodup - 1]) plist[pa][nodup++] = plist[pa][q];
        if (nodup > cnt) nodup = cnt;
        if (nodup > cell_max) cell_max = nodup;
        size_t cell_count_idx = len;
        size_t chunk_count_idx = 0;
        uint16_t last_cell = 0;
        if (!cfg->random_access) cc[len++] = 0;
        uint32_t n_sym = 0;
        for (uint32_t q = 0; q < nodup; q++) {
          uint16_t enc = (uint16_t)(plist[pa][q] >> 16);
          uint16_t pb = (uint16_t)(plist[pa][q] & 0xFFFFu);
          if (cfg->dual_particle_offset) {
            int64_t ri = (int64_t)(enc & 31) - 16, rj = (int64_t)((enc >> 5) & 31) - 16,
                    rk = (int64_t)((enc >> 10) & 31) - 16;
            int64_t cb = (int64_t)ca + ((rk * d[1]) + rj) * d[0] + ri;
            if (cb < (int64_t)ca || (cb == (int64_t)ca && pb < pa)) ++n_sym;
          }
          if (cfg->random_access) {
            cc[len++] = enc;
            cc[len++] = pb;
          } else {
            if (enc != last_cell) {
              ++cc[cell_count_idx];
              last_cell = enc;
              cc[len++] = last_cell;
              chunk_count_idx = len;
              cc[len++] = 0;
            }
            ++cc[chunk_count_idx];
            cc[len++] = pb;
          }
        }
        if (cfg->dual_particle_offset) {
          cc[((size_t)na + 1 + pa) * 2 + 0] = (uint16_t)n_sym;
          cc[((size_t)na + 1 + pa) * 2 + 1] = (uint16_t)(n_sym >> 16);
        }
      }
      if (na > 0 && build_offset) {
        uint32_t off = (uint32_t)(len - table + num_tables);
        cc[(size_t)na * 2 + 0] = (uint16_t)off;
        cc[(size_t)na * 2 + 1] = (uint16_t)(off >> 16);
      }
      if (na > 0 && cell_max > max_nbh) max_nbh = cell_max;
      if (len > 0) {
        s->stream[ca] = (uint16_t*)malloc(len * sizeof(uint16_t));
        memcpy(s->stream[ca], cc, len * sizeof(uint16_t));
      }
      s->stream_len[ca] = len;
    }
    for (uint32_t q = 0; q < palloc; q++) free(plist[q]);
    free(plist); free(pn); free(pcap); free(cc);
  }
  s->max_neighbors = max_nbh;
  s->streams_valid = 1;
  return OP_OK;
}

size_t op_stream_size(const op_sim* s, size_t c) { return c < s->ncells ? s->stream_len[c] : 0; }
const uint16_t* op_stream_data(const op_sim* s, size_t c) { return c < s->ncells ? s->stream[c] : NULL; }
uint32_t op_max_neighbors(const op_sim* s) { return s->max_neighbors; }

/* chunknbh_stream_info (chunk_neighbors.h:172-186) */
static const uint16_t* stream_info(const uint16_t* st, uint32_t n, const uint32_t** off, int32_t* shift) {
  *off = NULL;
  *shift = 0;
  if (!st || n == 0) return NULL;
  if (st[0] <= 2 && st[1] == 0) {
    int nt = st[0];
    size_t ts = nt >= 1 ? ((size_t)n * nt + 1) * 2 : 0;
    *off = nt >= 1 ? (const uint32_t*)st : NULL;
    *shift = -nt;
    return st + ts;
  }
  return st;
}

/* ------------------------------------------------------------- forces */

/* zero_particle_force (src/compute/zero_particle_force.cu:19-55) */
int op_zero_force(op_sim* s, int ghost) {
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < s->ncells; c++) {
    if (!ghost && is_ghost_cell(s, c)) continue;
    op_cell* C = &s->cells[c];
    for (uint32_t p = 0; p < C->n; p++) { C->fx[p] = 0.; C->fy[p] = 0.; C->fz[p] = 0.; C->ep[p] = 0.; }
  }
  return OP_OK;
}

/* lj_compute_energy (contribs/md/lennard_jones/lennard_jones.cu:46-56) */
static inline void lj_compute_energy(double eps, double sigma, double r, double* e, double* de) {
  const double inv_r = 1.0 / r;
  const double ratio = sigma * inv_r;
  const double ratio2 = ratio * ratio;
  const double ratio6 = ratio2 * ratio2 * ratio2;
  const double ratio12 = ratio6 * ratio6;
  *e = 4. * eps * (ratio12 - ratio6);
  *de = (-24. * eps * (2. * ratio12 - ratio6)) * inv_r;
}

/* lennard_jones_force: compute_cell_particle_pairs2 over inner cells (all cells if ghost)
 * (compute_cell_particle_pairs.h:119-173), CPU compute-buffer path of
 * compute_cell_particle_pairs_cell (compute_cell_particle_pairs_impl_default.h:37-241):
 * per particle, neighbours decoded in stream order, accepted iff 0 < d2 <= rcut^2, the
 * buffer then reduced by LennardJonesForceFunctor (lennard_jones.cu:67-102) and added to f.
 * Per-atom energy ep = sum of pair e in the same order (the reference discards e). */
int op_lj_force(op_sim* s, double epsilon, double sigma, double rcut, int ghost) {
  if (!s->streams_valid) return OP_E_STATE;
  const double rcut2 = rcut * rcut;
  const int64_t* d = s->dims;
  int err = 0;
#pragma omp parallel reduction(| : err)
  {
    size_t cap = 256;
    double* drx = (double*)malloc(cap * sizeof(double));
    double* dry = (double*)malloc(cap * sizeof(double));
    double* drz = (double*)malloc(cap * sizeof(double));
    double* bd2 = (double*)malloc(cap * sizeof(double));
#pragma omp for schedule(dynamic)
    for (size_t ca = 0; ca < s->ncells; ca++) {
      if (!ghost && is_ghost_cell(s, ca)) continue;
      op_cell* A = &s->cells[ca];
      const uint32_t na = A->n;
      const uint32_t* off;
      int32_t shift;
      const uint16_t* base = stream_info(s->stream[ca], na, &off, &shift);
      const uint16_t* st = base;
      for (uint32_t pa = 0; pa < na; pa++) {
        if (off) st = base + off[pa] + shift;
        size_t count = 0;
        unsigned cg = *(st++);
        for (; cg > 0; cg--) {
          uint16_t enc = *(st++);
          unsigned nch = *(st++);
          int64_t ri = (int64_t)(enc & 31) - 16, rj = (int64_t)((enc >> 5) & 31) - 16,
                  rk = (int64_t)((enc >> 10) & 31) - 16;
          int64_t cb = (int64_t)ca + ((rk * d[1]) + rj) * d[0] + ri;
          const op_cell* B = &s->cells[cb];
          for (; nch > 0; nch--) {
            uint32_t pb = *(st++);
            double dx = B->rx[pb] - A->rx[pa];
            double dy = B->ry[pb] - A->ry[pa];
            double dz = B->rz[pb] - A->rz[pa];
            double d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 > 0.0 && d2 <= rcut2) {
              if (count == cap) {
                cap *= 2;
                drx = (double*)realloc(drx, cap * sizeof(double));
                dry = (double*)realloc(dry, cap * sizeof(double));
                drz = (double*)realloc(drz, cap * sizeof(double));
                bd2 = (double*)realloc(bd2, cap * sizeof(double));
              }
              drx[count] = dx; dry[count] = dy; drz[count] = dz; bd2[count] = d2;
              ++count;
            }
          }
        }
        double tl_fx = 0.0, tl_fy = 0.0, tl_fz = 0.0, tl_e = 0.0;
        for (size_t q = 0; q < count; q++) {
          const double r = sqrt(bd2[q]);
          double e = 0.0, de = 0.0;
          lj_compute_energy(epsilon, sigma, r, &e, &de);
          de *= 1.0 / r; /* interaction_weight / r with weight 1.0 */
          tl_fx += de * drx[q];
          tl_fy += de * dry[q];
          tl_fz += de * drz[q];
          tl_e += e;
        }
        A->fx[pa] += tl_fx;
        A->fy[pa] += tl_fy;
        A->fz[pa] += tl_fz;
        A->ep[pa] += tl_e;
      }
    }
    free(drx); free(dry); free(drz); free(bd2);
  }
  return err ? OP_E_ARG : OP_OK;
}

/* gravitational_force: compute_cell_particle_pairs over inner cells (all cells if ghost) with
 * the buffer-less GravitationalForceFunctor (contribs/pi/gravitational_force.cu:56-82): per
 * particle, neighbours in stream order, accepted iff 0 < d2 <= rcut^2
 * (compute_cell_particle_pairs_impl_default.h:183-186), each pair added straight into the
 * particle's force field: gravitational_compute_energy (:44-50) gives de = G m_a m_b / r^2,
 * then de *= w / r (w = 1) and f += de * dr.  A type without a mass is an error. */
int op_gravitational_force(op_sim* s, double G, int ntypes, const double* type_mass, double rcut, int ghost) {
  if (!s->streams_valid) return OP_E_STATE;
  if (ntypes < 1 || !type_mass) return OP_E_ARG;
  for (size_t c = 0; c < s->ncells; c++) {
    if (!ghost && is_ghost_cell(s, c)) continue;
    for (uint32_t p = 0; p < s->cells[c].n; p++)
      if (s->cells[c].type[p] >= ntypes) return OP_E_ARG;
  }
  const double rcut2 = rcut * rcut;
  const int64_t* d = s->dims;
#pragma omp parallel for schedule(dynamic)
  for (size_t ca = 0; ca < s->ncells; ca++) {
    if (!ghost && is_ghost_cell(s, ca)) continue;
    op_cell* A = &s->cells[ca];
    const uint32_t na = A->n;
    const uint32_t* off;
    int32_t shift;
    const uint16_t* base = stream_info(s->stream[ca], na, &off, &shift);
    const uint16_t* st = base;
    for (uint32_t pa = 0; pa < na; pa++) {
      if (off) st = base + off[pa] + shift;
      const double mass_a = type_mass[A->type[pa]];
      unsigned cg = *(st++);
      for (; cg > 0; cg--) {
        uint16_t enc = *(st++);
        unsigned nch = *(st++);
        int64_t ri = (int64_t)(enc & 31) - 16, rj = (int64_t)((enc >> 5) & 31) - 16, rk = (int64_t)((enc >> 10) & 31) - 16;
        int64_t cb = (int64_t)ca + ((rk * d[1]) + rj) * d[0] + ri;
        const op_cell* B = &s->cells[cb];
        for (; nch > 0; nch--) {
          uint32_t pb = *(st++);
          double dx = B->rx[pb] - A->rx[pa];
          double dy = B->ry[pb] - A->ry[pa];
          double dz = B->rz[pb] - A->rz[pa];
          double d2 = (dx * dx + dy * dy) + dz * dz;
          if (d2 > 0.0 && d2 <= rcut2) {
            const double mass_b = type_mass[B->type[pb]];
            const double r = sqrt(d2);
            const double inv_r = 1.0 / r;
            double de = G * mass_a * mass_b * inv_r * inv_r;
            de *= 1.0 / r;
            A->fx[pa] += de * dx;
            A->fy[pa] += de * dy;
            A->fz[pa] += de * dz;
          }
        }
      }
    }
  }
  return OP_OK;
}

/* unique_pair_id (src/core/include/exanb/core/particle_type_pair.h:39-55) */
static inline unsigned unique_pair_id(unsigned a, unsigned b) {
  if (a > b) { unsigned t = a; a = b; b = t; }
  return (b * (b + 1)) / 2 + a;
}

/* Multi-material traversal (compute_pair_multimat.h:100-253): sequential stream decode,
 * accepted iff d2 <= rcut[pair]^2 (no d2 > 0 test; self excluded by index), one buffer per
 * type pair, then per pair id in increasing order the LJ buffer functor
 * (lennard_jones.cu:67-102) with that pair's parameters adds its sum to f. */
int op_lj_force_multimat(op_sim* s, int ntypes, const double* eps_pair,
                         const double* sigma_pair, const double* rcut_pair) {
  if (!s->streams_valid || ntypes <= 0 || ntypes > 16) return OP_E_ARG;
  const unsigned npairs = (unsigned)(ntypes * (ntypes + 1) / 2);
  const int64_t* d = s->dims;
#pragma omp parallel
  {
    size_t cap[136];
    size_t cnt[136];
    double* buf[136];
    for (unsigned q = 0; q < npairs; q++) { cap[q] = 64; buf[q] = (double*)malloc(cap[q] * 4 * sizeof(double)); }
#pragma omp for schedule(dynamic)
    for (size_t ca = 0; ca < s->ncells; ca++) {
      if (is_ghost_cell(s, ca)) continue;
      op_cell* A = &s->cells[ca];
      const uint32_t na = A->n;
      const uint32_t* off;
      int32_t shift;
      const uint16_t* st = stream_info(s->stream[ca], na, &off, &shift);
      for (uint32_t pa = 0; pa < na; pa++) {
        for (unsigned q = 0; q < npairs; q++) cnt[q] = 0;
        const unsigned ta = A->type[pa];
        unsigned cg = *(st++);
        for (; cg > 0; cg--) {
          uint16_t enc = *(st++);
          unsigned nch = *(st++);
          int64_t ri = (int64_t)(enc & 31) - 16, rj = (int64_t)((enc >> 5) & 31) - 16,
                  rk = (int64_t)((enc >> 10) & 31) - 16;
          int64_t cb = (int64_t)ca + ((rk * d[1]) + rj) * d[0] + ri;
          const op_cell* B = &s->cells[cb];
          for (; nch > 0; nch--) {
            uint32_t pb = *(st++);
            if (pb < B->n && (cb != (int64_t)ca || pb != pa)) {
              const unsigned tb = B->type[pb];
              double dx = B->rx[pb] - A->rx[pa];
              double dy = B->ry[pb] - A->ry[pa];
              double dz = B->rz[pb] - A->rz[pa];
              double d2 = (dx * dx + dy * dy) + dz * dz;
              unsigned pid = unique_pair_id(ta, tb);
              double rc = rcut_pair[pid];
              double rc2 = rc * rc;
              if (d2 <= rc2) {
                if (cnt[pid] == cap[pid]) {
                  cap[pid] *= 2;
                  buf[pid] = (double*)realloc(buf[pid], cap[pid] * 4 * sizeof(double));
                }
                double* b = buf[pid] + 4 * cnt[pid];
                b[0] = dx; b[1] = dy; b[2] = dz; b[3] = d2;
                cnt[pid]++;
              }
            }
          }
        }
        for (unsigned q = 0; q < npairs; q++) {
          if (cnt[q] == 0) continue;
          double tl_fx = 0.0, tl_fy = 0.0, tl_fz = 0.0, tl_e = 0.0;
          for (size_t m = 0; m < cnt[q]; m++) {
            const double* b = buf[q] + 4 * m;
            const double r = sqrt(b[3]);
            double e = 0.0, de = 0.0;
            lj_compute_energy(eps_pair[q], sigma_pair[q], r, &e, &de);
            de *= 1.0 / r;
            tl_fx += de * b[0];
            tl_fy += de * b[1];
            tl_fz += de * b[2];
            tl_e += e;
          }
          A->fx[pa] += tl_fx;
          A->fy[pa] += tl_fy;
          A->fz[pa] += tl_fz;
          A->ep[pa] += tl_e;
        }
      }
    }
    for (unsigned q = 0; q < npairs; q++) free(buf[q]);
  }
  return OP_OK;
}

/* ---------------------------------------------------------- integrators */

/* push_f_v_r = PushVec3SecondOrderXForm identity branch (src/defbox/include/exanb/defbox/
 * push_vec3_2nd_order.h:33-38,85-111): dt' = dt*scale, dt2 = dt'*dt'*0.5, x += v*dt' + f*dt2 */
int op_push_f_v_r(op_sim* s, double dt, double dt_scale) {
  const double delta_t = dt * dt_scale;
  const double delta_t2 = delta_t * delta_t * 0.5;
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < s->ncells; c++) {
    if (is_ghost_cell(s, c)) continue;
    op_cell* C = &s->cells[c];
    for (uint32_t p = 0; p < C->n; p++) {
      C->rx[p] += C->vx[p] * delta_t + C->fx[p] * delta_t2;
      C->ry[p] += C->vy[p] * delta_t + C->fy[p] * delta_t2;
      C->rz[p] += C->vz[p] * delta_t + C->fz[p] * delta_t2;
    }
  }
  return OP_OK;
}

/* push_f_v = PushVec3FirstOrder identity branch (push_vec3_1st_order.h:33-38,79-97): v += f*dt' */
int op_push_f_v(op_sim* s, double dt, double dt_scale) {
  const double delta_t = dt * dt_scale;
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < s->ncells; c++) {
    if (is_ghost_cell(s, c)) continue;
    op_cell* C = &s->cells[c];
    for (uint32_t p = 0; p < C->n; p++) {
      C->vx[p] += C->fx[p] * delta_t;
      C->vy[p] += C->fy[p] * delta_t;
      C->vz[p] += C->fz[p] * delta_t;
    }
  }
  return OP_OK;
}

/* ------------------------------------------------------------ composite */

/* parallel_update_particles after move_particles (update-particles.msp:47-53,59-65):
 * rebuild_amr (sub_grid_density 6.5, update-particles.msp:8-11), backup_r,
 * ghost_comm_scheme + ghost_update_all, chunk_neighbors */
static int full_update(op_sim* s) {
  int e;
  if ((e = op_move_particles(s))) return e;
  if ((e = op_rebuild_amr(s, OP_SUB_GRID_DENSITY))) return e;
  if ((e = op_backup_r(s))) return e;
  if ((e = op_ghost_update_all(s))) return e;
  op_nbh_config cfg = {1, 0, 0, 0, 0};
  return op_chunk_neighbors(s, s->nbh_dist_lab, &cfg);
}

/* nbh_dist (src/particle_neighbors/nbh_dist.cpp:45-72) + init_particles + first forces
 * (compute-loop.msp:1-7, numerical-scheme.msp:8-11) */
int op_init(op_sim* s, double epsilon, double sigma, double rcut, double rcut_inc) {
  s->nbh_dist_lab = rcut + rcut_inc;
  s->max_displ = rcut_inc / 2.0;
  s->ghost_dist = rcut + rcut_inc;
  int e = full_update(s);
  if (e) return e;
  op_zero_force(s, 0);
  return op_lj_force(s, epsilon, sigma, rcut, 0);
}

int op_run(op_sim* s, op_run_params* p) {
  int e;
  p->n_rebuilds = 0;
  p->n_violations = 0;
  for (int64_t step = 1; step <= p->steps; step++) {
    op_push_f_v_r(s, p->dt, 1.0);
    op_push_f_v(s, p->dt, 0.5);
    int64_t over = op_particle_displ_over(s, s->max_displ);
    int trigger;
    if (p->rebuild_every > 0) {
      trigger = (step % p->rebuild_every) == 0;
      if (over > 0 && !trigger) p->n_violations++;
    } else {
      trigger = over > 0;
    }
    if (trigger) {
      if ((e = full_update(s))) return e;
      p->n_rebuilds++;
    } else {
      op_ghost_update_r(s);
    }
    op_zero_force(s, 0);
    if ((e = op_lj_force(s, p->epsilon, p->sigma, p->rcut, 0))) return e;
    op_push_f_v(s, p->dt, 0.5);
  }
  return OP_OK;
}

void op_energies(const op_sim* s, double* ke, double* pe) {
  double k = 0., u = 0.;
  for (size_t c = 0; c < s->ncells; c++) {
    if (is_ghost_cell(s, c)) continue;
    const op_cell* C = &s->cells[c];
    for (uint32_t p = 0; p < C->n; p++) {
      k += 0.5 * ((C->vx[p] * C->vx[p] + C->vy[p] * C->vy[p]) + C->vz[p] * C->vz[p]);
      u += 0.5 * C->ep[p];
    }
  }
  *ke = k;
  *pe = u;
}

int op_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
