// xnb_api.hip -- C ABI (include/xnb_amd.h) over the gfx950 kernels: one context per box,
// owning its device arrays and one HIP stream.  Host code orchestrates the reference
// operator sequence; every particle operation runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xnb_amd.h"
#include "xnb_ctx.h"

using namespace xnb;

namespace xnb_impl {

thread_local std::string g_err;

int fail(xnb_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_err = msg;
  return code;
}

inline uint32_t nb256(size_t n) { return (uint32_t)((n + 255) / 256); }

// ---- timers
hipEvent_t get_event(xnb_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);  // a failed timer event only loses that timing
  return e;
}

void flush_timers(xnb_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->st);  // errors surface on the next checked call
  for (auto& t : c->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    c->t_ms[t.id] += ms;
    c->t_n[t.id] += 1;
    c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->pending.clear();
}

void tstart(xnb_ctx* c, int id) {
  if (!c->timing) return;
  c->t_cur = id;
  c->t_a = get_event(c);
  (void)hipEventRecord(c->t_a, c->st);
}

void tstop(xnb_ctx* c) {
  if (!c->timing || c->t_cur < 0) return;
  hipEvent_t b = get_event(c);
  (void)hipEventRecord(b, c->st);
  c->pending.push_back({c->t_cur, c->t_a, b});
  c->t_cur = -1;
  if (c->pending.size() >= 2048) flush_timers(c);
}

// ---- geometry
int set_geometry(xnb_ctx* c, double ghost_dist) {
  const xnb_domain& d = c->dom;
  Geom g{};
  g.Di = (int)d.grid_dims[0];
  g.Dj = (int)d.grid_dims[1];
  g.Dk = (int)d.grid_dims[2];
  g.pi = d.periodic[0] ? 1 : 0;
  g.pj = d.periodic[1] ? 1 : 0;
  g.pk = d.periodic[2] ? 1 : 0;
  g.ox = d.bounds_min[0];
  g.oy = d.bounds_min[1];
  g.oz = d.bounds_min[2];
  g.ex = d.bounds_max[0];
  g.ey = d.bounds_max[1];
  g.ez = d.bounds_max[2];
  g.cs = d.cell_size;
  g.ghost_dist = ghost_dist;
  g.gl = (int)std::ceil(ghost_dist / d.cell_size);
  if ((g.pi && g.gl > g.Di) || (g.pj && g.gl > g.Dj) || (g.pk && g.gl > g.Dk))
    return fail(c, XNB_E_DOMAIN, "ghost layers exceed the periodic domain (one image shift per side)");
  g.di = g.Di + 2 * g.gl;
  g.dj = g.Dj + 2 * g.gl;
  g.dk = g.Dk + 2 * g.gl;
  const size_t nc = (size_t)g.di * g.dj * g.dk;
  if (nc >= (size_t)scan_max_items()) return fail(c, XNB_E_CAPACITY, "too many cells");
  g.ncells = (uint32_t)nc;
  c->g = g;
  int e;
  for (DArr<uint32_t>* a : {&c->cell_off, &c->cell_n, &c->cnt_out, &c->cnt_in, &c->new_n, &c->new_off, &c->in_cur, &c->gcnt,
                            &c->goff, &c->clen, &c->coff})
    if ((e = dalloc(c, *a, nc + 1))) return e;
  // ghost counts are written for ghost cells only; inner cells keep this 0
  HIPCHK(c, hipMemsetAsync(c->gcnt.p, 0, (nc + 1) * sizeof(uint32_t), c->st));
  return XNB_OK;
}

void bind_states(xnb_ctx* c) {
  c->A = PState{c->rA.p, c->vA[0].p, c->vA[1].p, c->vA[2].p, c->fA[0].p, c->fA[1].p, c->fA[2].p, c->eA.p, c->idA.p, {}, 0};
  c->B = PState{c->rB.p, c->vB[0].p, c->vB[1].p, c->vB[2].p, c->fB[0].p, c->fB[1].p, c->fB[2].p, c->eB.p, c->idB.p, {}, 0};
  for (int k = 0; k < c->n_extra; k++) {
    c->A.x[k] = c->xA[k].p;
    c->B.x[k] = c->xB[k].p;
  }
  c->A.nx = c->B.nx = (uint32_t)c->n_extra;
}

// (Re)allocate particle buffers to hold ncap slots; the first keep slots of buffer A are kept.
int ensure_capacity(xnb_ctx* c, size_t ncap, size_t keep) {
  if (ncap <= c->cap && c->rA.p) return XNB_OK;
  size_t newcap = std::max(ncap, c->cap + c->cap / 4);
  newcap = (newcap + 255) & ~(size_t)255;
  // grow A preserving [0, keep)
  auto grow = [&](auto& arr, size_t elems) -> int {
    using T = std::remove_pointer_t<decltype(arr.p)>;
    DArr<T> na;
    int e = dalloc(c, na, elems);
    if (e) return e;
    hipError_t he = hipSuccess;
    if (arr.p && keep) he = hipMemcpyAsync(na.p, arr.p, keep * sizeof(T), hipMemcpyDeviceToDevice, c->st);
    if (he == hipSuccess) he = hipStreamSynchronize(c->st);
    if (he != hipSuccess) {
      dfree(c, na);
      return fail(c, XNB_E_HIP, std::string("capacity copy failed: ") + hipGetErrorString(he));
    }
    dfree(c, arr);
    arr = na;
    return XNB_OK;
  };
  int e;
  if ((e = grow(c->rA, newcap))) return e;
  for (int a = 0; a < 3; a++) {
    if ((e = grow(c->vA[a], newcap))) return e;
    if ((e = grow(c->fA[a], newcap))) return e;
  }
  if ((e = grow(c->eA, newcap))) return e;
  if ((e = grow(c->idA, newcap))) return e;
  for (int k = 0; k < c->n_extra; k++)
    if ((e = grow(c->xA[k], newcap))) return e;
  if ((e = grow(c->slot_cell_A, newcap))) return e;
  // B and per-slot scratch need no content
  dfree(c, c->rB);
  if ((e = dalloc(c, c->rB, newcap))) return e;
  for (int a = 0; a < 3; a++) {
    dfree(c, c->vB[a]);
    dfree(c, c->fB[a]);
    if ((e = dalloc(c, c->vB[a], newcap))) return e;
    if ((e = dalloc(c, c->fB[a], newcap))) return e;
  }
  dfree(c, c->eB);
  dfree(c, c->idB);
  dfree(c, c->slot_cell_B);
  if ((e = dalloc(c, c->eB, newcap))) return e;
  if ((e = dalloc(c, c->idB, newcap))) return e;
  for (int k = 0; k < c->n_extra; k++) {
    dfree(c, c->xB[k]);
    if ((e = dalloc(c, c->xB[k], newcap))) return e;
  }
  if ((e = dalloc(c, c->slot_cell_B, newcap))) return e;
  for (DArr<uint32_t>* a : {&c->mv_dst, &c->pk, &c->newloc, &c->in_list, &c->plen, &c->pstart}) {
    dfree(c, *a);
    if ((e = dalloc(c, *a, newcap))) return e;
  }
  // the ghost scheme (one entry per ghost) stays valid across a growth: ghost_update_r reads it
  for (DArr<uint32_t>* a : {&c->ghost_src, &c->ghost_shift}) {
    DArr<uint32_t> na;
    if ((e = dalloc(c, na, newcap))) return e;
    const size_t ng = std::min<size_t>(c->G, a->n);
    if (a->p && ng) HIPCHK(c, hipMemcpyAsync(na.p, a->p, ng * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    dfree(c, *a);
    *a = na;
  }
  {  // backup_r of the kept (inner) slots survives: a ghost rebuild may grow the buffers after it
    DArr<uint32_t> nb;
    if ((e = dalloc(c, nb, 3 * newcap))) return e;
    if (c->backup.p && keep)
      HIPCHK(c, hipMemcpyAsync(nb.p, c->backup.p, 3 * keep * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    dfree(c, c->backup);
    c->backup = nb;
  }
  c->cap = newcap;
  bind_states(c);
  return XNB_OK;
}

void swap_buffers(xnb_ctx* c) {
  std::swap(c->rA, c->rB);
  for (int a = 0; a < 3; a++) {
    std::swap(c->vA[a], c->vB[a]);
    std::swap(c->fA[a], c->fB[a]);
  }
  std::swap(c->eA, c->eB);
  std::swap(c->idA, c->idB);
  for (int k = 0; k < c->n_extra; k++) std::swap(c->xA[k], c->xB[k]);
  std::swap(c->slot_cell_A, c->slot_cell_B);
  bind_states(c);
}

int read_u32(xnb_ctx* c, const uint32_t* dptr, uint32_t* out, int n = 1) {
  HIPCHK(c, hipMemcpyAsync(c->h_u32, dptr, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int q = 0; q < n; q++) out[q] = c->h_u32[q];
  return XNB_OK;
}

int scan(xnb_ctx* c, const uint32_t* in, uint32_t n, uint32_t add, uint32_t* out) {
  const uint32_t nb = (n + 2047) / 2048;
  int e;
  if ((e = dalloc(c, c->bsum, nb + 1))) return e;
  scan_exclusive(in, n, add, out, c->bsum.p, c->u32s.p + 1, c->st);
  return XNB_OK;
}

// ---- operators

// check false: the out-of-domain flag (u32s[0]) is left for the caller to read (op_ghost_all
// reads it with the ghost count in one transfer)
int op_move(xnb_ctx* c, bool forces = true, bool check = true) {
  const Geom& g = c->g;
  c->lists_valid = false;
  c->ghosts_valid = false;
  c->backup_valid = false;
  // the out-of-domain flag describes this move, also when there is nothing to move
  HIPCHK(c, hipMemsetAsync(c->u32s.p, 0, sizeof(uint32_t), c->st));
  if (c->N == 0) return XNB_OK;
  tstart(c, XNB_TIMER_REBIN);
  HIPCHK(c, hipMemsetAsync(c->cnt_out.p, 0, g.ncells * sizeof(uint32_t), c->st));
  HIPCHK(c, hipMemsetAsync(c->cnt_in.p, 0, g.ncells * sizeof(uint32_t), c->st));
  HIPCHK(c, hipMemsetAsync(c->in_cur.p, 0, g.ncells * sizeof(uint32_t), c->st));
  launch_move_flag(g, c->N, c->A.r, c->slot_cell_A.p, c->mv_dst.p, c->cnt_out.p, c->cnt_in.p, c->u32s.p, c->st);
  launch_move_pack(g, c->cell_off.p, c->cell_n.p, c->cnt_out.p, c->cnt_in.p, c->mv_dst.p, c->pk.p, c->newloc.p, c->new_n.p, c->st);
  int e;
  if ((e = scan(c, c->new_n.p, g.ncells, 0, c->new_off.p))) return e;
  launch_move_scatter(c->N, c->A, c->B, c->slot_cell_A.p, c->mv_dst.p, c->newloc.p, c->cell_off.p, c->new_off.p, c->cell_n.p,
                      c->cnt_out.p, c->in_cur.p, c->in_list.p, c->slot_cell_B.p, forces, c->st);
  launch_move_incoming(g, c->A, c->B, c->new_off.p, c->cell_n.p, c->cnt_out.p, c->cnt_in.p, c->in_list.p, c->slot_cell_B.p, forces,
                       c->st);
  launch_set_inner_cells(g, c->new_n.p, c->new_off.p, c->cell_n.p, c->cell_off.p, c->st);
  tstop(c);
  HIPCHK(c, hipGetLastError());
  uint32_t errflag = 0;
  if (check && (e = read_u32(c, c->u32s.p, &errflag))) return e;
  swap_buffers(c);
  // ghost cells are empty until ghost_update_all rebuilds them
  c->G = 0;
  if (errflag) return fail(c, XNB_E_OUTSIDE, "a particle left the domain along a non-periodic axis");
  return XNB_OK;
}

// move_check: also report the out-of-domain flag of a preceding op_move(c, ..., false).
// forces: images copy f and ep of their sources (false: f and ep of the images are 0)
int op_ghost_all(xnb_ctx* c, bool move_check, bool forces) {
  const Geom& g = c->g;
  int e;
  tstart(c, XNB_TIMER_GHOST);
  launch_ghost_count(g, c->A.r, c->cell_off.p, c->cell_n.p, c->gcnt.p, c->st);
  if ((e = scan(c, c->gcnt.p, g.ncells, 0, c->goff.p))) return e;
  tstop(c);
  uint32_t w[2] = {0, 0};  // u32s[0]: bit 1 out of the domain, bit 2 too dense; u32s[1]: ghost count
  if ((e = read_u32(c, c->u32s.p, w, 2))) return e;
  if (w[0] & 2u) return fail(c, XNB_E_CAPACITY, "rebuild_amr: a cell holds more than 4096 particles");
  if (move_check && w[0]) return fail(c, XNB_E_OUTSIDE, "a particle left the domain along a non-periodic axis");
  const uint32_t G = w[1];
  if ((e = ensure_capacity(c, (size_t)c->N + G, c->N))) return e;
  tstart(c, XNB_TIMER_GHOST);
  launch_ghost_fill(g, c->N, c->A, c->goff.p, c->cell_off.p, c->cell_n.p, c->slot_cell_A.p, c->ghost_src.p, c->ghost_shift.p,
                    (uint32_t)std::min<size_t>(c->cap, 0xffffffffu), c->u32s.p + 7, forces, c->st);
  tstop(c);
  HIPCHK(c, hipGetLastError());
  c->G = G;
  c->ghosts_valid = true;
  c->lists_valid = false;
  return XNB_OK;
}

// rebuild_amr after a move (ghost cells empty): dense inner cells sorted by sub-cell through
// the spare buffers; u32s[8] counts the listed cells, a cell too dense to sort sets bit 2 of
// u32s[0] (reported with the ghost count, or by xnb_rebuild_amr)
int op_amr(xnb_ctx* c, double density, bool forces) {
  if (c->N == 0) return XNB_OK;
  int e;
  if ((e = dalloc(c, c->amr_list, c->g.ncells))) return e;
  HIPCHK(c, hipMemsetAsync(c->u32s.p + 8, 0, sizeof(uint32_t), c->st));
  launch_rebuild_amr(c->g, density, c->A, c->B, c->cell_off.p, c->cell_n.p, c->amr_list.p, c->u32s.p + 8, c->u32s.p, forces, c->st);
  HIPCHK(c, hipGetLastError());
  return XNB_OK;
}

int op_backup(xnb_ctx* c) {
  launch_backup_r(c->g, c->N, c->A.r, c->slot_cell_A.p, c->backup.p, c->st);
  HIPCHK(c, hipGetLastError());
  c->backup_valid = true;
  return XNB_OK;
}

constexpr int NOT_APPLICABLE = 1;

inline uint32_t list_granule() { return std::max<uint32_t>(16u, std::max(lj_list_granule(), brick_list_granule())); }
inline uint32_t grow_cap(uint32_t mx) {
  const uint32_t gran = list_granule();  // force kernels load rows in batches of the granule
  return (mx + mx / 8 + gran - 1) / gran * gran;
}

// Brick dims, T and Rmax for the current cell occupancy; NOT_APPLICABLE if nothing fits LDS.
int choose_bricks(xnb_ctx* c, uint32_t cap, BrickParams& out) {
  // (4,3,2) and (3,3,2) only for batched boxes: under the batch's LDS cap they beat (4,2,2)
  // (512-box batch: 3.58e9 -> 3.90e9 atom-steps/s), while the two-species slab run alone is
  // faster on (4,2,2) bricks
  static const int cand[][3] = {{4, 4, 2}, {4, 3, 2}, {3, 3, 2}, {4, 2, 2}, {2, 2, 2}, {4, 2, 1}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
  const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
  int e;
  for (int q = -1; q < ncand; q++) {
    if ((q == 1 || q == 2) && !c->batch_shapes) continue;
    int dims[3];
    if (q < 0) {
      if (c->brick_dims[0] <= 0) continue;
      dims[0] = c->brick_dims[0];
      dims[1] = c->brick_dims[1];
      dims[2] = c->brick_dims[2];
    } else {
      dims[0] = cand[q][0];
      dims[1] = cand[q][1];
      dims[2] = cand[q][2];
    }
    BrickParams bp = brick_params(c->g, dims[0], dims[1], dims[2]);
    HIPCHK(c, hipMemsetAsync(c->u32s.p + 3, 0, 2 * sizeof(uint32_t), c->st));
    launch_brick_stats(c->g, bp, c->cell_n.p, c->u32s.p + 3, c->st);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_u32, c->u32s.p + 3, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    // head room for later rebuilds of a batched box (brick_margin > 1): atoms may regroup
    const uint32_t nb_max = (uint32_t)std::ceil(c->h_u32[0] * c->brick_margin);
    const uint32_t r_max = (uint32_t)std::ceil(c->h_u32[1] * c->brick_margin);
    uint32_t T = std::max<uint32_t>({nb_max, (uint32_t)bp.nrc, (uint32_t)bp.nbc, 64u});
    T = (T + 63) / 64 * 64;
    if (T > 1024 || r_max >= 65535u) continue;
    bp.T = T;
    bp.Rmax = std::max<uint32_t>(r_max, 1u);
    bp.cap = cap;
    bp.status = c->u32s.p + 6;
    if (brick_layout(bp) > c->lds_limit) continue;
    out = bp;
    return XNB_OK;
  }
  (void)e;
  return NOT_APPLICABLE;
}

int op_nbh_brick(xnb_ctx* c, double nbh_dist, const xnb_chunk_neighbors_config& cfg, const PairCut& pc) {
  const double max2 = nbh_dist * nbh_dist;
  int e;
  uint32_t cap = (c->nbr_cap + list_granule() - 1) / list_granule() * list_granule();
  for (int attempt = 0; attempt < 3; attempt++) {
    BrickParams bp;
    if ((e = choose_bricks(c, cap, bp))) return e;
    if ((e = dalloc(c, c->blist, (size_t)bp.nbricks * bp.cap * bp.T))) return e;
    if ((e = dalloc(c, c->bcnt, (size_t)bp.nbricks * bp.T))) return e;
    HIPCHK(c, hipMemsetAsync(c->u32s.p + 2, 0, sizeof(uint32_t), c->st));
    HIPCHK(c, hipMemsetAsync(c->u32s.p + 6, 0, sizeof(uint32_t), c->st));
    tstart(c, XNB_TIMER_NBH);
    hipError_t he = launch_nbh_brick(c->g, bp, c->A.r, c->cell_off.p, c->cell_n.p, max2, pc, cfg.half_symmetric, cfg.skip_ghosts,
                                     c->nbh_prefilter, c->blist.p, c->bcnt.p, c->u32s.p + 2, c->st);
    tstop(c);
    if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick neighbour build: ") + hipGetErrorString(he));
    uint32_t w[5];  // u32s[2]: longest list, u32s[6]: brick status
    if ((e = read_u32(c, c->u32s.p + 2, w, 5))) return e;
    const uint32_t mx = w[0], bad = w[4];
    if (bad) return fail(c, XNB_E_STATE, "brick sizes changed between statistics and build");
    c->max_neighbors = mx;
    if (mx <= bp.cap) {
      c->bp = bp;
      c->nbr_cap = bp.cap;
      c->list_kind = 1;
      c->slot_lists_current = false;
      c->list_atoms = c->N;
      c->list_cfg = cfg;
      c->lists_valid = true;
      c->lists_pruned = pc.nt > 0;
      return XNB_OK;
    }
    cap = grow_cap(mx);
  }
  return NOT_APPLICABLE;
}

int op_nbh(xnb_ctx* c, double nbh_dist, const xnb_chunk_neighbors_config& cfg, int all_cells, const PairCut* pc) {
  const int lay = (int)std::ceil(nbh_dist / c->g.cs);  // cell_layers (amr_grid_algorithm.h:451)
  if (lay > 15) return fail(c, XNB_E_ARG, "neighbour distance spans more than 15 cells");
  int e;
  c->lists_pruned = false;
  if (!all_cells && lay == 1 && c->g.gl >= 1 && c->use_bricks && c->N > 0) {
    const PairCut none{};
    e = op_nbh_brick(c, nbh_dist, cfg, (pc && !cfg.half_symmetric && !cfg.skip_ghosts) ? *pc : none);
    if (e != NOT_APPLICABLE) return e;
  }
  const uint32_t n_atoms = all_cells ? c->N + c->G : c->N;
  const double max2 = nbh_dist * nbh_dist;
  c->nbr_cap = (c->nbr_cap + list_granule() - 1) / list_granule() * list_granule();
  for (int attempt = 0; attempt < 3; attempt++) {
    const size_t rows = ((size_t)n_atoms + 63) / 64 * 64;
    if ((e = dalloc(c, c->nbr, rows * c->nbr_cap))) return e;
    if ((e = dalloc(c, c->nbr_cnt, rows))) return e;
    HIPCHK(c, hipMemsetAsync(c->u32s.p + 2, 0, sizeof(uint32_t), c->st));
    tstart(c, XNB_TIMER_NBH);
    launch_nbh_build(c->g, n_atoms, c->A.r, c->slot_cell_A.p, c->cell_off.p, c->cell_n.p, max2, lay, cfg.half_symmetric,
                     cfg.skip_ghosts, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, c->u32s.p + 2, c->st);
    tstop(c);
    HIPCHK(c, hipGetLastError());
    uint32_t mx = 0;
    if ((e = read_u32(c, c->u32s.p + 2, &mx))) return e;
    c->max_neighbors = mx;
    if (mx <= c->nbr_cap) {
      c->list_kind = 0;
      c->slot_lists_current = true;
      c->list_atoms = n_atoms;
      c->list_cfg = cfg;
      c->lists_valid = true;
      return XNB_OK;
    }
    c->nbr_cap = grow_cap(mx);
  }
  return fail(c, XNB_E_OVERFLOW, "neighbour list capacity");
}

// Generic per-slot lists (nbr, nbr_cnt) for the stream export and generic traversals.
int ensure_slot_lists(xnb_ctx* c) {
  if (c->list_kind == 0 || c->slot_lists_current) return XNB_OK;
  const size_t rows = ((size_t)c->N + 63) / 64 * 64;
  int e;
  if ((e = dalloc(c, c->nbr, rows * c->bp.cap))) return e;
  if ((e = dalloc(c, c->nbr_cnt, rows))) return e;
  hipError_t he = launch_brick_to_slots(c->g, c->bp, c->cell_off.p, c->cell_n.p, c->blist.p, c->bcnt.p, c->nbr.p, c->nbr_cnt.p,
                                        c->bp.cap, c->st);
  if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick list conversion: ") + hipGetErrorString(he));
  c->nbr_cap = c->bp.cap;
  c->slot_lists_current = true;
  return XNB_OK;
}

// LJ force over the current lists: brick kernel for brick lists, generic kernel otherwise.
// With next (brick lists only, no accum), the kernel also advances the next step's first
// half into buffer B's position array, which the caller then swaps in.
int launch_force(xnb_ctx* c, int precision, bool accum, bool push, bool ghost, double eps, double sigma, double rcut, double dth,
                 const NextStep* next = nullptr) {
  if (c->list_kind == 1 && !ghost) {
    const int epi = next ? EPI_NEXT : (push ? EPI_KICK : EPI_NONE);
    hipError_t he = launch_lj_brick(precision, accum, epi, c->g, c->bp, c->A, c->cell_off.p, c->cell_n.p, c->blist.p, c->bcnt.p, eps,
                                    sigma, rcut, dth, next, c->st);
    if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick LJ kernel: ") + hipGetErrorString(he));
    return XNB_OK;
  }
  if (next) return fail(c, XNB_E_STATE, "fused next-step push needs brick lists");
  const uint32_t n_atoms = ghost ? c->N + c->G : c->N;
  if (c->list_kind != 0 || c->list_atoms < n_atoms) return fail(c, XNB_E_STATE, "neighbour lists do not cover the requested cells");
  launch_lj(precision, accum, push, n_atoms, c->A, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, eps, sigma, rcut, dth, c->st);
  HIPCHK(c, hipGetLastError());
  return XNB_OK;
}

// Multimat force of a run-loop step: zero + compute_pair_multimat (+ the epilogue) in one brick
// kernel, or, on generic lists, the separate kernels (no fusion).
int launch_force_multimat(xnb_ctx* c, int precision, const LJMulti& mm, int ntypes, double dth, const NextStep* next) {
  if (c->list_kind == 1) {
    const int epi = next ? EPI_NEXT : EPI_KICK;
    hipError_t he = launch_ljmm_brick(precision, ntypes, epi, c->g, c->bp, c->A, c->cell_off.p, c->cell_n.p, c->blist.p, c->bcnt.p, mm,
                                      dth, next, c->st);
    if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick multimat kernel: ") + hipGetErrorString(he));
    return XNB_OK;
  }
  if (next) return fail(c, XNB_E_STATE, "fused next-step push needs brick lists");
  launch_zero_force(c->N, c->A, c->st);
  if (launch_lj_multimat(precision, c->N, c->A, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, ntypes, mm, c->st))
    return fail(c, XNB_E_ARG, "too many type pairs");
  launch_push_f_v(c->N, c->A, dth, c->st);
  HIPCHK(c, hipGetLastError());
  return XNB_OK;
}

int full_update(xnb_ctx* c, bool forces, const PairCut* pc) {
  int e;
  if ((e = op_move(c, forces, false))) return e;
  if ((e = op_amr(c, kSubGridDensity, forces))) return e;
  if ((e = op_backup(c))) return e;
  if ((e = op_ghost_all(c, true, forces))) return e;
  xnb_chunk_neighbors_config cfg{1, 0, 0, 0, 0};
  return op_nbh(c, c->nbh_dist_lab, cfg, 0, pc);
}

// Lists a multi-material run built with per type-pair radii hold only the pairs of their
// pair's radius; any other consumer gets full lists (nbh_dist_lab, inner cells) rebuilt on the
// same positions first.
int ensure_full_lists(xnb_ctx* c) {
  if (!c->lists_valid || !c->lists_pruned) return XNB_OK;
  xnb_chunk_neighbors_config cfg{1, 0, 0, 0, 0};
  return op_nbh(c, c->nbh_dist_lab, cfg, 0, nullptr);
}

int check_ctx(xnb_ctx* c) {
  if (!c) return XNB_E_ARG;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(c, XNB_E_HIP, hipGetErrorString(e));
  return XNB_OK;
}

}  // namespace xnb_impl

using namespace xnb_impl;

// =========================================================================== C ABI

extern "C" {

int xnb_version(void) { return 100; }

const char* xnb_error_string(int code) {
  switch (code) {
    case XNB_OK: return "ok";
    case XNB_E_ARG: return "invalid argument";
    case XNB_E_DOMAIN: return "invalid domain";
    case XNB_E_OUTSIDE: return "particle outside the domain";
    case XNB_E_STATE: return "operator called in an invalid state";
    case XNB_E_HIP: return "HIP runtime error";
    case XNB_E_NOMEM: return "device allocation failed";
    case XNB_E_OVERFLOW: return "neighbour list overflow";
    case XNB_E_CAPACITY: return "capacity exceeded";
    default: return "unknown error";
  }
}

const char* xnb_last_error(const xnb_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int xnb_device_count(int* count) {
  if (!count) return XNB_E_ARG;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return XNB_OK;
}

int xnb_create(xnb_ctx** out, const xnb_domain* d, int device) {
  if (!out || !d) return fail(nullptr, XNB_E_ARG, "null argument");
  *out = nullptr;
  for (int a = 0; a < 3; a++) {
    if (d->grid_dims[a] <= 0 || !(d->cell_size > 0.)) return fail(nullptr, XNB_E_ARG, "bad grid dims or cell size");
    // check_domain (src/core/lib/domain.cpp:78-85)
    const double gsz = (double)d->grid_dims[a] * d->cell_size;
    const double dsz = d->bounds_max[a] - d->bounds_min[a];
    if (!(std::fabs(1. - gsz / dsz) < 1.e-15)) return fail(nullptr, XNB_E_DOMAIN, "grid_dims*cell_size != domain size");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(nullptr, XNB_E_HIP, "no HIP device");
  if (device < 0 || device >= ndev) return fail(nullptr, XNB_E_ARG, "bad device index");
  xnb_ctx* c = new xnb_ctx();
  c->device = device;
  c->dom = *d;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, XNB_E_HIP, "stream creation failed");
  }
  int e;
  if ((e = set_geometry(c, 0.0)) || (e = dalloc(c, c->u32s, 16)) || (e = dalloc(c, c->u64s, 8)) || (e = dalloc(c, c->red_out, 2))) {
    xnb_destroy(c);
    return e;
  }
  // flags and counters start at 0 (hipMalloc does not clear)
  if (hipMemsetAsync(c->u32s.p, 0, c->u32s.n * sizeof(uint32_t), c->st) != hipSuccess ||
      hipMemsetAsync(c->u64s.p, 0, c->u64s.n * sizeof(unsigned long long), c->st) != hipSuccess) {
    xnb_destroy(c);
    return fail(nullptr, XNB_E_HIP, "scalar buffer clear failed");
  }
  if (hipHostMalloc((void**)&c->h_u32, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_u64, 64 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    xnb_destroy(c);
    return fail(nullptr, XNB_E_NOMEM, "pinned allocation failed");
  }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_limit = (uint32_t)lds;
  if (const char* v = std::getenv("XNB_BRICKS")) {  // "0" disables the brick path, "i,j,k" fixes brick dims
    if (v[0] == '0' && v[1] == 0) c->use_bricks = false;
    else std::sscanf(v, "%d,%d,%d", &c->brick_dims[0], &c->brick_dims[1], &c->brick_dims[2]);
  }
  if (const char* v = std::getenv("XNB_FUSE_PUSH")) c->fuse_push = v[0] != '0';
  if (const char* v = std::getenv("XNB_NBH_PREFILTER")) c->nbh_prefilter = v[0] != '0';
  HIPCHK(c, hipMemsetAsync(c->cell_n.p, 0, c->cell_n.n * sizeof(uint32_t), c->st));
  HIPCHK(c, hipMemsetAsync(c->cell_off.p, 0, c->cell_off.n * sizeof(uint32_t), c->st));
  *out = c;
  return XNB_OK;
}

int xnb_destroy(xnb_ctx* c) {
  if (!c) return XNB_OK;
  // teardown: failures here leave nothing to recover, so results are not checked
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto ev : c->pool) (void)hipEventDestroy(ev);
  dfree(c, c->slot_cell_A); dfree(c, c->slot_cell_B); dfree(c, c->rA); dfree(c, c->rB);
  for (int a = 0; a < 3; a++) { dfree(c, c->vA[a]); dfree(c, c->vB[a]); dfree(c, c->fA[a]); dfree(c, c->fB[a]); }
  dfree(c, c->eA); dfree(c, c->eB); dfree(c, c->idA); dfree(c, c->idB);
  for (int k = 0; k < kMaxExtraFields; k++) { dfree(c, c->xA[k]); dfree(c, c->xB[k]); }
  for (DArr<uint32_t>* a : {&c->cell_off, &c->cell_n, &c->cnt_out, &c->cnt_in, &c->new_n, &c->new_off, &c->in_cur, &c->gcnt, &c->goff,
                            &c->clen, &c->coff, &c->mv_dst, &c->pk, &c->newloc, &c->in_list, &c->plen, &c->pstart, &c->ghost_src,
                            &c->ghost_shift, &c->backup, &c->nbr, &c->nbr_cnt, &c->bsum, &c->u32s, &c->amr_list})
    dfree(c, *a);
  dfree(c, c->blist); dfree(c, c->bcnt);
  dfree(c, c->stream_buf); dfree(c, c->u64s); dfree(c, c->step_counts); dfree(c, c->red_part); dfree(c, c->red_out);
  if (c->h_u32) (void)hipHostFree(c->h_u32);
  if (c->h_u64) (void)hipHostFree(c->h_u64);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return XNB_OK;
}

int xnb_nbh_dist(xnb_ctx* c, double rcut_max, double rcut_inc, double* nbh_dist_lab, double* max_displ, double* ghost_dist) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!(rcut_max >= 0.) || !(rcut_inc >= 0.)) return fail(c, XNB_E_ARG, "negative distance");
  // nbh_dist.cpp:47-71 with identity xform and no bonds
  const double lab = rcut_max + rcut_inc;
  const double md = rcut_inc / 2.0;
  const double gd = std::max(rcut_inc, rcut_max + rcut_inc);
  const int new_gl = (int)std::ceil(gd / c->dom.cell_size);
  if (c->N > 0 && new_gl != c->g.gl) {
    // re-grid: keep particles, in cell order (per-cell order preserved by stable binning)
    std::vector<double> r(3 * (size_t)c->N), v(3 * (size_t)c->N);
    std::vector<uint64_t> id(c->N);
    std::vector<uint8_t> ty(c->N);
    if ((e = xnb_get_particles(c, 0, nullptr, r.data(), v.data(), nullptr, nullptr, id.data(), ty.data()))) return e;
    const size_t n = c->N;
    if ((e = set_geometry(c, gd))) return e;
    c->nbh_set = true;
    if ((e = xnb_set_particles(c, n, r.data(), v.data(), id.data(), ty.data()))) return e;
  } else if ((e = set_geometry(c, gd))) {
    return e;
  } else if (c->N == 0) {
    HIPCHK(c, hipMemsetAsync(c->cell_n.p, 0, c->cell_n.n * sizeof(uint32_t), c->st));
    HIPCHK(c, hipMemsetAsync(c->cell_off.p, 0, c->cell_off.n * sizeof(uint32_t), c->st));
  }
  c->nbh_set = true;
  c->nbh_dist_lab = lab;
  c->max_displ = md;
  c->ghost_dist = gd;
  c->lists_valid = false;
  if (nbh_dist_lab) *nbh_dist_lab = lab;
  if (max_displ) *max_displ = md;
  if (ghost_dist) *ghost_dist = gd;
  return XNB_OK;
}

int xnb_set_particles(xnb_ctx* c, size_t n, const double* r, const double* v, const uint64_t* id, const uint8_t* type) {
  return xnb_set_particles_with_fields(c, n, r, v, id, type, 0, nullptr);
}

int xnb_set_particles_with_fields(xnb_ctx* c, size_t n, const double* r, const double* v, const uint64_t* id,
                                  const uint8_t* type, int n_extra, const double* const* extra) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (n > 0 && !r) return fail(c, XNB_E_ARG, "null positions");
  if (n >= 0x7FFFFFFFu) return fail(c, XNB_E_CAPACITY, "too many particles");
  if (n_extra < 0 || n_extra > kMaxExtraFields) return fail(c, XNB_E_ARG, "0 to 4 extra fields");
  for (int k = 0; k < n_extra; k++)
    if (n > 0 && (!extra || !extra[k])) return fail(c, XNB_E_ARG, "null extra field");
  const Geom& g = c->g;
  // host binning in input order (domain_periodic_location, domain.h:160-191)
  std::vector<uint32_t> cell(n), count(g.ncells, 0);
  std::vector<double> wr(3 * n);
  for (size_t p = 0; p < n; p++) {
    double x = r[3 * p], y = r[3 * p + 1], z = r[3 * p + 2];
    int li, lj, lk;
    periodic_location(g, x, y, z, li, lj, lk);
    if (li < 0 || lj < 0 || lk < 0 || li >= g.Di || lj >= g.Dj || lk >= g.Dk)
      return fail(c, XNB_E_OUTSIDE, "particle outside a non-periodic domain");
    wr[3 * p] = x;
    wr[3 * p + 1] = y;
    wr[3 * p + 2] = z;
    cell[p] = cell_index(g, li + g.gl, lj + g.gl, lk + g.gl);
    count[cell[p]]++;
  }
  std::vector<uint32_t> off(g.ncells + 1, 0);
  uint32_t acc = 0;
  for (uint32_t q = 0; q < g.ncells; q++) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    off[q] = acc;
    if (!is_ghost(g, i, j, k)) acc += count[q];
  }
  for (uint32_t q = 0; q < g.ncells; q++) {  // ghost cells: empty, placed after inner slots
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    if (is_ghost(g, i, j, k)) off[q] = acc;
  }
  std::vector<uint32_t> cur(off.begin(), off.end() - 1);
  std::vector<double4> hr(n);
  std::vector<double> hv[3];
  for (int a = 0; a < 3; a++) hv[a].assign(n, 0.0);
  std::vector<uint64_t> hid(n);
  std::vector<uint32_t> hsc(n);
  std::vector<double> hx[kMaxExtraFields];
  for (int k = 0; k < n_extra; k++) hx[k].resize(n);
  for (size_t p = 0; p < n; p++) {
    const uint32_t t = cur[cell[p]]++;
    hr[t] = make_double4(wr[3 * p], wr[3 * p + 1], wr[3 * p + 2], type ? (double)type[p] : 0.0);
    for (int a = 0; a < 3; a++) hv[a][t] = v ? v[3 * p + a] : 0.0;
    hid[t] = id ? id[p] : (uint64_t)p;
    hsc[t] = cell[p];
    for (int k = 0; k < n_extra; k++) hx[k][t] = extra[k][p];
  }
  c->N = (uint32_t)n;
  c->G = 0;
  const size_t est = n + std::max<size_t>(n / 2, 4096);
  c->n_extra = n_extra;
  if ((e = ensure_capacity(c, est, 0))) return e;
  for (int k = 0; k < n_extra; k++)  // fields declared on a context that already had capacity
    if ((e = dalloc(c, c->xA[k], c->cap)) || (e = dalloc(c, c->xB[k], c->cap))) return e;
  bind_states(c);
  for (int k = 0; k < n_extra; k++)
    if (n > 0) HIPCHK(c, hipMemcpyAsync(c->xA[k].p, hx[k].data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(c->A.r, hr.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->A.vx, hv[0].data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->A.vy, hv[1].data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->A.vz, hv[2].data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->A.id, hid.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->slot_cell_A.p, hsc.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
    launch_zero_force((uint32_t)n, c->A, c->st);
  }
  std::vector<uint32_t> cn(g.ncells);
  for (uint32_t q = 0; q < g.ncells; q++) {
    int i, j, k;
    cell_ijk(g, q, i, j, k);
    cn[q] = is_ghost(g, i, j, k) ? 0u : count[q];
  }
  HIPCHK(c, hipMemcpyAsync(c->cell_n.p, cn.data(), g.ncells * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->cell_off.p, off.data(), g.ncells * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->lists_valid = c->ghosts_valid = c->backup_valid = false;
  return XNB_OK;
}

// ---- snapshots: the inner particles in slot order with everything the step loop carries
// from one step to the next (r with type, v, f, ep, id, per-cell counts, backup_r), so a box
// restored from one continues bitwise as if it had not stopped.
namespace {
struct SnapHeader {
  char magic[8];  // "XNBSNAP1"
  uint32_t version, n, ncells, n_extra;  // n_extra: extra fields after the backup block
  double nbh_dist_lab, max_displ, ghost_dist;
  xnb_domain domain;
};
constexpr uint32_t kSnapVersion = 1;

bool same_domain(const xnb_domain& a, const xnb_domain& b) {
  for (int q = 0; q < 3; q++)
    if (a.bounds_min[q] != b.bounds_min[q] || a.bounds_max[q] != b.bounds_max[q] || a.grid_dims[q] != b.grid_dims[q] ||
        (a.periodic[q] != 0) != (b.periodic[q] != 0))
      return false;
  return a.cell_size == b.cell_size;
}

size_t snap_bytes(uint32_t n, uint32_t ncells, uint32_t n_extra) {
  return sizeof(SnapHeader) + (size_t)ncells * 4 + (size_t)n * (32 + 24 + 24 + 8 + 8 + 12 + 8 * (size_t)n_extra);
}
}  // namespace

int xnb_snapshot_size(xnb_ctx* c, size_t* bytes) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!bytes) return fail(c, XNB_E_ARG, "null argument");
  *bytes = snap_bytes(c->N, c->g.ncells, (uint32_t)c->n_extra);
  return XNB_OK;
}

int xnb_snapshot_save(xnb_ctx* c, void* buf, size_t bytes) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!buf || bytes < snap_bytes(c->N, c->g.ncells, (uint32_t)c->n_extra)) return fail(c, XNB_E_CAPACITY, "snapshot buffer too small");
  if (!c->backup_valid || !c->nbh_set) return fail(c, XNB_E_STATE, "snapshot needs an initialised box (xnb_init)");
  const uint32_t n = c->N, nc = c->g.ncells;
  SnapHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, "XNBSNAP1", 8);
  h.version = kSnapVersion;
  h.n = n;
  h.ncells = nc;
  h.n_extra = (uint32_t)c->n_extra;
  h.nbh_dist_lab = c->nbh_dist_lab;
  h.max_displ = c->max_displ;
  h.ghost_dist = c->ghost_dist;
  h.domain = c->dom;
  char* o = (char*)buf;
  std::memcpy(o, &h, sizeof(h));
  o += sizeof(h);
  std::vector<uint32_t> cn(nc);
  HIPCHK(c, hipMemcpyAsync(cn.data(), c->cell_n.p, nc * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (uint32_t q = 0; q < nc; q++) {  // ghost cells are rebuilt on load
    int i, j, k;
    cell_ijk(c->g, q, i, j, k);
    if (is_ghost(c->g, i, j, k)) cn[q] = 0;
  }
  std::memcpy(o, cn.data(), nc * sizeof(uint32_t));
  o += (size_t)nc * 4;
  auto d2h = [&](const void* src, size_t sz) -> int {
    HIPCHK(c, hipMemcpyAsync(o, src, sz, hipMemcpyDeviceToHost, c->st));
    o += sz;
    return XNB_OK;
  };
  if ((e = d2h(c->A.r, (size_t)n * 32)) || (e = d2h(c->A.vx, (size_t)n * 8)) || (e = d2h(c->A.vy, (size_t)n * 8)) ||
      (e = d2h(c->A.vz, (size_t)n * 8)) || (e = d2h(c->A.fx, (size_t)n * 8)) || (e = d2h(c->A.fy, (size_t)n * 8)) ||
      (e = d2h(c->A.fz, (size_t)n * 8)) || (e = d2h(c->A.ep, (size_t)n * 8)) || (e = d2h(c->A.id, (size_t)n * 8)) ||
      (e = d2h(c->backup.p, (size_t)n * 12)))
    return e;
  for (int k = 0; k < c->n_extra; k++)
    if ((e = d2h(c->xA[k].p, (size_t)n * 8))) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

// Load into a context created (xnb_create) with the snapshot's domain; replaces its particles,
// rebuilds ghosts and neighbour lists from the restored slots (no re-binning).
int xnb_snapshot_load(xnb_ctx* c, const void* buf, size_t bytes) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!buf || bytes < sizeof(SnapHeader)) return fail(c, XNB_E_ARG, "bad snapshot buffer");
  SnapHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, "XNBSNAP1", 8) != 0 || h.version != kSnapVersion) return fail(c, XNB_E_ARG, "not a snapshot");
  if (!same_domain(h.domain, c->dom)) return fail(c, XNB_E_DOMAIN, "snapshot of another domain");
  if (h.n_extra > (uint32_t)kMaxExtraFields) return fail(c, XNB_E_ARG, "snapshot with more extra fields than supported");
  if (bytes < snap_bytes(h.n, h.ncells, h.n_extra)) return fail(c, XNB_E_ARG, "truncated snapshot");
  if ((e = set_geometry(c, h.ghost_dist))) return e;
  if (h.ncells != c->g.ncells) return fail(c, XNB_E_DOMAIN, "snapshot grid differs");
  c->nbh_set = true;
  c->nbh_dist_lab = h.nbh_dist_lab;
  c->max_displ = h.max_displ;
  c->ghost_dist = h.ghost_dist;
  const uint32_t n = h.n, nc = h.ncells;
  const char* o = (const char*)buf + sizeof(h);
  std::vector<uint32_t> cn(nc), off(nc + 1), sc(n);
  std::memcpy(cn.data(), o, nc * sizeof(uint32_t));
  o += (size_t)nc * 4;
  uint32_t acc = 0;
  for (uint32_t q = 0; q < nc; q++) {
    off[q] = acc;
    acc += cn[q];  // ghost cells hold 0
    for (uint32_t p = 0; p < cn[q] && off[q] + p < n; p++) sc[off[q] + p] = q;
  }
  if (acc != n) return fail(c, XNB_E_ARG, "snapshot cell counts do not match its particle count");
  for (uint32_t q = 0; q < nc; q++) {  // ghost cells: empty, after the inner slots
    int i, j, k;
    cell_ijk(c->g, q, i, j, k);
    if (is_ghost(c->g, i, j, k)) off[q] = n;
  }
  c->N = n;
  c->G = 0;
  c->n_extra = (int)h.n_extra;
  if ((e = ensure_capacity(c, (size_t)n + std::max<size_t>(n / 2, 4096), 0))) return e;
  for (int k = 0; k < c->n_extra; k++)
    if ((e = dalloc(c, c->xA[k], c->cap)) || (e = dalloc(c, c->xB[k], c->cap))) return e;
  bind_states(c);
  auto h2d = [&](void* dst, size_t sz) -> int {
    HIPCHK(c, hipMemcpyAsync(dst, o, sz, hipMemcpyHostToDevice, c->st));
    o += sz;
    return XNB_OK;
  };
  if ((e = h2d(c->A.r, (size_t)n * 32)) || (e = h2d(c->A.vx, (size_t)n * 8)) || (e = h2d(c->A.vy, (size_t)n * 8)) ||
      (e = h2d(c->A.vz, (size_t)n * 8)) || (e = h2d(c->A.fx, (size_t)n * 8)) || (e = h2d(c->A.fy, (size_t)n * 8)) ||
      (e = h2d(c->A.fz, (size_t)n * 8)) || (e = h2d(c->A.ep, (size_t)n * 8)) || (e = h2d(c->A.id, (size_t)n * 8)) ||
      (e = h2d(c->backup.p, (size_t)n * 12)))
    return e;
  for (int k = 0; k < c->n_extra; k++)
    if ((e = h2d(c->xA[k].p, (size_t)n * 8))) return e;
  HIPCHK(c, hipMemcpyAsync(c->slot_cell_A.p, sc.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->cell_n.p, cn.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->cell_off.p, off.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->backup_valid = true;
  c->lists_valid = false;
  if ((e = op_ghost_all(c))) return e;
  xnb_chunk_neighbors_config cfg{1, 0, 0, 0, 0};
  if ((e = op_nbh(c, c->nbh_dist_lab, cfg, 0))) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_num_particles(const xnb_ctx* c, size_t* n_inner, size_t* n_ghost) {
  if (!c) return XNB_E_ARG;
  if (n_inner) *n_inner = c->N;
  if (n_ghost) *n_ghost = c->G;
  return XNB_OK;
}

int xnb_grid_info(const xnb_ctx* c, int64_t dims[3], int* ghost_layers, size_t* n_cells) {
  if (!c) return XNB_E_ARG;
  if (dims) {
    dims[0] = c->g.di;
    dims[1] = c->g.dj;
    dims[2] = c->g.dk;
  }
  if (ghost_layers) *ghost_layers = c->g.gl;
  if (n_cells) *n_cells = c->g.ncells;
  return XNB_OK;
}

int xnb_get_cell_counts(xnb_ctx* c, uint32_t* counts) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!counts) return fail(c, XNB_E_ARG, "null output");
  HIPCHK(c, hipMemcpyAsync(counts, c->cell_n.p, c->g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_field_count(const xnb_ctx* c, int* n_extra) {
  if (!c || !n_extra) return XNB_E_ARG;
  *n_extra = c->n_extra;
  return XNB_OK;
}

int xnb_get_fields(xnb_ctx* c, int include_ghosts, int n_extra, double* const* out) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (n_extra < 0 || n_extra > c->n_extra || (n_extra && !out)) return fail(c, XNB_E_ARG, "more fields than the grid holds");
  const size_t tot = (size_t)c->N + c->G;
  const Geom& g = c->g;
  std::vector<double> hx[kMaxExtraFields];
  std::vector<uint32_t> off(g.ncells), cn(g.ncells);
  for (int k = 0; k < n_extra; k++) {
    hx[k].resize(tot);
    if (tot) HIPCHK(c, hipMemcpyAsync(hx[k].data(), c->xA[k].p, tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(off.data(), c->cell_off.p, g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(cn.data(), c->cell_n.p, g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  size_t q2 = 0;
  for (uint32_t q = 0; q < g.ncells; q++) {  // the order of xnb_get_particles
    int i, j, kk;
    cell_ijk(g, q, i, j, kk);
    if (!include_ghosts && is_ghost(g, i, j, kk)) continue;
    for (uint32_t p = 0; p < cn[q]; p++, q2++)
      for (int k = 0; k < n_extra; k++)
        if (out[k]) out[k][q2] = hx[k][(size_t)off[q] + p];
  }
  return XNB_OK;
}

int xnb_get_particles(xnb_ctx* c, int include_ghosts, uint32_t* cell, double* r, double* v, double* f, double* ep, uint64_t* id,
                      uint8_t* type) {
  int e;
  if ((e = check_ctx(c))) return e;
  const size_t tot = (size_t)c->N + c->G;
  const Geom& g = c->g;
  std::vector<double4> hr(tot);
  std::vector<double> hv[3], hf[3], he(tot);
  std::vector<uint64_t> hid(tot);
  std::vector<uint32_t> off(g.ncells), cn(g.ncells);
  for (int a = 0; a < 3; a++) {
    hv[a].resize(tot);
    hf[a].resize(tot);
  }
  if (tot) {
    HIPCHK(c, hipMemcpyAsync(hr.data(), c->A.r, tot * sizeof(double4), hipMemcpyDeviceToHost, c->st));
    double* dv[3] = {c->A.vx, c->A.vy, c->A.vz};
    double* df[3] = {c->A.fx, c->A.fy, c->A.fz};
    for (int a = 0; a < 3; a++) {
      HIPCHK(c, hipMemcpyAsync(hv[a].data(), dv[a], tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipMemcpyAsync(hf[a].data(), df[a], tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipMemcpyAsync(he.data(), c->A.ep, tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(hid.data(), c->A.id, tot * sizeof(uint64_t), hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(off.data(), c->cell_off.p, g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(cn.data(), c->cell_n.p, g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  size_t k = 0;
  for (uint32_t q = 0; q < g.ncells; q++) {
    int i, j, kk;
    cell_ijk(g, q, i, j, kk);
    if (!include_ghosts && is_ghost(g, i, j, kk)) continue;
    for (uint32_t p = 0; p < cn[q]; p++, k++) {
      const size_t s = (size_t)off[q] + p;
      if (cell) cell[k] = q;
      if (r) {
        r[3 * k] = hr[s].x;
        r[3 * k + 1] = hr[s].y;
        r[3 * k + 2] = hr[s].z;
      }
      for (int a = 0; a < 3; a++) {
        if (v) v[3 * k + a] = hv[a][s];
        if (f) f[3 * k + a] = hf[a][s];
      }
      if (ep) ep[k] = he[s];
      if (id) id[k] = hid[s];
      if (type) type[k] = (uint8_t)hr[s].w;
    }
  }
  return XNB_OK;
}

int xnb_move_particles(xnb_ctx* c) {
  int e;
  if ((e = check_ctx(c))) return e;
  return op_move(c);
}

int xnb_backup_r(xnb_ctx* c) {
  int e;
  if ((e = check_ctx(c))) return e;
  return op_backup(c);
}

int xnb_rebuild_amr(xnb_ctx* c, double sub_grid_density, int enforced_ordering) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!(sub_grid_density > 0.)) return fail(c, XNB_E_ARG, "sub_grid_density must be positive");
  if (enforced_ordering > 1) return fail(c, XNB_E_ARG, "rebuild_amr: only enforced_ordering 1 (one sub-cell level) is implemented");
  if (c->G > 0) return fail(c, XNB_E_STATE, "rebuild_amr runs after move_particles, before ghost_update_all");
  if ((e = op_amr(c, sub_grid_density, true))) return e;
  c->lists_valid = false;
  c->ghosts_valid = false;
  c->backup_valid = false;
  uint32_t flag = 0;
  if ((e = read_u32(c, c->u32s.p, &flag))) return e;
  if (flag & 2u) return fail(c, XNB_E_CAPACITY, "rebuild_amr: a cell holds more than 4096 particles");
  return XNB_OK;
}

int xnb_ghost_update_all(xnb_ctx* c) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!c->nbh_set) return fail(c, XNB_E_STATE, "nbh_dist must run before ghost_update_all");
  return op_ghost_all(c);
}

int xnb_ghost_update_r(xnb_ctx* c) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!c->ghosts_valid) return fail(c, XNB_E_STATE, "ghost_update_r needs a ghost scheme (ghost_update_all)");
  tstart(c, XNB_TIMER_GHOST);
  launch_ghost_update_r(c->g, c->N, c->G, c->A.r, c->ghost_src.p, c->ghost_shift.p, c->st);
  tstop(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_particle_displ_over(xnb_ctx* c, double threshold, uint64_t* count) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!count) return fail(c, XNB_E_ARG, "null output");
  if (!c->backup_valid) return fail(c, XNB_E_STATE, "particle_displ_over needs backup_r");
  HIPCHK(c, hipMemsetAsync(c->u64s.p, 0, sizeof(unsigned long long), c->st));
  tstart(c, XNB_TIMER_TRIGGER);
  launch_displ_over(c->g, c->N, c->A.r, c->slot_cell_A.p, c->backup.p, threshold * threshold, c->u64s.p, c->st);
  tstop(c);
  HIPCHK(c, hipMemcpyAsync(c->h_u64, c->u64s.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  *count = c->h_u64[0];
  return XNB_OK;
}

int xnb_chunk_neighbors(xnb_ctx* c, double nbh_dist_lab, const xnb_chunk_neighbors_config* cfg, int all_cells) {
  int e;
  if ((e = check_ctx(c))) return e;
  xnb_chunk_neighbors_config k = cfg ? *cfg : xnb_chunk_neighbors_config{1, 0, 0, 0, 0};
  if (k.dual_particle_offset) k.build_particle_offset = 1;  // chunk_neighbors_config.h:76-80
  return op_nbh(c, nbh_dist_lab, k, all_cells);
}

int xnb_chunk_neighbors_stream(xnb_ctx* c, uint16_t* out, size_t cap, uint64_t* cell_offsets, size_t* total,
                               uint32_t* max_neighbors) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!c->lists_valid) return fail(c, XNB_E_STATE, "no neighbour lists (chunk_neighbors)");
  if ((e = ensure_full_lists(c)) || (e = ensure_slot_lists(c))) return e;
  const Geom& g = c->g;
  const int ntables = c->list_cfg.build_particle_offset ? (c->list_cfg.dual_particle_offset ? 2 : 1) : 0;
  launch_stream_plen(c->list_atoms, c->slot_cell_A.p, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, c->list_cfg.random_access, c->plen.p,
                     c->st);
  launch_stream_cell_len(g, c->list_atoms, c->cell_off.p, c->cell_n.p, c->plen.p, ntables, c->pstart.p, c->clen.p, c->st);
  if ((e = scan(c, c->clen.p, g.ncells, 0, c->coff.p))) return e;
  uint32_t tot = 0;
  if ((e = read_u32(c, c->u32s.p + 1, &tot))) return e;
  if (total) *total = tot;
  if (max_neighbors) *max_neighbors = c->max_neighbors;
  if (!out) return XNB_OK;
  if (cap < tot) return fail(c, XNB_E_CAPACITY, "stream output buffer too small");
  if ((e = dalloc(c, c->stream_buf, (size_t)tot + 1))) return e;
  launch_stream_fill(g, c->list_atoms, c->slot_cell_A.p, c->cell_off.p, c->cell_n.p, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, c->pstart.p,
                     c->plen.p, c->coff.p, ntables, c->list_cfg.random_access, c->stream_buf.p, c->st);
  HIPCHK(c, hipGetLastError());
  std::vector<uint32_t> co(g.ncells);
  HIPCHK(c, hipMemcpyAsync(out, c->stream_buf.p, (size_t)tot * sizeof(uint16_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(co.data(), c->coff.p, g.ncells * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (cell_offsets) {
    for (uint32_t q = 0; q < g.ncells; q++) cell_offsets[q] = co[q];
    cell_offsets[g.ncells] = tot;
  }
  return XNB_OK;
}

int xnb_zero_particle_force(xnb_ctx* c, int ghost) {
  int e;
  if ((e = check_ctx(c))) return e;
  launch_zero_force(ghost ? c->N + c->G : c->N, c->A, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_divide_force_by_type_scalar(xnb_ctx* c, int ntypes, const double* values, int ghost) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (ntypes < 1 || ntypes > 256 || !values) return fail(c, XNB_E_ARG, "divide_force_by_type_scalar: 1..256 type values required");
  const uint32_t n = ghost ? c->N + c->G : c->N;
  if (n > 0) {  // every particle's type indexes the table
    HIPCHK(c, hipMemsetAsync(c->u32s.p + 9, 0, sizeof(uint32_t), c->st));
    launch_max_type(n, c->A.r, c->u32s.p + 9, c->st);
    HIPCHK(c, hipGetLastError());
    uint32_t tmax = 0;
    if ((e = read_u32(c, c->u32s.p + 9, &tmax))) return e;
    if ((int)tmax >= ntypes) return fail(c, XNB_E_ARG, "divide_force_by_type_scalar: a particle type has no value");
  }
  launch_divide_force_by_type(n, c->A, values, ntypes, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_lennard_jones_force(xnb_ctx* c, const xnb_lj_params* p, double rcut, int precision, int ghost) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!p || (precision != XNB_PRECISION_FP64 && precision != XNB_PRECISION_MIXED)) return fail(c, XNB_E_ARG, "bad LJ arguments");
  if (!c->lists_valid) return fail(c, XNB_E_STATE, "no neighbour lists (chunk_neighbors)");
  if ((e = ensure_full_lists(c))) return e;
  c->last_rcut = rcut;
  tstart(c, XNB_TIMER_FORCE);
  e = launch_force(c, precision, true, false, ghost != 0, p->epsilon, p->sigma, rcut, 0.0);
  tstop(c);
  if (e) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_lennard_jones_force_multimat(xnb_ctx* c, int ntypes, const double* eps, const double* sigma, const double* rcut,
                                     int precision) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (ntypes <= 0 || ntypes > 4 || !eps || !sigma || !rcut) return fail(c, XNB_E_ARG, "bad multimat arguments (1..4 types)");
  if (!c->lists_valid) return fail(c, XNB_E_STATE, "no neighbour lists");
  if ((e = ensure_full_lists(c))) return e;
  LJMulti mm;
  if (make_lj_multi(ntypes, eps, sigma, rcut, mm)) return fail(c, XNB_E_ARG, "too many type pairs");
  tstart(c, XNB_TIMER_FORCE);
  if (c->list_kind == 1) {
    hipError_t he = launch_ljmm_brick(precision, ntypes, EPI_NONE, c->g, c->bp, c->A, c->cell_off.p, c->cell_n.p, c->blist.p, c->bcnt.p,
                                      mm, 0.0, nullptr, c->st);
    if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick multimat kernel: ") + hipGetErrorString(he));
  } else {
    if ((e = ensure_slot_lists(c))) return e;
    if (launch_lj_multimat(precision, c->N, c->A, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, ntypes, mm, c->st))
      return fail(c, XNB_E_ARG, "too many type pairs");
  }
  tstop(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_gravitational_force(xnb_ctx* c, double G, int ntypes, const double* type_mass, double rcut, int ghost) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (ntypes < 1 || ntypes > kGravityTypes || !type_mass || !(rcut > 0.0))
    return fail(c, XNB_E_ARG, "bad gravitational_force arguments (1..16 types, rcut > 0)");
  if (!c->lists_valid) return fail(c, XNB_E_STATE, "no neighbour lists (chunk_neighbors)");
  if ((e = ensure_full_lists(c))) return e;
  const uint32_t n_atoms = ghost ? c->N + c->G : c->N;
  // every type indexes the mass table (the reference reads m_type_mass[type] unchecked)
  HIPCHK(c, hipMemsetAsync(c->u32s.p + 9, 0, sizeof(uint32_t), c->st));
  launch_max_type(n_atoms, c->A.r, c->u32s.p + 9, c->st);
  HIPCHK(c, hipGetLastError());
  uint32_t tmax = 0;
  if ((e = read_u32(c, c->u32s.p + 9, &tmax))) return e;
  if ((int)tmax >= ntypes) return fail(c, XNB_E_ARG, "a particle type has no mass");
  GravitationalForceFunctor gf{};
  gf.G = G;
  for (int q = 0; q < ntypes; q++) gf.mass[q] = type_mass[q];
  const double rc2 = rcut * rcut;
  tstart(c, XNB_TIMER_FORCE);
  if (c->list_kind == 1 && !ghost) {
    hipError_t he = launch_gravity_brick(c->g, c->bp, c->A, c->cell_off.p, c->cell_n.p, c->blist.p, c->bcnt.p, gf, rc2, c->st);
    if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("brick gravitational kernel: ") + hipGetErrorString(he));
  } else {
    if ((e = ensure_slot_lists(c))) return e;
    if (c->list_atoms < n_atoms) return fail(c, XNB_E_STATE, "neighbour lists do not cover the requested cells");
    launch_gravity_slots(n_atoms, c->A, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, gf, rc2, c->st);
  }
  tstop(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_push_f_v_r(xnb_ctx* c, double dt, double dt_scale) {
  int e;
  if ((e = check_ctx(c))) return e;
  const double delta_t = dt * dt_scale;
  const double delta_t2 = delta_t * delta_t * 0.5;
  tstart(c, XNB_TIMER_PUSH);
  launch_push_f_v_r(c->N, c->A, delta_t, delta_t2, c->st);
  tstop(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_push_f_v(xnb_ctx* c, double dt, double dt_scale) {
  int e;
  if ((e = check_ctx(c))) return e;
  tstart(c, XNB_TIMER_PUSH);
  launch_push_f_v(c->N, c->A, dt * dt_scale, c->st);
  tstop(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_simulation_energies(xnb_ctx* c, double* ke, double* pe) {
  int e;
  if ((e = check_ctx(c))) return e;
  const uint32_t nparts = std::max<uint32_t>(1, std::min<uint32_t>(1024, nb256(c->N)));
  if ((e = dalloc(c, c->red_part, 2 * (size_t)nparts))) return e;
  launch_energies(c->N, c->A, c->red_part.p, nparts, c->red_out.p, c->st);
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, c->red_out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (ke) *ke = h[0];
  if (pe) *pe = h[1];
  return XNB_OK;
}

}  // extern "C"

// ---- energies without allocation: the reduction partials live in a fixed device array
namespace {
constexpr uint32_t kEnergyParts = 1024;  // (ke, pe) partial pairs: one per reduction workgroup
// Each GPU the library runs on holds its own copy (16 KiB) until the process ends.  Host code
// never uses the name: kernels get the address hipGetSymbolAddress gives per device.
static __device__ __align__(256) double scratch[2 * kEnergyParts];
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

// the device address of `scratch` on the current device (c->device), looked up on the first
// call on that device
int scratch_for(xnb_ctx* c, double** out) {
  if (c->device < 0 || c->device >= kMaxDevices) return fail(c, XNB_E_ARG, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[c->device]) {
    void* ptr = nullptr;
    const hipError_t he = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (he != hipSuccess || !ptr) return fail(c, XNB_E_HIP, std::string("scratch address: ") + hipGetErrorString(he));
    scratch_dev[c->device] = (double*)ptr;
  }
  *out = scratch_dev[c->device];
  return XNB_OK;
}
}  // namespace

extern "C" {

int xnb_simulation_energies_noalloc(xnb_ctx* c, double* ke_pe) {
  int e;
  if ((e = check_ctx(c))) return e;  // makes the context's device current
  if (!ke_pe) return fail(c, XNB_E_ARG, "null output");
  // the same partial count as xnb_simulation_energies, so the sums are the same bit for bit;
  // the reduction strides over the atoms, so every particle count fits kEnergyParts partials
  const uint32_t nparts = std::max<uint32_t>(1, std::min<uint32_t>(1024, nb256(c->N)));
  if (nparts > kEnergyParts) return fail(c, XNB_E_CAPACITY, "more reduction partials than the scratch array holds");
  double* part = nullptr;
  if ((e = scratch_for(c, &part))) return e;
  // every partial read by the final sum is written first by this call: nothing to clear
  launch_energies(c->N, c->A, part, nparts, ke_pe, c->st);
  HIPCHK(c, hipGetLastError());
  return XNB_OK;
}

int xnb_init(xnb_ctx* c, const xnb_lj_params* p, double rcut, double rcut_inc, int precision) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!p) return fail(c, XNB_E_ARG, "null LJ parameters");
  if ((e = xnb_nbh_dist(c, rcut, rcut_inc, nullptr, nullptr, nullptr))) return e;
  if ((e = full_update(c))) return e;
  c->last_rcut = rcut;
  tstart(c, XNB_TIMER_FORCE);
  e = launch_force(c, precision, false, false, false, p->epsilon, p->sigma, rcut, 0.0);
  tstop(c);
  if (e) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

// Velocity-Verlet loop (numerical-scheme.msp: push_f_v_r, push_f_v, check_and_update,
// compute_force, push_f_v).  When the lists are brick lists, the force kernel of step s < steps
// also performs step s+1's first half and displacement count, writing positions to buffer B
// whose position array is then swapped with A's; the arithmetic is the same as the separate
// push kernel's.  Forces and per-atom energies are stored on the last step and on steps
// followed by a rebuild; the state at return equals the unfused loop's bit for bit.
// Displacement counts go to step_counts[step - 1].
}  // extern "C"

namespace xnb_impl {

// The step loop of xnb_run / xnb_run_multimat (mm == nullptr: single-material LJ).
int run_loop(xnb_ctx* c, xnb_run_params* p, const LJMulti* mm, int ntypes, const double* rcut_pair) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!p || p->steps < 0) return fail(c, XNB_E_ARG, "bad run parameters");
  if (!c->lists_valid || !c->ghosts_valid || !c->backup_valid) return fail(c, XNB_E_STATE, "xnb_init must run first");
  // A multi-material run rebuilds its lists with each type pair's radius rc_ij + 2 max_displ
  // (at most nbh_dist): pairs beyond rc_ij add +-0 to the sums, so the forces are bitwise those
  // of full lists, and the skin argument holds pair by pair.
  PairCut pc{};
  if (mm) {
    pc.nt = ntypes;
    for (int a = 0; a < ntypes; a++)
      for (int b = 0; b < ntypes; b++) {
        const double rl = std::min(c->nbh_dist_lab, rcut_pair[unique_pair_id((unsigned)a, (unsigned)b)] + 2.0 * c->max_displ);
        pc.thr2[a][b] = rl * rl;
      }
  }
  // lists pruned by another run's radii: start from full lists
  if (c->lists_pruned && (!mm || std::memcmp(&pc, &c->pc_used, sizeof(PairCut)) != 0))
    if ((e = ensure_full_lists(c))) return e;
  c->pc_used = pc;
  const double dt = p->dt * 1.0, dt2 = dt * dt * 0.5, dth = p->dt * 0.5;
  const double thr2 = c->max_displ * c->max_displ;
  const bool interval = p->rebuild_every > 0;
  c->last_rcut = p->rcut;
  p->n_rebuilds = 0;
  p->n_violations = 0;
  c->last_run_steps = 0;
  if (p->steps > 0) {
    if ((e = dalloc(c, c->step_counts, (size_t)p->steps))) return e;
    HIPCHK(c, hipMemsetAsync(c->step_counts.p, 0, (size_t)p->steps * sizeof(unsigned long long), c->st));
  }
  std::vector<char> rebuilt(interval ? (size_t)p->steps : 0, 0);
  bool first_half_done = false;  // step's first half already done by the previous force kernel
  for (int64_t step = 1; step <= p->steps; step++) {
    unsigned long long* cnt = c->step_counts.p + (step - 1);
    if (!first_half_done) {
      tstart(c, XNB_TIMER_PUSH);
      launch_verlet_first_half(c->g, c->N, c->A, dt, dt2, dth, c->slot_cell_A.p, c->backup.p, thr2, cnt, c->st);
      tstop(c);
    }
    bool trigger;
    if (interval) {
      trigger = (step % p->rebuild_every) == 0;
    } else {
      HIPCHK(c, hipMemcpyAsync(c->h_u64, cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      trigger = c->h_u64[0] > 0;
    }
    if (trigger) {
      if ((e = full_update(c, false, mm ? &pc : nullptr))) return e;  // f, ep: overwritten by this step's force kernel
      p->n_rebuilds++;
      if (interval) rebuilt[(size_t)step - 1] = 1;
    } else {
      tstart(c, XNB_TIMER_GHOST);
      launch_ghost_update_r(c->g, c->N, c->G, c->A.r, c->ghost_src.p, c->ghost_shift.p, c->st);
      tstop(c);
    }
    const bool fuse = c->fuse_push && step < p->steps && c->list_kind == 1;
    // A fused step (EPI_NEXT) stores no forces or energies: the next step overwrites them before
    // anything reads them (its force kernel sets f = 0 + sum for every inner atom; a rebuild in
    // between does not carry them, and the first half-step of the next step is done by this
    // kernel from the force in registers).  The last step of a run is not fused and stores both.
    NextStep nx{c->rB.p, c->backup.p, dt, dt2, thr2, cnt + 1};
    tstart(c, XNB_TIMER_FORCE);
    if (mm) e = launch_force_multimat(c, p->precision, *mm, ntypes, dth, fuse ? &nx : nullptr);
    else e = launch_force(c, p->precision, false, true, false, p->lj.epsilon, p->lj.sigma, p->rcut, dth, fuse ? &nx : nullptr);
    tstop(c);
    if (e) return e;
    if (fuse) {
      std::swap(c->rA, c->rB);
      bind_states(c);
    }
    first_half_done = fuse;
  }
  HIPCHK(c, hipGetLastError());
  if (interval && p->steps > 0) {
    std::vector<unsigned long long> h((size_t)p->steps);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->step_counts.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t q = 0; q < h.size(); q++)
      if (h[q] > 0 && !rebuilt[q]) p->n_violations++;
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->last_run_steps = p->steps;
  return XNB_OK;
}

}  // namespace xnb_impl

extern "C" {

int xnb_run(xnb_ctx* c, xnb_run_params* p) { return run_loop(c, p, nullptr, 0, nullptr); }

int xnb_run_step_counts(xnb_ctx* c, uint64_t* counts, size_t n) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (n > (size_t)c->last_run_steps || (n && !counts)) return fail(c, XNB_E_ARG, "more counts than steps of the last run");
  if (!n) return XNB_OK;
  HIPCHK(c, hipMemcpyAsync(counts, c->step_counts.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

int xnb_run_multimat(xnb_ctx* c, xnb_run_params* p, int ntypes, const double* eps, const double* sigma, const double* rcut) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (ntypes <= 0 || ntypes > 4 || !eps || !sigma || !rcut) return fail(c, XNB_E_ARG, "bad multimat arguments (1..4 types)");
  LJMulti mm;
  if (make_lj_multi(ntypes, eps, sigma, rcut, mm)) return fail(c, XNB_E_ARG, "too many type pairs");
  return run_loop(c, p, &mm, ntypes, rcut);
}

// ---- autotuner: brick shape (cells per workgroup tile) by measurement, cached per class
namespace {
struct TuneKey {
  int di, dj, dk, occ, precision;
  bool operator<(const TuneKey& o) const {
    if (di != o.di) return di < o.di;
    if (dj != o.dj) return dj < o.dj;
    if (dk != o.dk) return dk < o.dk;
    if (occ != o.occ) return occ < o.occ;
    return precision < o.precision;
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 3>> g_tune;
}  // namespace

int xnb_autotune(xnb_ctx* c, const xnb_lj_params* p, double rcut, int precision, int* dims_out) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!p) return fail(c, XNB_E_ARG, "null parameters");
  if (!c->lists_valid || !c->ghosts_valid) return fail(c, XNB_E_STATE, "xnb_init must run first");
  if ((e = ensure_full_lists(c))) return e;
  if (c->list_kind != 1) {  // not on the brick path: nothing to tune
    if (dims_out) dims_out[0] = dims_out[1] = dims_out[2] = 0;
    return XNB_OK;
  }
  const Geom& g = c->g;
  // class: cells per side (large boxes alike), mean atoms per inner cell, precision
  const double occ = (double)c->N / ((double)g.Di * g.Dj * g.Dk);
  const TuneKey key{std::min(g.Di, 16), std::min(g.Dj, 16), std::min(g.Dk, 16), (int)std::lround(occ), precision};
  std::array<int, 3> best{0, 0, 0};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      best = it->second;
      cached = true;
    }
  }
  const int saved[3] = {c->brick_dims[0], c->brick_dims[1], c->brick_dims[2]};
  xnb_chunk_neighbors_config cfg{1, 0, 0, 0, 0};
  if (!cached) {
    static const int cand[][3] = {{4, 4, 2}, {4, 2, 2}, {2, 4, 2}, {4, 4, 1}, {2, 2, 2}, {4, 2, 1}};
    struct Events {  // destroyed on every return path
      hipEvent_t a = nullptr, b = nullptr;
      ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
      }
    } ev;
    struct Dims {  // the caller's shape is back on every exit from the trials (the winner is set below)
      xnb_ctx* c;
      int d[3];
      ~Dims() { std::copy(d, d + 3, c->brick_dims); }
    } restore{c, {saved[0], saved[1], saved[2]}};
    HIPCHK(c, hipEventCreate(&ev.a));
    HIPCHK(c, hipEventCreate(&ev.b));
    hipEvent_t t0 = ev.a, t1 = ev.b;
    float best_ms = 1e30f;
    for (const auto& d : cand) {
      c->brick_dims[0] = d[0];
      c->brick_dims[1] = d[1];
      c->brick_dims[2] = d[2];
      if ((e = op_nbh(c, c->nbh_dist_lab, cfg, 0))) break;
      if (c->list_kind != 1 || c->bp.Bi != d[0] || c->bp.Bj != d[1] || c->bp.Bk != d[2]) continue;  // did not fit
      // forces are rewritten with the values they had (the shape does not change results)
      if ((e = launch_force(c, precision, false, false, false, p->epsilon, p->sigma, rcut, 0.0))) break;
      HIPCHK(c, hipEventRecord(t0, c->st));
      for (int q = 0; q < 3 && !e; q++) e = launch_force(c, precision, false, false, false, p->epsilon, p->sigma, rcut, 0.0);
      HIPCHK(c, hipEventRecord(t1, c->st));
      HIPCHK(c, hipEventSynchronize(t1));
      if (e) break;
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, t0, t1));
      if (ms < best_ms) {
        best_ms = ms;
        best = {d[0], d[1], d[2]};
      }
    }
    if (e) return e;
    if (best[0] > 0) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = best;
    }
  }
  if (best[0] > 0) {
    c->brick_dims[0] = best[0];
    c->brick_dims[1] = best[1];
    c->brick_dims[2] = best[2];
  } else {
    c->brick_dims[0] = saved[0];
    c->brick_dims[1] = saved[1];
    c->brick_dims[2] = saved[2];
  }
  if ((e = op_nbh(c, c->nbh_dist_lab, cfg, 0))) return e;
  if ((e = launch_force(c, precision, false, false, false, p->epsilon, p->sigma, rcut, 0.0))) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (dims_out) {
    dims_out[0] = c->list_kind == 1 ? c->bp.Bi : 0;
    dims_out[1] = c->list_kind == 1 ? c->bp.Bj : 0;
    dims_out[2] = c->list_kind == 1 ? c->bp.Bk : 0;
  }
  return XNB_OK;
}

int xnb_timer_enable(xnb_ctx* c, int enable) {
  if (!c) return XNB_E_ARG;
  c->timing = enable != 0;
  return XNB_OK;
}

int xnb_timer_reset(xnb_ctx* c) {
  if (!c) return XNB_E_ARG;
  flush_timers(c);
  for (int q = 0; q < XNB_TIMER_COUNT; q++) {
    c->t_ms[q] = 0.;
    c->t_n[q] = 0;
  }
  return XNB_OK;
}

int xnb_timer_read(xnb_ctx* c, int timer, double* total_ms, int64_t* launches) {
  if (!c || timer < 0 || timer >= XNB_TIMER_COUNT) return XNB_E_ARG;
  flush_timers(c);
  if (total_ms) *total_ms = c->t_ms[timer];
  if (launches) *launches = c->t_n[timer];
  return XNB_OK;
}

int xnb_pair_counts(xnb_ctx* c, uint64_t* list_entries, uint64_t* cutoff_pairs) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (!c->lists_valid) return fail(c, XNB_E_STATE, "no neighbour lists");
  if ((e = ensure_full_lists(c)) || (e = ensure_slot_lists(c))) return e;
  HIPCHK(c, hipMemsetAsync(c->u64s.p + 1, 0, 2 * sizeof(unsigned long long), c->st));
  launch_pair_counts(c->N, c->A.r, c->nbr.p, c->nbr_cnt.p, c->nbr_cap, c->last_rcut * c->last_rcut, c->u64s.p + 1, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_u64, c->u64s.p + 1, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (list_entries) *list_entries = c->h_u64[0];
  if (cutoff_pairs) *cutoff_pairs = c->h_u64[1];
  return XNB_OK;
}

int xnb_copy_probe(xnb_ctx* c, size_t bytes, int iters, double* gbps) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (bytes < 16 || iters <= 0 || !gbps) return fail(c, XNB_E_ARG, "bad copy probe arguments");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // every step's status folds into he; the buffers and events are released on all paths
  hipError_t he = hipMalloc(&a, bytes);
  if (he == hipSuccess) he = hipMalloc(&b, bytes);
  if (he == hipSuccess) he = hipMemsetAsync(a, 1, bytes, c->st);
  if (he == hipSuccess) he = hipEventCreate(&t0);
  if (he == hipSuccess) he = hipEventCreate(&t1);
  float ms = 0.f;
  if (he == hipSuccess) {
    launch_copy_probe(a, b, bytes, c->st);  // warm-up
    he = hipEventRecord(t0, c->st);
    for (int q = 0; q < iters && he == hipSuccess; q++) launch_copy_probe(a, b, bytes, c->st);
    if (he == hipSuccess) he = hipGetLastError();
    if (he == hipSuccess) he = hipEventRecord(t1, c->st);
    if (he == hipSuccess) he = hipEventSynchronize(t1);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, t0, t1);
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("copy probe: ") + hipGetErrorString(he));
  *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  return XNB_OK;
}

int xnb_fp64_core(xnb_ctx* c, const double* x, size_t n, double* sqrt_out, double* recip_out) {
  int e;
  if ((e = check_ctx(c))) return e;
  if (n == 0) return XNB_OK;
  if (!x || !sqrt_out || !recip_out) return fail(c, XNB_E_ARG, "null argument");
  double* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, 3 * n * sizeof(double)));
  hipError_t he = hipMemcpyAsync(d, x, n * sizeof(double), hipMemcpyHostToDevice, c->st);
  if (he == hipSuccess) {
    launch_fp64_core(d, n, d + n, d + 2 * n, c->st);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipMemcpyAsync(sqrt_out, d + n, n * sizeof(double), hipMemcpyDeviceToHost, c->st);
  if (he == hipSuccess) he = hipMemcpyAsync(recip_out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, c->st);
  if (he == hipSuccess) he = hipStreamSynchronize(c->st);
  if (d) (void)hipFree(d);
  if (he != hipSuccess) return fail(c, XNB_E_HIP, std::string("fp64 core: ") + hipGetErrorString(he));
  return XNB_OK;
}

int xnb_device_bytes(const xnb_ctx* c, size_t* current, size_t* peak) {
  if (!c) return XNB_E_ARG;
  if (current) *current = c->dev_bytes;
  if (peak) *peak = c->dev_peak;
  return XNB_OK;
}

int xnb_synchronize(xnb_ctx* c) {
  int e;
  if ((e = check_ctx(c))) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return XNB_OK;
}

}  // extern "C"
