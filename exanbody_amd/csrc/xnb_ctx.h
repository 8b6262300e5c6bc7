// xnb_ctx.h -- host-side state of one box (an xnb_ctx) and the internal helpers shared by
// the single-box C ABI (xnb_api.hip) and the batched engine (xnb_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/xnb_amd.h"
#include "xnb_kernels.h"

using namespace xnb;

namespace xnb_impl {

template <class T>
struct DArr {
  T* p = nullptr;
  size_t n = 0;
};

struct TimerRec {
  int id;
  hipEvent_t a, b;
};

}  // namespace xnb_impl

using xnb_impl::DArr;
using xnb_impl::TimerRec;

struct xnb_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  Geom g{};
  xnb_domain dom{};
  bool nbh_set = false;
  double nbh_dist_lab = 0., max_displ = 0., ghost_dist = 0.;

  uint32_t N = 0, G = 0;
  size_t cap = 0;  // particle slots allocated in each buffer
  PState A{}, B{};
  DArr<uint32_t> slot_cell_A, slot_cell_B;
  DArr<double4> rA, rB;
  DArr<double> vA[3], vB[3], fA[3], fB[3], eA, eB;
  DArr<uint64_t> idA, idB;
  int n_extra = 0;  // extra per-particle fields (xnb_set_particles_with_fields)
  DArr<double> xA[kMaxExtraFields], xB[kMaxExtraFields];

  // per cell
  DArr<uint32_t> cell_off, cell_n, cnt_out, cnt_in, new_n, new_off, in_cur, gcnt, goff, clen, coff;
  // per slot scratch
  DArr<uint32_t> mv_dst, pk, newloc, in_list, plen, pstart;
  DArr<uint32_t> ghost_src, ghost_shift, backup;
  // neighbour lists
  DArr<uint32_t> nbr, nbr_cnt;
  uint32_t nbr_cap = 96;
  uint32_t list_atoms = 0;
  uint32_t max_neighbors = 0;
  xnb_chunk_neighbors_config list_cfg{1, 0, 0, 0, 0};
  bool lists_valid = false, ghosts_valid = false, backup_valid = false;
  bool lists_pruned = false;  // built with per type-pair radii inside xnb_run_multimat
  PairCut pc_used{};          // the radii of the last run (pc_used.nt = 0: single material)
  DArr<uint16_t> stream_buf;

  // scan / scalar scratch
  DArr<uint32_t> bsum, u32s;            // u32s: [0] err, [1] scan total, [2] max count, [3..4] brick stats, [6] brick status,
                                        // [7] ghost status, [8] rebuild_amr cell count, [9] largest particle type
  DArr<uint32_t> amr_list;              // rebuild_amr: cells to sort
  DArr<unsigned long long> u64s;        // [0] displacement count, [1..2] pair counts
  DArr<unsigned long long> step_counts; // per-step displacement counts of xnb_run
  int64_t last_run_steps = 0;           // steps of the last xnb_run / xnb_run_multimat
  DArr<double> red_part, red_out;
  uint32_t* h_u32 = nullptr;            // pinned
  unsigned long long* h_u64 = nullptr;  // pinned

  // timers
  bool timing = false;
  std::vector<TimerRec> pending;
  std::vector<hipEvent_t> pool;
  double t_ms[XNB_TIMER_COUNT] = {};
  int64_t t_n[XNB_TIMER_COUNT] = {};
  int t_cur = -1;
  hipEvent_t t_a = nullptr;

  // brick-tiled production lists (xnb_brick.hip)
  int list_kind = 0;  // 0: generic per-slot u32 lists in nbr; 1: brick u16 lists in blist
  bool use_bricks = true;
  BrickParams bp{};
  int brick_dims[3] = {0, 0, 0};  // 0: choose automatically
  DArr<uint16_t> blist;
  DArr<uint32_t> bcnt;
  bool slot_lists_current = false;  // nbr/nbr_cnt mirror the brick lists
  bool fuse_push = true;            // force kernel performs the next step's first half (brick lists)
  bool nbh_prefilter = true;        // unfiltered brick builds test in FP32 with an FP64 band (same lists)
  uint32_t lds_limit = 65536;
  bool batch_shapes = false;  // brick shapes offered to batched boxes only
  double brick_margin = 1.0;        // brick T / Rmax head room (batched boxes rebuild without re-choosing)

  double last_rcut = 0.;
  size_t dev_bytes = 0, dev_peak = 0;  // device bytes the context holds now / at most so far
  std::string err;
};


namespace xnb_impl {

int fail(xnb_ctx* c, int code, const std::string& msg);

#define HIPCHK(ctx, expr)                                                                            \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess)                                                                            \
      return ::xnb_impl::fail((ctx), XNB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

// Device arrays of a context: every allocation and free is counted in c->dev_bytes (and its
// running maximum c->dev_peak), the memory the serving loop's per-job estimate must bound.
template <class T>
void dfree(xnb_ctx* c, DArr<T>& a) {
  if (a.p) {
    (void)hipFree(a.p);  // hipFree waits for work still using the buffer
    c->dev_bytes -= a.n * sizeof(T);
  }
  a.p = nullptr;
  a.n = 0;
}

template <class T>
int dalloc(xnb_ctx* c, DArr<T>& a, size_t n) {
  if (a.n >= n && a.p) return XNB_OK;
  dfree(c, a);
  size_t nn = n ? n : 1;
  hipError_t e = hipMalloc((void**)&a.p, nn * sizeof(T));
  if (e != hipSuccess) {
    a.p = nullptr;
    return fail(c, XNB_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  a.n = nn;
  c->dev_bytes += nn * sizeof(T);
  if (c->dev_bytes > c->dev_peak) c->dev_peak = c->dev_bytes;
  return XNB_OK;
}


void bind_states(xnb_ctx* c);
void swap_buffers(xnb_ctx* c);
int ensure_capacity(xnb_ctx* c, size_t ncap, size_t keep);
int check_ctx(xnb_ctx* c);
int op_ghost_all(xnb_ctx* c, bool move_check = false, bool forces = true);
int op_backup(xnb_ctx* c);
int op_amr(xnb_ctx* c, double density, bool forces);
int op_nbh(xnb_ctx* c, double nbh_dist, const xnb_chunk_neighbors_config& cfg, int all_cells, const PairCut* pc = nullptr);
// forces false: f and ep neither carried by the move nor copied to ghosts; pc: per type-pair list radii
int full_update(xnb_ctx* c, bool forces = true, const PairCut* pc = nullptr);
int ensure_full_lists(xnb_ctx* c);
void tstart(xnb_ctx* c, int id);
void tstop(xnb_ctx* c);

}  // namespace xnb_impl
