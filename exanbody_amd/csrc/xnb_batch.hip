// xnb_batch.hip -- many independent boxes advanced together: every kernel of a step (first
// half, rebuild, ghost refresh, fused force) is one launch over all boxes of the batch.
//
// A box is an initialised xnb_ctx (xnb_create, xnb_set_particles, xnb_nbh_dist, xnb_init)
// attached with its run parameters; it keeps its own device arrays, which the batch addresses
// through a device table of BoxFrame entries.  Each box follows exactly the operator sequence
// of xnb_run (same kernels' per-atom bodies, same order of operations per atom), so a box
// advanced in a batch ends bitwise equal to the same box advanced alone.  Boxes may be
// attached and detached between steps (continuous batching); their step counters are their
// own.
//
// Host synchronisation: none on a step without rebuilds; one read-back of four status words
// per box at the end of a step where some box rebuilds (ghost counts for the next work
// tables, overflow flags).  A box whose rebuild overflowed the sizes its frame was laid out
// for (ghost slots, brick workgroup or region, list capacity) is left untouched by that step's
// force kernel, then redone with the one-box operators, which re-size its arrays, and given
// its force in a launch of its own.  A step with no box starting or rebuilt on displacement
// is one launch sequence (rebuilds, ghost refresh, force, clock), replayed from a captured
// HIP graph once the same sequence has come twice (quiet steps and rebuild steps each have
// their own graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xnb_ctx.h"

using namespace xnb;
using namespace xnb_impl;

namespace {

struct BatchBox {
  xnb_ctx* c = nullptr;
  bool used = false, finished = false, failed = false;
  BoxFrame f{};  // host mirror of the device frame
  int64_t steps = 0, done = 0;
  int every = 10;         // rebuild every this many steps; 0: when an atom passed max_displ
  bool fresh = true;      // the next step's first half still has to be launched
  uint32_t last_par = 0;  // parity of the box's last step
  uint32_t ghost_span = 0;  // ghost slots covered by its ghost-refresh blocks (>= G)
  int64_t n_rebuilds = 0, n_violations = 0, n_recover = 0;
  std::vector<char> rebuilt;
  DArr<unsigned long long> counts;
  xnb_run_params p{};
  std::string err;
};

// brick T / Rmax head room of attached boxes; under the batch's LDS cap it also decides
// which brick shape a box gets.  512-box batch, 300 steps, atom-steps/s: 1.06 3.83e9, 1.12
// 3.92e9, 1.15 4.27e9, 1.18 4.21e9, 1.21 3.94e9, 1.35 3.88e9, 1.5 3.45e9; 1.0 recovers
// thousands of times.
constexpr double kBrickMargin = 1.15;
constexpr double kSlotMargin = 1.3;    // slot capacity head room for ghosts

// One launch of brick work: its table range, workgroup lanes and dynamic LDS.
struct BrickLaunch {
  int kind = kBatchNbh;  // which batched brick kernel
  size_t o = 0, n = 0;
  uint32_t T = 64, lds = 0;
  bool operator==(const BrickLaunch& b) const { return kind == b.kind && o == b.o && n == b.n && T == b.T && lds == b.lds; }
};

// The launches of one step after its first half, as ranges of the step's work table.
struct StepLayout {
  uint32_t par = 0;   // parity of a rebuild step (its rebuild kernels take it as an argument)
  uint32_t nreb = 0;  // rebuilding boxes, listed at o_list
  size_t o_rs = 0, n_rs = 0, o_rc = 0, n_rc = 0, o_rg = 0, n_rg = 0, o_list = 0, o_zero = 0;
  std::vector<BrickLaunch> lb;  // neighbour builds of the rebuilding boxes
  size_t o_g = 0, n_g = 0;      // ghost refresh of the other boxes
  std::vector<BrickLaunch> lf;  // force of every box
  bool operator==(const StepLayout& x) const {
    return par == x.par && nreb == x.nreb && o_rs == x.o_rs && n_rs == x.n_rs && o_rc == x.o_rc && n_rc == x.n_rc &&
           o_rg == x.o_rg && n_rg == x.n_rg && o_list == x.o_list && o_zero == x.o_zero && lb == x.lb && o_g == x.o_g &&
           n_g == x.n_g && lf == x.lf;
  }
};

// A captured step: its graph with the table and layout it was captured from (the table in
// buf[cur]), and the table and layout of the last step of its kind.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<uint32_t> tab, last_tab;
  StepLayout lay, last_lay;
  bool have_last = false;
  DArr<uint32_t> buf[2];
  int cur = 0;
};

}  // namespace

struct xnb_batch {
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<BatchBox> box;
  DArr<BoxFrame> dF;
  BoxFrame* hF = nullptr;  // pinned staging of frame uploads, one entry per slot
  DArr<uint32_t> dwords;   // 4 words per slot
  uint32_t* hwords = nullptr;
  DArr<uint32_t> dtrig;    // displacement trigger per slot for the step about to run
  uint32_t* htrig = nullptr;
  DArr<uint32_t> dwork;
  uint32_t* hwork[2] = {nullptr, nullptr};
  size_t hwork_cap = 0;
  hipEvent_t hev[2] = {nullptr, nullptr};
  int hcur = 0;
  uint32_t par = 0;
  uint32_t gstep = 1;  // global step number of the next step
  DArr<uint32_t> dclock;  // {par, gstep} of the step being run, then of the previous step
  DArr<uint32_t> amr_list, amr_count;  // rebuild_amr: cells of rebuilding boxes to sort, their number (+3 words)
  uint32_t lds_limit = 65536;
  BatchKernelInfo kinfo[3] = {};  // per batched brick kernel (kBatchLjFp64, kBatchLjMixed, kBatchNbh)
  // steps replayed from captured HIP graphs (XNB_BATCH_GRAPH=0: every step launched eagerly)
  bool use_graph = true;
  // head room of attached boxes (XNB_BATCH_MARGINS="brick,slot" at create; 1,1 leaves none, so
  // most rebuilds that grow a brick or the ghosts take the one-box recovery)
  double brick_margin = kBrickMargin, slot_margin = kSlotMargin;
  // LDS bytes a box's brick workgroup may take: half the CU's by default, so that two brick
  // workgroups share a CU (512-box batch, 100 steps: 3.15e9 -> 3.54e9 atom-steps/s against no
  // cap; a third of the CU is slower again); XNB_BATCH_LDS="fraction" at create
  uint32_t brick_lds_cap = 0;
  double lds_frac = 0.5;  // brick_lds_cap as a fraction of the CU's LDS (xnb_batch_autotune may change it)
  // brick work launched per class of boxes (XNB_BATCH_CLASSES=0: one launch sized for the
  // widest box)
  bool use_classes = true;
  // a step's force launches (disjoint boxes) alternate between st and st2, forked from and
  // joined back into st, so one class's tail overlaps another's work (XNB_BATCH_STREAMS=0:
  // all on st); both streams at the default priority
  bool split_streams = true;
  hipStream_t st2 = nullptr;
  hipEvent_t fork_ev[2] = {nullptr, nullptr}, join_ev[2] = {nullptr, nullptr};  // neighbour build, force
  // graphs of quiet steps (no box starting, rebuilding or rebuilt on displacement) and of
  // rebuild steps (fixed-interval rebuilds only), one per step parity, which the rebuild
  // kernels take as an argument (an odd interval alternates it)
  StepGraph gq, gr[2];
  std::vector<hipGraphExec_t> retired;  // destroyed after the next stream synchronisation
  int64_t n_graph_steps = 0, n_captures = 0;
  std::string err;
};

namespace {

int bfail(xnb_batch* B, int code, const std::string& msg) {
  if (B) B->err = msg;
  return code;
}

#define BHIP(B, expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) return bfail((B), XNB_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ---- frames <-> contexts

void frame_from_ctx(xnb_batch* B, int i, const xnb_run_params& p) {
  BatchBox& bb = B->box[i];
  xnb_ctx* c = bb.c;
  BoxFrame f{};
  f.g = c->g;
  f.bp = c->bp;
  f.bp.status = B->dwords.p + 4 * (size_t)i;
  f.bpn = f.bp;
  brick_layout(f.bpn, true);
  f.N = c->N;
  f.G = c->G;
  f.cap = (uint32_t)std::min<size_t>(c->cap, 0xffffffffu);
  f.ab = 0;
  f.rcur = B->par;  // current positions: rbuf[rcur ^ par] = rbuf[0] = the context's rA
  f.mixed = p.precision == XNB_PRECISION_MIXED ? 1u : 0u;
  f.A = c->A;
  f.B = c->B;
  f.rbuf[0] = c->rA.p;
  f.rbuf[1] = c->rB.p;
  f.slot_cell[0] = c->slot_cell_A.p;
  f.slot_cell[1] = c->slot_cell_B.p;
  f.cell_off = c->cell_off.p;
  f.cell_n = c->cell_n.p;
  f.cnt_out = c->cnt_out.p;
  f.cnt_in = c->cnt_in.p;
  f.new_n = c->new_n.p;
  f.new_off = c->new_off.p;
  f.in_cur = c->in_cur.p;
  f.gcnt = c->gcnt.p;
  f.goff = c->goff.p;
  f.mv_dst = c->mv_dst.p;
  f.pk = c->pk.p;
  f.newloc = c->newloc.p;
  f.in_list = c->in_list.p;
  f.ghost_src = c->ghost_src.p;
  f.ghost_shift = c->ghost_shift.p;
  f.backup = c->backup.p;
  f.blist = c->blist.p;
  f.bcnt = c->bcnt.p;
  f.lj = make_lj_const(p.lj.epsilon, p.lj.sigma, p.rcut);
  f.dt = p.dt * 1.0;  // the same expressions as xnb_run
  f.dt2 = f.dt * f.dt * 0.5;
  f.dth = p.dt * 0.5;
  f.thr2 = c->max_displ * c->max_displ;
  f.max2 = c->nbh_dist_lab * c->nbh_dist_lab;
  f.words = B->dwords.p + 4 * (size_t)i;
  f.counts = bb.counts.p;
  f.gstep0 = B->gstep - (uint32_t)bb.done;  // local step k runs at global step gstep0 + k - 1
  bb.f = f;
}

int upload_frame(xnb_batch* B, int i) {
  B->hF[i] = B->box[i].f;
  BHIP(B, hipMemcpyAsync(B->dF.p + i, B->hF + i, sizeof(BoxFrame), hipMemcpyHostToDevice, B->st));
  return XNB_OK;
}

// Hand the box's current arrays back to its context (positions in rbuf[rcur ^ par]).
void ctx_from_frame(xnb_batch* B, int i, uint32_t par) {
  BatchBox& bb = B->box[i];
  xnb_ctx* c = bb.c;
  const BoxFrame& f = bb.f;
  const uint32_t cur = (f.rcur ^ par) & 1u;
  if (cur) std::swap(c->rA, c->rB);  // rbuf[0] was rA when the frame was built
  if (f.ab) {
    for (int a = 0; a < 3; a++) {
      std::swap(c->vA[a], c->vB[a]);
      std::swap(c->fA[a], c->fB[a]);
    }
    std::swap(c->eA, c->eB);
    std::swap(c->idA, c->idB);
    for (int k = 0; k < c->n_extra; k++) std::swap(c->xA[k], c->xB[k]);
    std::swap(c->slot_cell_A, c->slot_cell_B);
  }
  bind_states(c);
  c->G = f.G;
}

// ---- work tables

struct Tables {
  std::vector<uint32_t> v;
  size_t add(const std::vector<uint32_t>& t) {
    const size_t o = v.size();
    v.insert(v.end(), t.begin(), t.end());
    return o;
  }
};

std::vector<uint32_t> blocks_table(const std::vector<int>& boxes, const std::vector<uint32_t>& items) {
  std::vector<uint32_t> t;
  for (size_t q = 0; q < boxes.size(); q++) {
    const uint32_t nb = (items[q] + kBatchSlotBlock - 1) / kBatchSlotBlock;
    for (uint32_t b = 0; b < nb; b++) t.push_back(((uint32_t)boxes[q] << kWorkBoxShift) | b);
  }
  return t;
}

// bricks of all listed boxes, concatenated, dealt to the 8 XCDs in contiguous runs
std::vector<uint32_t> bricks_table(const std::vector<int>& boxes, const std::vector<uint32_t>& nbricks, const std::vector<uint32_t>& tag) {
  std::vector<uint32_t> all;
  for (size_t q = 0; q < boxes.size(); q++)
    for (uint32_t b = 0; b < nbricks[q]; b++) all.push_back(work_brick_entry((uint32_t)boxes[q], tag[q], b));
  const size_t n = all.size(), per = (n + kXcds - 1) / kXcds;
  std::vector<uint32_t> t(per * kXcds, NONE);
  for (size_t j = 0; j < t.size(); j++) {
    const size_t g = (j % kXcds) * per + j / kXcds;
    if (g < n) t[j] = all[g];
  }
  return t;
}

uint32_t brick_lds_bytes(BrickParams bp, bool f32_build = false) { return brick_layout(bp, f32_build); }

// Workgroups of a box's bricks that fit one CU at once (LDS bytes and the waves the
// kernel's registers allow), 1..4.
uint32_t brick_class(uint32_t T, uint32_t lds, uint32_t cu_lds, uint32_t waves_per_cu) {
  const uint32_t by_lds = lds ? cu_lds / lds : 4u;
  const uint32_t by_waves = waves_per_cu / ((T + 63u) / 64u);
  return std::max(1u, std::min({by_lds, by_waves, 4u}));
}

// Bricks of the listed boxes for the batched kernel `kind` (for the force: the boxes of that
// precision), one launch per class of boxes whose workgroups fit the same number of times on
// a CU, each sized for its own widest box, so that boxes with small bricks do not run in the
// lanes and LDS of the largest one.  Results do not depend on the class: a brick's lanes
// beyond its box's T take no atoms.
void brick_launches(const xnb_batch* B, Tables& T, int kind, const std::vector<int>& boxes, const std::vector<uint32_t>& tag,
                    std::vector<BrickLaunch>& out) {
  const BatchKernelInfo& ki = B->kinfo[kind];
  for (uint32_t k = 1; k <= 4; k++) {
    std::vector<int> sub;
    std::vector<uint32_t> nbr, tg;
    BrickLaunch L;
    L.kind = kind;
    for (size_t q = 0; q < boxes.size(); q++) {
      const BatchBox& bb = B->box[boxes[q]];
      if (kind != kBatchNbh && (kind == kBatchLjMixed) != (bb.f.mixed != 0)) continue;
      const uint32_t lds = brick_lds_bytes(bb.f.bp, kind == kBatchNbh);
      if (B->use_classes && brick_class(bb.f.bp.T, lds + ki.static_lds, B->lds_limit, ki.waves_per_cu) != k) continue;
      sub.push_back(boxes[q]);
      nbr.push_back(bb.f.bp.nbricks);
      tg.push_back(tag[q]);
      L.T = std::max(L.T, bb.f.bp.T);
      L.lds = std::max(L.lds, lds);
    }
    if (!sub.empty()) {
      L.o = T.add(bricks_table(sub, nbr, tg));
      L.n = T.v.size() - L.o;
      out.push_back(L);
    }
    if (!B->use_classes) break;
  }
}

// pinned staging (two buffers) and B->dwork hold at least n words
int ensure_staging(xnb_batch* B, size_t n) {
  if (n > B->hwork_cap) {
    BHIP(B, hipStreamSynchronize(B->st));
    for (int k = 0; k < 2; k++) {
      if (B->hwork[k]) (void)hipHostFree(B->hwork[k]);  // after the synchronisation above
      B->hwork[k] = nullptr;
    }
    const size_t cap = n + n / 2 + 1024;
    for (int k = 0; k < 2; k++) BHIP(B, hipHostMalloc((void**)&B->hwork[k], cap * sizeof(uint32_t), hipHostMallocDefault));
    if (B->dwork.p) (void)hipFree(B->dwork.p);
    B->dwork.p = nullptr;
    BHIP(B, hipMalloc((void**)&B->dwork.p, cap * sizeof(uint32_t)));
    B->dwork.n = cap;
    B->hwork_cap = cap;
  }
  return XNB_OK;
}

// T through a staging buffer into dst (n words), in stream order
int stage_copy(xnb_batch* B, const Tables& T, uint32_t* dst) {
  const size_t n = T.v.size();
  const int k = B->hcur;
  B->hcur ^= 1;
  BHIP(B, hipEventSynchronize(B->hev[k]));  // the copy that last read this staging buffer is done
  std::memcpy(B->hwork[k], T.v.data(), n * sizeof(uint32_t));
  BHIP(B, hipMemcpyAsync(dst, B->hwork[k], n * sizeof(uint32_t), hipMemcpyHostToDevice, B->st));
  BHIP(B, hipEventRecord(B->hev[k], B->st));
  return XNB_OK;
}

int upload_tables(xnb_batch* B, const Tables& T) {
  const size_t n = T.v.size();
  if (n == 0) return XNB_OK;
  int e;
  if ((e = ensure_staging(B, n))) return e;
  return stage_copy(B, T, B->dwork.p);
}

void destroy_retired(xnb_batch* B) {  // after a stream synchronisation
  for (hipGraphExec_t x : B->retired) (void)hipGraphExecDestroy(x);
  B->retired.clear();
}

// Capture the quiet step of table T2 (force table at o_f, ghost refresh at o_g) into a new
// graph.  Its table goes to the graph buffer not used by the current graph: graphs launched
// earlier that read that buffer run before the copy (stream order).
// n launches over disjoint boxes (one per class), alternately on st and st2 when there are
// several: st2 waits for st's work so far and st waits for st2's launches before anything
// after them.  launch(i, stream) issues the i-th.
// (ev: the event pair of this group of launches; a step has two such groups)
template <class Launch>
hipError_t launch_split(xnb_batch* B, size_t n, int ev, Launch&& launch) {
  const bool split = B->split_streams && n > 1;
  hipError_t he = hipSuccess;
  if (split) {
    if ((he = hipEventRecord(B->fork_ev[ev], B->st)) != hipSuccess) return he;
    if ((he = hipStreamWaitEvent(B->st2, B->fork_ev[ev], 0)) != hipSuccess) return he;
  }
  for (size_t i = 0; i < n && he == hipSuccess; i++) he = launch(i, split && (i & 1) ? B->st2 : B->st);
  if (split) {
    const hipError_t a = hipEventRecord(B->join_ev[ev], B->st2);
    const hipError_t b = a == hipSuccess ? hipStreamWaitEvent(B->st, B->join_ev[ev], 0) : a;
    if (he == hipSuccess) he = b;
  }
  return he;
}

// the step's force launches at the device clock `clock`; set_lds: set each kernel's LDS
// attribute first (not under graph capture)
hipError_t launch_force_classes(xnb_batch* B, const uint32_t* W, const std::vector<BrickLaunch>& lf, bool set_lds,
                                const uint32_t* clock) {
  return launch_split(B, lf.size(), 1, [&](size_t i, hipStream_t s) {
    const BrickLaunch& L = lf[i];
    return launch_batch_lj(L.kind == kBatchLjMixed, B->dF.p, W + L.o, (uint32_t)L.n, L.T, L.lds, clock, s, set_lds);
  });
}

// Everything of a step after its first half, from work table W laid out as L: the rebuilds
// (move_particles, rebuild_amr, backup_r, ghosts, chunk_neighbors), the ghost refresh of the
// other boxes, the force of every box, the clock advance.  No host synchronisation and no
// memset nodes (the rebuild_amr counter is cleared by a kernel), so it can be captured.
hipError_t launch_step(xnb_batch* B, const uint32_t* W, const StepLayout& L, bool set_lds) {
  if (L.nreb) {
    launch_batch_zero_words(B->dwords.p, W + L.o_list, L.nreb, B->st);
    launch_batch_zero_words(B->amr_count.p, W + L.o_zero, 1, B->st);  // its entry is box 0: words 0..3
    launch_batch_move(B->dF.p, W + L.o_rs, (uint32_t)L.n_rs, W + L.o_rc, (uint32_t)L.n_rc, W + L.o_list, L.nreb, L.par, B->st);
    launch_batch_ghosts(B->dF.p, W + L.o_rs, (uint32_t)L.n_rs, W + L.o_rc, (uint32_t)L.n_rc, W + L.o_rg, (uint32_t)L.n_rg,
                        W + L.o_list, L.nreb, B->amr_list.p, B->amr_count.p, L.par, B->st);
    const hipError_t he = launch_split(B, L.lb.size(), 0, [&](size_t i, hipStream_t s) {
      return launch_batch_nbh(B->dF.p, W + L.lb[i].o, (uint32_t)L.lb[i].n, L.lb[i].T, L.lb[i].lds, L.par, s, set_lds);
    });
    if (he != hipSuccess) return he;
  }
  launch_batch_ghost_update_r(B->dF.p, W + L.o_g, (uint32_t)L.n_g, B->dclock.p, B->st);
  const hipError_t he = launch_force_classes(B, W, L.lf, set_lds, B->dclock.p);
  if (he != hipSuccess) return he;
  launch_batch_advance(B->dclock.p, B->st);
  return hipGetLastError();
}

void retire_graph(xnb_batch* B, StepGraph& G) {
  if (G.exec) B->retired.push_back(G.exec);
  G.exec = nullptr;
  G.tab.clear();
  G.have_last = false;
}

// Capture the step of table T laid out as L into a new graph of G.  Its table goes to the
// buffer of G not used by G's current graph: graphs launched earlier that read that buffer run
// before the copy (stream order).
int capture_step(xnb_batch* B, StepGraph& G, const Tables& T, const StepLayout& L) {
  const int j = G.cur ^ 1;
  const size_t n = T.v.size();
  int e;
  if ((e = ensure_staging(B, n))) return e;
  if (G.buf[j].n < n) {
    BHIP(B, hipStreamSynchronize(B->st));
    destroy_retired(B);
    if (G.buf[j].p) (void)hipFree(G.buf[j].p);
    G.buf[j].p = nullptr;
    G.buf[j].n = 0;
    BHIP(B, hipMalloc((void**)&G.buf[j].p, B->hwork_cap * sizeof(uint32_t)));
    G.buf[j].n = B->hwork_cap;
  }
  if ((e = stage_copy(B, T, G.buf[j].p))) return e;
  for (int kind : {kBatchLjFp64, kBatchLjMixed, kBatchNbh}) {  // attributes set outside the capture
    uint32_t lds = 0;
    for (const auto* v : {&L.lf, &L.lb})
      for (const BrickLaunch& x : *v)
        if (x.kind == kind) lds = std::max(lds, x.lds);
    if (lds) BHIP(B, batch_set_lds(kind, lds));
  }
  BHIP(B, hipStreamBeginCapture(B->st, hipStreamCaptureModeRelaxed));
  const hipError_t le = launch_step(B, G.buf[j].p, L, false);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(B->st, &g);
  if (le != hipSuccess || ce != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(B, XNB_E_HIP, std::string("step capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return bfail(B, XNB_E_HIP, std::string("step graph: ") + hipGetErrorString(ie));
  if (G.exec) B->retired.push_back(G.exec);
  G.exec = x;
  G.cur = j;
  G.tab = T.v;
  G.lay = L;
  B->n_captures++;
  return XNB_OK;
}

bool active(const BatchBox& bb) { return bb.used && !bb.finished && !bb.failed; }

// LDS bytes a brick workgroup may take for a cap of `frac` of the CU's LDS, less the
// kernels' static LDS
uint32_t lds_cap_for(const xnb_batch* B, double frac) {
  uint32_t st_lds = 0;
  for (int k = 0; k < 3; k++) st_lds = std::max(st_lds, B->kinfo[k].static_lds);
  return (uint32_t)(frac * B->lds_limit) - std::min<uint32_t>(st_lds + 64u, B->lds_limit / 8);
}

// chunk_neighbors of a box for the batch: brick head room and the batch's LDS cap per brick
// workgroup, or the context's own limit when no brick fits the cap (restored afterwards)
int batch_box_nbh(xnb_batch* B, xnb_ctx* c) {
  c->brick_margin = B->brick_margin;
  const uint32_t keep = c->lds_limit;
  xnb_chunk_neighbors_config cfg{1, 0, 0, 0, 0};
  int e = XNB_OK;
  bool done = false;
  c->batch_shapes = true;
  if (B->brick_lds_cap < keep) {
    c->lds_limit = B->brick_lds_cap;
    e = op_nbh(c, c->nbh_dist_lab, cfg, 0);
    c->lds_limit = keep;
    done = e || c->list_kind == 1;
  }
  if (!done) e = op_nbh(c, c->nbh_dist_lab, cfg, 0);
  c->batch_shapes = false;
  return e;
}

// One-box recovery after a batched rebuild overflowed the box's frame sizes: ghosts and
// lists again with the one-box operators (which re-size), then a new frame.
int recover_box(xnb_batch* B, int i, const xnb_run_params& p) {
  BatchBox& bb = B->box[i];
  xnb_ctx* c = bb.c;
  ctx_from_frame(B, i, B->par);
  int e;
  c->brick_margin = B->brick_margin;
  if ((e = ensure_capacity(c, (size_t)std::ceil((c->N + c->G) * B->slot_margin) + 256, c->N))) return e;
  if ((e = op_ghost_all(c, false, false))) return e;  // f and ep of the run's boxes are not carried
  if ((e = batch_box_nbh(B, c))) return e;
  if (c->list_kind != 1) return fail(c, XNB_E_STATE, "box no longer fits the brick path");
  if (c->bp.nbricks > kWorkBrickMask + 1) return fail(c, XNB_E_CAPACITY, "box has too many bricks for a batch");
  BHIP(B, hipStreamSynchronize(c->st));
  frame_from_ctx(B, i, p);
  int e2;
  if ((e2 = upload_frame(B, i))) return e2;
  // the overflow the batched rebuild left in the status words no longer holds
  BHIP(B, hipMemsetAsync(B->dwords.p + 4 * (size_t)i, 0, 4 * sizeof(uint32_t), B->st));
  return XNB_OK;
}

}  // namespace

extern "C" {

int xnb_batch_create(xnb_batch** out, int device) {
  if (!out) return XNB_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return XNB_E_HIP;
  if (hipSetDevice(device) != hipSuccess) return XNB_E_HIP;
  xnb_batch* B = new xnb_batch;
  B->device = device;
  B->box.resize(kMaxBatchBoxes);
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    B->lds_limit = (uint32_t)lds;
  for (int k = 0; k < 3; k++) B->kinfo[k] = batch_kernel_info(k);
  bool ok = hipStreamCreate(&B->st) == hipSuccess && hipStreamCreate(&B->st2) == hipSuccess;
  for (int k = 0; k < 2; k++)
    ok = ok && hipEventCreateWithFlags(&B->fork_ev[k], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&B->join_ev[k], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipMalloc((void**)&B->dF.p, kMaxBatchBoxes * sizeof(BoxFrame)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&B->hF, kMaxBatchBoxes * sizeof(BoxFrame), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc((void**)&B->dwords.p, 4 * kMaxBatchBoxes * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&B->hwords, 4 * kMaxBatchBoxes * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc((void**)&B->dtrig.p, kMaxBatchBoxes * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&B->htrig, kMaxBatchBoxes * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipEventCreate(&B->hev[0]) == hipSuccess && hipEventCreate(&B->hev[1]) == hipSuccess;
  ok = ok && hipMalloc((void**)&B->dclock.p, 4 * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMalloc((void**)&B->amr_count.p, 4 * sizeof(uint32_t)) == hipSuccess;
  if (ok) {
    B->dclock.n = 4;
    B->amr_count.n = 4;
    const uint32_t clock0[4] = {B->par, B->gstep, B->par, B->gstep};
    ok = hipMemcpy(B->dclock.p, clock0, sizeof(clock0), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (const char* s = std::getenv("XNB_BATCH_GRAPH")) B->use_graph = std::atoi(s) != 0;
  if (const char* s = std::getenv("XNB_BATCH_CLASSES")) B->use_classes = std::atoi(s) != 0;
  if (const char* s = std::getenv("XNB_BATCH_STREAMS")) B->split_streams = std::atoi(s) != 0;
  double lds_frac = 0.5;
  if (const char* s = std::getenv("XNB_BATCH_LDS")) {
    const double f = std::atof(s);
    if (f > 0.0 && f <= 1.0) lds_frac = f;
  }
  B->lds_frac = lds_frac;
  B->brick_lds_cap = lds_cap_for(B, lds_frac);
  if (const char* s = std::getenv("XNB_BATCH_MARGINS")) {
    double bm = 0, sm = 0;
    if (std::sscanf(s, "%lf,%lf", &bm, &sm) != 2 || !(bm >= 1.0 && bm <= 4.0) || !(sm >= 1.0 && sm <= 4.0)) {
      xnb_batch_destroy(B);
      return XNB_E_ARG;
    }
    B->brick_margin = bm;
    B->slot_margin = sm;
  }
  if (ok) ok = hipMemset(B->dwords.p, 0, 4 * kMaxBatchBoxes * sizeof(uint32_t)) == hipSuccess;
  if (ok) ok = hipMemset(B->dtrig.p, 0, kMaxBatchBoxes * sizeof(uint32_t)) == hipSuccess;
  // both staging events start recorded, so the first waits on them return at once
  if (ok) ok = hipEventRecord(B->hev[0], B->st) == hipSuccess && hipEventRecord(B->hev[1], B->st) == hipSuccess;
  if (!ok) {
    xnb_batch_destroy(B);  // frees whatever was created
    return XNB_E_HIP;
  }
  *out = B;
  return XNB_OK;
}

int xnb_batch_destroy(xnb_batch* B) {
  if (!B) return XNB_E_ARG;
  // teardown: failures here leave nothing to recover, so results are not checked
  (void)hipSetDevice(B->device);
  if (B->st) (void)hipStreamSynchronize(B->st);
  if (B->st2) (void)hipStreamSynchronize(B->st2);  // joined into st, unless a launch failed in between
  for (auto& bb : B->box)
    if (bb.counts.p) (void)hipFree(bb.counts.p);
  if (B->dF.p) (void)hipFree(B->dF.p);
  if (B->hF) (void)hipHostFree(B->hF);
  if (B->dwords.p) (void)hipFree(B->dwords.p);
  if (B->hwords) (void)hipHostFree(B->hwords);
  if (B->dtrig.p) (void)hipFree(B->dtrig.p);
  if (B->htrig) (void)hipHostFree(B->htrig);
  if (B->dwork.p) (void)hipFree(B->dwork.p);
  destroy_retired(B);
  for (StepGraph* G : {&B->gq, &B->gr[0], &B->gr[1]}) {
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    for (int k = 0; k < 2; k++)
      if (G->buf[k].p) (void)hipFree(G->buf[k].p);
  }
  if (B->dclock.p) (void)hipFree(B->dclock.p);
  if (B->amr_count.p) (void)hipFree(B->amr_count.p);
  if (B->amr_list.p) (void)hipFree(B->amr_list.p);
  for (int k = 0; k < 2; k++) {
    if (B->hwork[k]) (void)hipHostFree(B->hwork[k]);
    if (B->hev[k]) (void)hipEventDestroy(B->hev[k]);
    if (B->fork_ev[k]) (void)hipEventDestroy(B->fork_ev[k]);
    if (B->join_ev[k]) (void)hipEventDestroy(B->join_ev[k]);
  }
  if (B->st2) (void)hipStreamDestroy(B->st2);
  if (B->st) (void)hipStreamDestroy(B->st);
  delete B;
  return XNB_OK;
}

const char* xnb_batch_error(const xnb_batch* B) { return B ? B->err.c_str() : "null batch"; }

int xnb_batch_attach(xnb_batch* B, xnb_ctx* c, const xnb_run_params* p, int* slot) {
  if (!B || !c || !p || !slot || p->steps < 1) return bfail(B, XNB_E_ARG, "bad attach arguments");
  int e;
  if ((e = check_ctx(c))) return e;
  if (c->device != B->device) return bfail(B, XNB_E_ARG, "box on another device");
  if (!c->lists_valid || !c->ghosts_valid || !c->backup_valid) return bfail(B, XNB_E_STATE, "xnb_init must run first");
  if (ensure_full_lists(c)) return bfail(B, XNB_E_STATE, std::string("full neighbour lists: ") + c->err);
  // work and rebuild_amr list entries hold a cell or block index in kWorkBoxShift bits
  if (c->g.ncells > kWorkBlockMask) return bfail(B, XNB_E_CAPACITY, "box has too many cells for a batch");
  int i = -1;
  for (int q = 0; q < (int)kMaxBatchBoxes; q++)
    if (!B->box[q].used) {
      i = q;
      break;
    }
  if (i < 0) return bfail(B, XNB_E_CAPACITY, "batch is full");
  // head room: ghost slots, and bricks / lists rebuilt with margins
  if ((e = ensure_capacity(c, (size_t)std::ceil((c->N + c->G) * B->slot_margin) + 256, (size_t)c->N + c->G))) return e;
  if ((e = batch_box_nbh(B, c))) return e;
  if (c->list_kind != 1) return bfail(B, XNB_E_STATE, "box geometry is not on the brick path (cell edge < nbh_dist)");
  if (brick_lds_bytes(c->bp) > B->lds_limit) return bfail(B, XNB_E_CAPACITY, "box bricks exceed LDS");
  // force-table entries hold the brick index in kWorkEpiShift bits
  if (c->bp.nbricks > kWorkBrickMask + 1) return bfail(B, XNB_E_CAPACITY, "box has too many bricks for a batch");
  if (hipStreamSynchronize(c->st) != hipSuccess) return bfail(B, XNB_E_HIP, "box stream");
  BatchBox& bb = B->box[i];
  bb = BatchBox{};
  bb.c = c;
  bb.used = true;
  bb.steps = p->steps;
  bb.every = std::max(p->rebuild_every, 0);  // <= 0: displacement trigger, as in xnb_run
  bb.rebuilt.assign((size_t)p->steps, 0);
  if (hipMalloc((void**)&bb.counts.p, (size_t)p->steps * sizeof(unsigned long long)) != hipSuccess) {
    bb.used = false;
    return bfail(B, XNB_E_NOMEM, "step counts");
  }
  bb.counts.n = (size_t)p->steps;
  BHIP(B, hipMemsetAsync(bb.counts.p, 0, (size_t)p->steps * sizeof(unsigned long long), B->st));
  // the slot's status words: no overflow (a previous box may have left flags)
  BHIP(B, hipMemsetAsync(B->dwords.p + 4 * (size_t)i, 0, 4 * sizeof(uint32_t), B->st));
  bb.p = *p;
  c->last_rcut = p->rcut;
  frame_from_ctx(B, i, *p);
  if ((e = upload_frame(B, i))) return e;
  *slot = i;
  return XNB_OK;
}

// Advance every unfinished box by up to nsteps steps (fewer for boxes that reach their own
// step count).  *n_active receives the number of boxes still unfinished afterwards.
int xnb_batch_step(xnb_batch* B, int nsteps, int* n_active) {
  if (!B || nsteps < 0) return XNB_E_ARG;
  BHIP(B, hipSetDevice(B->device));
  int e;
  for (int s = 0; s < nsteps; s++) {
    std::vector<int> act;
    for (int i = 0; i < (int)kMaxBatchBoxes; i++)
      if (active(B->box[i])) act.push_back(i);
    if (act.empty()) break;
    const uint32_t par = B->par, gstep = B->gstep;
    // 1. first half of boxes whose previous force kernel did not do it; then the trigger of
    // boxes rebuilt on displacement: this step's particle_displ_over count, written by the
    // previous force kernel or by that first half (one host read per step while such boxes run)
    std::vector<int> fresh, disp, reb, quiet;
    for (int i : act) {
      if (B->box[i].fresh) fresh.push_back(i);
      if (B->box[i].every == 0) disp.push_back(i);
    }
    {
      Tables T0;
      std::vector<uint32_t> nF;
      for (int i : fresh) nF.push_back(B->box[i].f.N);
      const size_t o_fresh = T0.add(blocks_table(fresh, nF));
      const size_t n_fresh = T0.v.size() - o_fresh;
      const size_t o_disp = T0.add(std::vector<uint32_t>(disp.begin(), disp.end()));
      if ((e = upload_tables(B, T0))) return e;
      const uint32_t* W0 = B->dwork.p;
      if (n_fresh) launch_batch_first_half(B->dF.p, W0 + o_fresh, (uint32_t)n_fresh, par, gstep, B->st);
      for (int i : fresh) B->box[i].fresh = false;
      if (!disp.empty()) {
        launch_batch_trigger(B->dF.p, W0 + o_disp, (uint32_t)disp.size(), gstep, B->dtrig.p, B->st);
        BHIP(B, hipMemcpyAsync(B->htrig, B->dtrig.p, kMaxBatchBoxes * sizeof(uint32_t), hipMemcpyDeviceToHost, B->st));
        BHIP(B, hipStreamSynchronize(B->st));
      }
    }
    for (int i : act) {
      const BatchBox& bb = B->box[i];
      const int64_t k = bb.done + 1;
      const bool trigger = bb.every > 0 ? k % bb.every == 0 : B->htrig[i] != 0;
      if (trigger) reb.push_back(i);
      else quiet.push_back(i);
    }
    // 2. One work table for the rest of the step: the rebuilds of the boxes that rebuild, the
    // ghost refresh of the others, the force of every box (launch_step).
    Tables T;
    StepLayout L;
    if (!reb.empty()) {
      std::vector<uint32_t> nN, nC, nGh;
      for (int i : reb) {
        nN.push_back(B->box[i].f.N);
        nC.push_back(B->box[i].f.g.ncells);
        nGh.push_back(ghost_cell_count(B->box[i].f.g) * 64u);  // one wave per ghost cell
      }
      L.par = par;
      L.nreb = (uint32_t)reb.size();
      L.o_rs = T.add(blocks_table(reb, nN));
      L.n_rs = T.v.size() - L.o_rs;
      L.o_rc = T.add(blocks_table(reb, nC));
      L.n_rc = T.v.size() - L.o_rc;
      L.o_rg = T.add(blocks_table(reb, nGh));
      L.n_rg = T.v.size() - L.o_rg;
      brick_launches(B, T, kBatchNbh, reb, std::vector<uint32_t>(reb.size(), 0u), L.lb);
      L.o_list = T.add(std::vector<uint32_t>(reb.begin(), reb.end()));
      L.o_zero = T.add(std::vector<uint32_t>{0u});
      // the move swaps each rebuilding box's buffer roles on the device; the host mirror follows
      for (int i : reb) {
        B->box[i].f.ab ^= 1u;
        B->box[i].f.rcur ^= 1u;
        B->box[i].f.G = 0;
      }
      size_t cells = 0;
      for (uint32_t x : nC) cells += x;
      if (B->amr_list.n < cells) {
        retire_graph(B, B->gr[0]);  // captured with the old list
        retire_graph(B, B->gr[1]);
        BHIP(B, hipStreamSynchronize(B->st));
        destroy_retired(B);
        if (B->amr_list.p) (void)hipFree(B->amr_list.p);
        B->amr_list.p = nullptr;
        B->amr_list.n = 0;
        BHIP(B, hipMalloc((void**)&B->amr_list.p, cells * sizeof(uint32_t)));
        B->amr_list.n = cells;
      }
    }
    // Ghost refresh blocks of a box cover 10 % more than the largest ghost count it has had
    // (the kernel stops at the current count), so the table stays the same over rebuilds that
    // move ghost counts a little.
    std::vector<int> q2;
    std::vector<uint32_t> nG;
    for (int i : quiet) {
      BatchBox& bb = B->box[i];
      if (bb.f.G > bb.ghost_span) bb.ghost_span = bb.f.G + bb.f.G / 10 + 1;
      if (bb.ghost_span) {
        q2.push_back(i);
        nG.push_back(bb.ghost_span);
      }
    }
    L.o_g = T.add(blocks_table(q2, nG));
    L.n_g = T.v.size() - L.o_g;
    std::vector<int> frc;
    std::vector<uint32_t> tag;
    for (int i : act) {
      BatchBox& bb = B->box[i];
      if (bb.failed) continue;
      const int64_t k = bb.done + 1;
      // a fused step stores no forces or energies: only the last step's are read (see run_loop)
      const bool fuse = k < bb.steps;
      frc.push_back(i);
      tag.push_back(fuse ? (uint32_t)EPI_NEXT : (uint32_t)EPI_KICK);
    }
    for (int kind : {kBatchLjMixed, kBatchLjFp64}) brick_launches(B, T, kind, frc, tag, L.lf);
    // 3. Launch it, from a captured graph when a step of the same kind (quiet, or rebuilding
    // at fixed intervals) had the same table and layout before.
    const bool plain = fresh.empty() && disp.empty() && !T.v.empty();
    StepGraph& G = reb.empty() ? B->gq : B->gr[par & 1u];
    bool graphed = false;
    if (plain && B->use_graph) {
      auto same = [&]() { return G.exec && T.v == G.tab && L == G.lay; };
      if (!same() && G.have_last && T.v == G.last_tab && L == G.last_lay && (e = capture_step(B, G, T, L))) return e;
      if (same()) {
        BHIP(B, hipGraphLaunch(G.exec, B->st));
        graphed = true;
        B->n_graph_steps++;
      }
      G.last_tab = T.v;
      G.last_lay = L;
      G.have_last = true;
      if (B->retired.size() > 4) {
        BHIP(B, hipStreamSynchronize(B->st));
        destroy_retired(B);
      }
    }
    // quiet graphs are captured from two quiet steps in a row
    if (!plain || !reb.empty()) B->gq.have_last = false;
    if (!graphed) {
      if ((e = upload_tables(B, T))) return e;
      const hipError_t he = launch_step(B, B->dwork.p, L, true);
      if (he != hipSuccess) return bfail(B, XNB_E_HIP, std::string("batched step: ") + hipGetErrorString(he));
    }
    // 4. After a rebuild step: the status words of the rebuilt boxes (ghost count, longest
    // list, overflow and failure flags).  A box whose rebuild overflowed was left untouched by
    // the force kernel; it is redone with the one-box operators and given its force at this
    // step's clock.
    if (!reb.empty()) {
      BHIP(B, hipMemcpyAsync(B->hwords, B->dwords.p, 4 * kMaxBatchBoxes * sizeof(uint32_t), hipMemcpyDeviceToHost, B->st));
      BHIP(B, hipStreamSynchronize(B->st));
      destroy_retired(B);
      std::vector<int> late;
      std::vector<uint32_t> late_tag;
      for (int i : reb) {
        BatchBox& bb = B->box[i];
        const uint32_t* w = B->hwords + 4 * (size_t)i;
        bb.f.G = w[3];
        bb.c->max_neighbors = w[1];
        bb.c->slot_lists_current = false;
        bb.n_rebuilds++;
        bb.rebuilt[(size_t)bb.done] = 1;
        if (w[2]) {
          bb.failed = true;
          bb.err = "a particle left the domain along a non-periodic axis";
          continue;
        }
        if (w[0] & 8u) {
          bb.failed = true;
          bb.err = "rebuild_amr: a cell holds more than 4096 particles";
          continue;
        }
        if (w[0] || w[1] > bb.f.bp.cap) {
          bb.n_recover++;
          if ((e = recover_box(B, i, bb.p))) {
            bb.failed = true;
            bb.err = std::string("rebuild recovery failed: ") + bb.c->err;
            continue;
          }
          late.push_back(i);
          late_tag.push_back(bb.done + 1 < bb.steps ? (uint32_t)EPI_NEXT : (uint32_t)EPI_KICK);
        }
      }
      if (!late.empty()) {
        Tables TL;
        std::vector<BrickLaunch> ll;
        for (int kind : {kBatchLjMixed, kBatchLjFp64}) brick_launches(B, TL, kind, late, late_tag, ll);
        if ((e = upload_tables(B, TL))) return e;
        const hipError_t he = launch_force_classes(B, B->dwork.p, ll, true, B->dclock.p + 2);
        if (he != hipSuccess) return bfail(B, XNB_E_HIP, std::string("batched force after recovery: ") + hipGetErrorString(he));
      }
    }
    for (size_t q = 0; q < frc.size(); q++) {
      BatchBox& bb = B->box[frc[q]];
      if (bb.failed) continue;  // failed in this step's rebuild: its force did not run
      bb.done++;
      if (tag[q] == (uint32_t)EPI_KICK) {
        bb.finished = true;
        bb.last_par = par;
      }
    }
    B->par ^= 1u;
    B->gstep++;
  }
  if (n_active) {
    int n = 0;
    for (auto& bb : B->box)
      if (active(bb)) n++;
    *n_active = n;
  }
  BHIP(B, hipGetLastError());
  return XNB_OK;
}

// Progress of the box in a slot: steps done, finished (1), failed (-1, message in
// xnb_batch_error after the call).
int xnb_batch_box_status(xnb_batch* B, int slot, int64_t* done, int* state) {
  if (!B || slot < 0 || slot >= (int)kMaxBatchBoxes || !B->box[slot].used) return XNB_E_ARG;
  const BatchBox& bb = B->box[slot];
  if (done) *done = bb.done;
  if (state) *state = bb.failed ? -1 : (bb.finished ? 1 : 0);
  if (bb.failed) B->err = bb.err;
  return XNB_OK;
}

// Rebuilds of the box in a slot that outgrew its frame and were redone with the one-box
// operators.
int xnb_batch_box_recoveries(xnb_batch* B, int slot, int64_t* n) {
  if (!B || slot < 0 || slot >= (int)kMaxBatchBoxes || !B->box[slot].used || !n) return XNB_E_ARG;
  *n = B->box[slot].n_recover;
  return XNB_OK;
}

int xnb_batch_graph_stats(xnb_batch* B, int64_t* graph_steps, int64_t* captures) {
  if (!B) return XNB_E_ARG;
  if (graph_steps) *graph_steps = B->n_graph_steps;
  if (captures) *captures = B->n_captures;
  return XNB_OK;
}

// Batch autotuner.  The LDS bytes a brick workgroup may take decide each box's brick shape
// (cells per workgroup, lanes per workgroup) and how many of its workgroups share a CU.  For
// each candidate fraction of the CU's LDS, every attached box is laid out again under that
// cap (its lists rebuilt: lists and forces do not depend on the brick shape) and the batched
// neighbour build and force (force and energy stored, from the same positions, so the same
// values) are timed; the fraction with the least force time plus a tenth of the build time
// (a rebuild every 10 steps) is kept for the boxes and for later attaches and recoveries.
// Every attached box must be before its first step.
int xnb_batch_autotune(xnb_batch* B, double* lds_frac_out) {
  if (!B) return XNB_E_ARG;
  BHIP(B, hipSetDevice(B->device));
  std::vector<int> boxes;
  for (int i = 0; i < (int)kMaxBatchBoxes; i++) {
    const BatchBox& bb = B->box[i];
    if (!bb.used) continue;
    if (!bb.fresh || bb.done != 0 || bb.failed) return bfail(B, XNB_E_STATE, "autotune runs before the first step of every attached box");
    boxes.push_back(i);
  }
  if (boxes.empty()) return bfail(B, XNB_E_STATE, "autotune needs attached boxes");
  BHIP(B, hipStreamSynchronize(B->st));
  for (StepGraph* G : {&B->gq, &B->gr[0], &B->gr[1]}) retire_graph(B, *G);
  destroy_retired(B);
  int e;
  // every box laid out again under the cap of `frac`
  auto layout = [&](double frac) -> int {
    B->lds_frac = frac;
    B->brick_lds_cap = lds_cap_for(B, frac);
    for (int i : boxes) {
      BatchBox& bb = B->box[i];
      xnb_ctx* c = bb.c;
      ctx_from_frame(B, i, B->par);
      int e2;
      if ((e2 = batch_box_nbh(B, c))) return bfail(B, e2, std::string("autotune layout: ") + c->err);
      if (c->list_kind != 1 || brick_lds_bytes(c->bp) > B->lds_limit || c->bp.nbricks > kWorkBrickMask + 1)
        return bfail(B, XNB_E_CAPACITY, "autotune layout: a box left the batch's brick path");
      BHIP(B, hipStreamSynchronize(c->st));
      frame_from_ctx(B, i, bb.p);
      if ((e2 = upload_frame(B, i))) return e2;
    }
    BHIP(B, hipMemsetAsync(B->dwords.p, 0, 4 * kMaxBatchBoxes * sizeof(uint32_t), B->st));
    return XNB_OK;
  };
  struct Events {  // destroyed on every return path
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    }
  } E;
  for (hipEvent_t& x : E.ev) BHIP(B, hipEventCreate(&x));
  // force time + build time / 10 of the current layout, in ms per step; < 0 when a build overflowed
  auto score = [&](double& ms) -> int {
    Tables T;
    std::vector<BrickLaunch> lb, lf;
    brick_launches(B, T, kBatchNbh, boxes, std::vector<uint32_t>(boxes.size(), 0u), lb);
    for (int kind : {kBatchLjMixed, kBatchLjFp64})
      brick_launches(B, T, kind, boxes, std::vector<uint32_t>(boxes.size(), (uint32_t)EPI_NONE), lf);
    const size_t o_list = T.add(std::vector<uint32_t>(boxes.begin(), boxes.end()));
    int e2;
    if ((e2 = upload_tables(B, T))) return e2;
    const uint32_t* W = B->dwork.p;
    const uint32_t par = B->par;
    auto build = [&]() {
      launch_batch_zero_words(B->dwords.p, W + o_list, (uint32_t)boxes.size(), B->st);
      return launch_split(B, lb.size(), 0, [&](size_t i, hipStream_t st) {
        return launch_batch_nbh(B->dF.p, W + lb[i].o, (uint32_t)lb[i].n, lb[i].T, lb[i].lds, par, st, true);
      });
    };
    constexpr int kReps = 3;
    hipError_t he = build();  // untimed round
    if (he == hipSuccess) he = launch_force_classes(B, W, lf, true, B->dclock.p);
    if (he == hipSuccess) he = hipEventRecord(E.ev[0], B->st);
    for (int q = 0; q < kReps && he == hipSuccess; q++) he = build();
    if (he == hipSuccess) he = hipEventRecord(E.ev[1], B->st);
    for (int q = 0; q < kReps && he == hipSuccess; q++) he = launch_force_classes(B, W, lf, true, B->dclock.p);
    if (he == hipSuccess) he = hipEventRecord(E.ev[2], B->st);
    if (he != hipSuccess) return bfail(B, XNB_E_HIP, std::string("autotune launches: ") + hipGetErrorString(he));
    BHIP(B, hipMemcpyAsync(B->hwords, B->dwords.p, 4 * kMaxBatchBoxes * sizeof(uint32_t), hipMemcpyDeviceToHost, B->st));
    BHIP(B, hipStreamSynchronize(B->st));
    float tb = 0.f, tf = 0.f;
    BHIP(B, hipEventElapsedTime(&tb, E.ev[0], E.ev[1]));
    BHIP(B, hipEventElapsedTime(&tf, E.ev[1], E.ev[2]));
    ms = (tf + 0.1 * tb) / kReps;
    for (int i : boxes) {
      const uint32_t* w = B->hwords + 4 * (size_t)i;
      if (w[0] || w[1] > B->box[i].f.bp.cap || w[2]) ms = -1.0;
    }
    return XNB_OK;
  };
  static const double cand[] = {1.0 / 3.0, 0.5, 0.75, 1.0};
  double best = -1.0, best_ms = 0.0;
  for (double frac : cand) {
    double ms = 0.0;
    if ((e = layout(frac)) || (e = score(ms))) return e;
    if (ms >= 0.0 && (best < 0.0 || ms < best_ms)) {
      best = frac;
      best_ms = ms;
    }
  }
  if (best < 0.0) best = 0.5;
  // the boxes as they will run: laid out under the kept cap, lists, forces and status words
  // those of that layout
  if ((e = layout(best))) return e;
  BHIP(B, hipStreamSynchronize(B->st));
  if (lds_frac_out) *lds_frac_out = best;
  return XNB_OK;
}

// Detach a box (finished or not): its context holds the current state again, as after
// xnb_run; the run's rebuild and guard-violation counts go to p->n_rebuilds/n_violations.
int xnb_batch_detach(xnb_batch* B, int slot, xnb_run_params* p) {
  if (!B || slot < 0 || slot >= (int)kMaxBatchBoxes || !B->box[slot].used) return XNB_E_ARG;
  BHIP(B, hipSetDevice(B->device));
  BHIP(B, hipStreamSynchronize(B->st));
  BatchBox& bb = B->box[slot];
  ctx_from_frame(B, slot, bb.finished ? bb.last_par : B->par);
  // a box detached before its last step keeps the fused first half of its next step; the
  // one-box loop restarts from a completed step, so only finished boxes are returned as
  // such (unfinished boxes are for tear-down)
  int64_t viol = 0;
  if (bb.done > 0) {
    std::vector<unsigned long long> h((size_t)bb.done);
    BHIP(B, hipMemcpy(h.data(), bb.counts.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < h.size(); q++)
      if (h[q] > 0 && !bb.rebuilt[q]) viol++;
  }
  if (p) {
    p->n_rebuilds = bb.n_rebuilds;
    p->n_violations = viol;
  }
  if (!bb.finished) {
    // positions and velocities already hold the first half of the next step: the context is
    // not a completed step, so xnb_run and the force operators refuse it (XNB_E_STATE) until
    // it is initialised again
    bb.c->lists_valid = false;
    bb.c->backup_valid = false;
  }
  const int rc = bb.failed ? XNB_E_STATE : XNB_OK;
  if (bb.failed) B->err = bb.err;
  (void)hipFree(bb.counts.p);  // the batch stream was synchronised above
  bb = BatchBox{};
  return rc;
}

}  // extern "C"
