// xnb_kernels.h -- launch interface of the gfx950 kernels (host side of xnb_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xnb_device.h"
#include "xnb_pair.h"

namespace xnb {

// Brick tiling of the inner domain (see xnb_brick.hip) and the LDS layout of one workgroup.
// Workgroups are dispatched to the 8 XCDs of the chip round-robin; brick kernels give each
// XCD a contiguous run of bricks so that neighbouring bricks, which share halo cells, run
// under the same L2.
constexpr uint32_t kXcds = 8;
// Brick lists are padded per wave to a multiple of this many entries with the region's
// sentinel atom (far away, so every pair with it fails the cutoff test).
constexpr uint32_t kListBatch = 8;

struct BrickParams {
  int Bi, Bj, Bk;     // brick size in cells
  int nbx, nby, nbz;  // bricks per axis
  int RX, RY, RZ;     // region (brick + 1 cell each side)
  int nrc, nbc;       // region cells, brick cells
  uint32_t nbricks;
  uint32_t per_xcd;  // bricks per XCD: workgroup i runs brick (i % 8) * per_xcd + i / 8
  uint32_t T;     // threads per workgroup (>= brick atoms, nrc, nbc; multiple of 64)
  uint32_t Rmax;  // largest region atom count
  uint32_t cap;   // list rows per brick slab
  uint32_t* status;  // device word: bit 0 set when a region exceeds Rmax, bit 1 when a brick exceeds T
  uint32_t off_rcnt, off_rpref, off_roff, off_bpref, off_pos, off_type;  // LDS byte offsets
  uint32_t off_rdel, off_crow;  // staging tables: per region row, per 64 region atoms
  uint32_t off_wsum;            // 32 scan words, then the 2 wave-group counter words
};

// Per type-pair list radii of a multi-material run: a candidate b of an atom a is kept when
// d2 <= thr2[type a][type b] (each at most the build's nbh_dist^2); nt = 0: one radius for all.
struct PairCut {
  int nt;
  double thr2[4][4];
};

// Particle arrays of one storage buffer (slots [0, N) inner, [N, N+G) ghosts).
// Force-kernel epilogues (see k_lj_brick)
enum { EPI_NONE = 0, EPI_KICK = 1, EPI_NEXT = 2 };

// Next-step first half fused into the force kernel: positions go to r_next, the count of
// atoms displaced by >= sqrt(thr2) from their backup is added to *count.
struct NextStep {
  double4* r_next;
  const uint32_t* backup;
  double dt, dt2, thr2;
  unsigned long long* count;
};

// Extra per-particle FP64 fields a grid may carry beside the fixed ones (a field set with more
// fields, make_grid_variant_operator.h:104-164): moved, sorted and copied to ghost images with
// their particle, never read by the force kernels.
constexpr int kMaxExtraFields = 4;

struct PState {
  double4* r;  // x, y, z, type
  double *vx, *vy, *vz;
  double *fx, *fy, *fz;
  double* ep;  // per-atom potential energy (sum of pair energies over the full list)
  uint64_t* id;
  double* x[kMaxExtraFields];  // extra fields [0, nx)
  uint32_t nx;
};

// ------------------------------------------------------------------ batched boxes
// One box of a batch as the batched kernels see it (device memory, one entry per slot of the
// batch).  Kernels may update ab, rcur and G; the host keeps a mirror.
struct BoxFrame {
  Geom g;
  BrickParams bp;  // bp.status == words + 0
  BrickParams bpn;  // bp with the prefiltered neighbour build's LDS layout
  uint32_t N, G, cap;
  uint32_t ab;     // 1: the roles of the A and B field arrays are exchanged
  uint32_t rcur;   // current positions are rbuf[rcur ^ parity] (parity: argument of the launch)
  uint32_t mixed;  // 1: mixed precision force
  PState A, B;     // .r unused: positions live in rbuf
  double4* rbuf[2];
  uint32_t* slot_cell[2];  // [0] goes with A, [1] with B
  uint32_t *cell_off, *cell_n, *cnt_out, *cnt_in, *new_n, *new_off, *in_cur, *gcnt, *goff;
  uint32_t *mv_dst, *pk, *newloc, *in_list, *ghost_src, *ghost_shift, *backup;
  uint16_t* blist;
  uint32_t* bcnt;
  LJConst lj;
  double dt, dt2, dth, thr2, max2;
  uint32_t* words;             // [0] status bits (1, 2 brick sizes; 4 ghost capacity), [1] max list count, [2] left the domain, [3] ghost count
  unsigned long long* counts;  // displacement count of local step k at [k - 1]
  uint32_t gstep0;             // global step number of the box's local step 1
};

// work tables: one u32 per workgroup, box << 20 | block within the box, or NONE; force tables
// carry the box's epilogue in bits 18-19 (box << 20 | epi << 18 | brick)
constexpr uint32_t kWorkBoxShift = 20;
constexpr uint32_t kWorkBlockMask = (1u << kWorkBoxShift) - 1;
constexpr uint32_t kWorkEpiShift = 18;
constexpr uint32_t kWorkBrickMask = (1u << kWorkEpiShift) - 1;
constexpr uint32_t kMaxBatchBoxes = 4096;  // box index < 2^12
// force-table entry of brick `brick` of box `box` with epilogue `epi`; a batched box may have
// at most kWorkBrickMask + 1 bricks (attach and recovery reject more)
__host__ __device__ constexpr uint32_t work_brick_entry(uint32_t box, uint32_t epi, uint32_t brick) {
  return (box << kWorkBoxShift) | (epi << kWorkEpiShift) | brick;
}
__host__ __device__ constexpr uint32_t work_entry_box(uint32_t w) { return w >> kWorkBoxShift; }
__host__ __device__ constexpr uint32_t work_entry_epi(uint32_t w) { return (w & kWorkBlockMask) >> kWorkEpiShift; }
__host__ __device__ constexpr uint32_t work_entry_brick(uint32_t w) { return w & kWorkBrickMask; }
// round trips at the field limits (checked by the compiler)
static_assert(work_entry_box(work_brick_entry(kMaxBatchBoxes - 1, 3, kWorkBrickMask)) == kMaxBatchBoxes - 1, "box field");
static_assert(work_entry_epi(work_brick_entry(kMaxBatchBoxes - 1, 3, kWorkBrickMask)) == 3, "epilogue field");
static_assert(work_entry_brick(work_brick_entry(kMaxBatchBoxes - 1, 3, kWorkBrickMask)) == kWorkBrickMask, "brick field");
static_assert(work_entry_epi(work_brick_entry(0, 2, 0)) == 2 && work_entry_brick(work_brick_entry(1, 0, 1)) == 1, "fields");
static_assert((kMaxBatchBoxes - 1) <= (0xffffffffu >> kWorkBoxShift), "box index fits above the block field");

void launch_batch_zero_cells(const BoxFrame* F, const uint32_t* work, uint32_t nwork, hipStream_t st);
void launch_batch_zero_words(uint32_t* words, const uint32_t* boxes, uint32_t n, hipStream_t st);
void launch_batch_move(BoxFrame* F, const uint32_t* wslots, uint32_t nslots, const uint32_t* wcells, uint32_t ncells,
                       const uint32_t* boxes, uint32_t nboxes, uint32_t par, hipStream_t st);
// rebuild_amr, backup_r, ghost_comm_scheme + ghost_update_all of the listed boxes after their
// move; amr_list holds one word per cell of the listed boxes, *amr_count is 0 on entry; a box
// with a cell too dense for rebuild_amr gets status bit 8
void launch_batch_ghosts(BoxFrame* F, const uint32_t* wslots, uint32_t nslots, const uint32_t* wcells, uint32_t ncells,
                         const uint32_t* wghost, uint32_t nghost, const uint32_t* boxes, uint32_t nboxes, uint32_t* amr_list,
                         uint32_t* amr_count, uint32_t par, hipStream_t st);
// clock: device words {parity, global step} of the step being run, then those of the step run
// before it (kb_advance moves them on)
void launch_batch_ghost_update_r(const BoxFrame* F, const uint32_t* work, uint32_t nwork, const uint32_t* clock, hipStream_t st);
void launch_batch_advance(uint32_t* clock, hipStream_t st);
void launch_batch_first_half(const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t par, uint32_t gstep, hipStream_t st);
void launch_batch_trigger(const BoxFrame* F, const uint32_t* boxes, uint32_t n, uint32_t gstep, uint32_t* out, hipStream_t st);
hipError_t launch_batch_nbh(const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t T, uint32_t lds, uint32_t par,
                            hipStream_t st, bool set_lds = true);
// batched brick kernels: the force per precision and the neighbour build
enum { kBatchLjFp64 = 0, kBatchLjMixed = 1, kBatchNbh = 2 };
struct BatchKernelInfo {
  uint32_t static_lds;    // bytes of static LDS
  uint32_t waves_per_cu;  // waves one CU holds at the kernel's register use
};
BatchKernelInfo batch_kernel_info(int kind);
// set_lds false: the workgroup LDS attribute was already raised to lds (launches under capture)
hipError_t launch_batch_lj(bool mixed, const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t T, uint32_t lds,
                           const uint32_t* clock, hipStream_t st, bool set_lds = true);
hipError_t batch_set_lds(int kind, uint32_t lds);
constexpr uint32_t kBatchSlotBlock = 256;  // threads per workgroup of the per-slot / per-cell batched kernels

void scan_exclusive(const uint32_t* in, uint32_t n, uint32_t add, uint32_t* out, uint32_t* bsum, uint32_t* total, hipStream_t st);
uint32_t scan_max_items();

void launch_move_flag(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, uint32_t* mv_dst, uint32_t* cnt_out,
                      uint32_t* cnt_in, uint32_t* err, hipStream_t st);
void launch_move_pack(const Geom& g, const uint32_t* cell_off, const uint32_t* cell_n, const uint32_t* cnt_out,
                      const uint32_t* cnt_in, const uint32_t* mv_dst, uint32_t* pk, uint32_t* newloc, uint32_t* new_n,
                      hipStream_t st);
void launch_move_scatter(uint32_t N, const PState& A, const PState& B, const uint32_t* slot_cell, const uint32_t* mv_dst,
                         const uint32_t* newloc, const uint32_t* cell_off, const uint32_t* new_off, const uint32_t* cell_n,
                         const uint32_t* cnt_out, uint32_t* in_cur, uint32_t* in_list, uint32_t* slot_cell_B, bool forces,
                         hipStream_t st);
void launch_move_incoming(const Geom& g, const PState& A, const PState& B, const uint32_t* new_off, const uint32_t* cell_n,
                          const uint32_t* cnt_out, const uint32_t* cnt_in, uint32_t* in_list, uint32_t* slot_cell_B,
                          bool forces, hipStream_t st);
void launch_set_inner_cells(const Geom& g, const uint32_t* new_n, const uint32_t* new_off, uint32_t* cell_n, uint32_t* cell_off,
                            hipStream_t st);
// rebuild_amr: dense inner cells of A sorted by sub-cell, through B (spare) and back.  list: one
// word per cell; *count must be 0 on entry; a cell of more than 4096 particles sets bit 2 of
// *too_dense and stays unsorted
void launch_rebuild_amr(const Geom& g, double density, const PState& A, const PState& B, const uint32_t* cell_off,
                        const uint32_t* cell_n, uint32_t* list, uint32_t* count, uint32_t* too_dense, bool forces, hipStream_t st);
// sub_grid_density of rebuild_amr in the reference's update deck (update-particles.msp:8-11),
// used by the run loops and batches
constexpr double kSubGridDensity = 6.5;
void launch_ghost_count(const Geom& g, const double4* r, const uint32_t* cell_off, const uint32_t* cell_n, uint32_t* gcnt,
                        hipStream_t st);
void launch_ghost_fill(const Geom& g, uint32_t N, const PState& B, const uint32_t* goff, uint32_t* cell_off, uint32_t* cell_n,
                       uint32_t* slot_cell, uint32_t* ghost_src, uint32_t* ghost_shift, uint32_t cap, uint32_t* status, bool forces,
                       hipStream_t st);
void launch_ghost_update_r(const Geom& g, uint32_t N, uint32_t G, double4* r, const uint32_t* ghost_src, const uint32_t* ghost_shift,
                           hipStream_t st);
void launch_backup_r(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, uint32_t* backup, hipStream_t st);
void launch_displ_over(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, const uint32_t* backup, double thr2,
                       unsigned long long* count, hipStream_t st);
void launch_push_f_v_r(uint32_t N, const PState& A, double dt, double dt2, hipStream_t st);
void launch_push_f_v(uint32_t N, const PState& A, double dt, hipStream_t st);
void launch_verlet_first_half(const Geom& g, uint32_t N, const PState& A, double dt, double dt2, double dth,
                              const uint32_t* slot_cell, const uint32_t* backup, double thr2, unsigned long long* count,
                              hipStream_t st);
void launch_zero_force(uint32_t n, const PState& A, hipStream_t st);
void launch_divide_force_by_type(uint32_t n, const PState& A, const double* values, int ntypes, hipStream_t st);
void launch_nbh_build(const Geom& g, uint32_t n_atoms, const double4* r, const uint32_t* slot_cell, const uint32_t* cell_off,
                      const uint32_t* cell_n, double max2, int lay, int half_symmetric, int skip_ghosts, uint32_t* nbr,
                      uint32_t* nbr_cnt, uint32_t cap, uint32_t* max_cnt, hipStream_t st);
uint32_t lj_list_granule();
void launch_gravity_slots(uint32_t n_atoms, const PState& A, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                          const GravitationalForceFunctor& gf, double rc2, hipStream_t st);
hipError_t launch_gravity_brick(const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off, const uint32_t* cell_n,
                                const uint16_t* blist, const uint32_t* bcnt, const GravitationalForceFunctor& gf, double rc2,
                                hipStream_t st);
void launch_max_type(uint32_t n, const double4* r, uint32_t* out, hipStream_t st);
void launch_lj(int precision, bool accum, bool push, uint32_t n_atoms, const PState& A, const uint32_t* nbr,
               const uint32_t* nbr_cnt, uint32_t cap, double eps, double sigma, double rcut, double dth, hipStream_t st);
int launch_lj_multimat(int precision, uint32_t n_atoms, const PState& A, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                       int ntypes, const LJMulti& m, hipStream_t st);
void launch_pair_counts(uint32_t n_atoms, const double4* r, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap, double rc2,
                        unsigned long long* out, hipStream_t st);
void launch_stream_plen(uint32_t n_atoms, const uint32_t* slot_cell, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                        int random_access, uint32_t* plen, hipStream_t st);
void launch_stream_cell_len(const Geom& g, uint32_t n_atoms, const uint32_t* cell_off, const uint32_t* cell_n, const uint32_t* plen,
                            int ntables, uint32_t* pstart, uint32_t* clen, hipStream_t st);
void launch_stream_fill(const Geom& g, uint32_t n_atoms, const uint32_t* slot_cell, const uint32_t* cell_off, const uint32_t* cell_n,
                        const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap, const uint32_t* pstart, const uint32_t* plen,
                        const uint32_t* coff, int ntables, int random_access, uint16_t* out, hipStream_t st);
void launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t st);
void launch_fp64_core(const double* x, size_t n, double* sq, double* rc, hipStream_t st);

BrickParams brick_params(const Geom& g, int Bi, int Bj, int Bk);
// LDS layout of the brick kernels (offsets into bp), its bytes; f32_build: the prefiltered
// neighbour build's (FP32 positions; types only with f32_types, for per-pair radii)
uint32_t brick_layout(BrickParams& bp, bool f32_build = false, bool f32_types = false);
uint32_t brick_list_granule();
void launch_brick_stats(const Geom& g, const BrickParams& bp, const uint32_t* cell_n, uint32_t* out, hipStream_t st);
// prefilter: the unfiltered builds test candidates in FP32 with an FP64 band (the same lists)
hipError_t launch_nbh_brick(const Geom& g, BrickParams bp, const double4* r, const uint32_t* cell_off, const uint32_t* cell_n, double max2,
                            const PairCut& pc, int half_symmetric, int skip_ghosts, bool prefilter, uint16_t* blist, uint32_t* bcnt,
                            uint32_t* max_cnt, hipStream_t st);
hipError_t launch_lj_brick(int precision, bool accum, int epi, const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off,
                           const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, double eps, double sigma, double rcut,
                           double dth, const NextStep* next, hipStream_t st);
hipError_t launch_ljmm_brick(int precision, int ntypes, int epi, const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off,
                             const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, const LJMulti& mm, double dth,
                             const NextStep* next, hipStream_t st);
hipError_t launch_brick_to_slots(const Geom& g, BrickParams bp, const uint32_t* cell_off, const uint32_t* cell_n, const uint16_t* blist,
                                 const uint32_t* bcnt, uint32_t* nbr, uint32_t* nbr_cnt, uint32_t cap, hipStream_t st);
void launch_energies(uint32_t N, const PState& A, double* part, uint32_t nparts, double* out, hipStream_t st);

}  // namespace xnb
