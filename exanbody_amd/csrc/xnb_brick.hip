// xnb_brick.hip -- LDS-tiled neighbour build and LJ force for gfx950.
//
// The inner domain is tiled by bricks of Bi x Bj x Bk cells; one workgroup owns one brick.
// It stages the positions of the brick's region (the brick plus one cell on every side,
// ghost cells included) into LDS, region cells in k-major order, so a neighbour is addressed
// by its u16 LDS index.  Lanes are the brick's atoms in (cell k-major, particle) order.
//
// Neighbour lists are u16 LDS indices, one slab of cap entries x T lanes per brick.  Entries
// are grouped by four: entry q of lane t sits at ((q / 4) * T + t) * 4 + q % 4, so a lane
// reads four consecutive entries with one 8-byte load and a wave's loads are contiguous.
// Because region cells are k-major and candidate cells are visited in (dk, dj, di) order
// with particles in cell order, every list is in ascending (cell_enc, p_b) order -- the order
// of the reference stream (chunk_neighbors_execute.h:308) -- and the
// forces are summed in that order.
//
// Requires cell_layers == 1 (cell edge >= nbh_dist) and ghost layers >= 1.
#include <algorithm>
#include <mutex>
#include <vector>

#include "xnb_kernels.h"

namespace xnb {

// The brick kernels keep every LDS word in their dynamic block, so it starts at address 0 and
// a staged atom's address is one multiply of its index.

// Two exclusive scans in one pass (a in the low, b in the high 32 bits of a 64-bit sum; the
// low sums stay below 2^32): ea / eb the exclusive prefixes, ta / tb the workgroup totals.
// wsum: 32 words of LDS, not used again by the caller (so no barrier after its last read).
__device__ inline void block_scan_excl2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb, uint32_t& ta, uint32_t& tb,
                                        uint32_t* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned long long v = (unsigned long long)a | ((unsigned long long)b << 32);
  unsigned long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  unsigned long long* ws = reinterpret_cast<unsigned long long*>(wsum);
  if (lane == 63) ws[w] = x;
  __syncthreads();
  unsigned long long woff = 0, tot = 0;
  for (int i = 0; i < nw; i++) {
    const unsigned long long s = ws[i];
    if (i < w) woff += s;
    tot += s;
  }
  const unsigned long long ex = x - v + woff;
  ea = (uint32_t)ex;
  eb = (uint32_t)(ex >> 32);
  ta = (uint32_t)tot;
  tb = (uint32_t)(tot >> 32);
}

__device__ inline uint32_t wave_max_b(uint32_t v) { return wave_max_u32(v); }

// 2 m + (this lane's bit of lanes), one v_addc with the lane mask (a compare's result) as the
// carry in; the compiler's select-and-shift takes two instructions.
__device__ inline uint32_t shift_in(uint32_t m, unsigned long long lanes) {
  uint32_t r;
  unsigned long long carry_out;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(m), "s"(lanes));
  return r;
}

// largest k in [0, n) with pref[k] <= x (pref non-decreasing, pref[0] = 0 <= x)
__device__ inline uint32_t upper_index(const uint32_t* pref, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;  // invariant: pref[lo] <= x, answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pref[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

static constexpr int LJB_UNROLL_B = (int)kListBatch;  // list entries per force batch
// sentinel atoms staged after a region's atoms: index R is the force kernels' list padding;
// the build reads whole candidate chunks of 32 from a row start and needs 32 of them
static constexpr uint32_t kSentinels = 32;
static constexpr int32_t kNoDelta = INT32_MIN;  // a region row whose cells' slots do not follow each other
// region atoms a thread stages per memory round.  A 4x4x2 brick's region holds about 4.2
// times its own atoms (the workgroup's lanes), so four a round leave a second round trip for
// the last few; five or six a round measured the same for the mixed force kernel (other
// workgroups hide the round trip), 2 % faster for the FP64 one (six) and 8 % slower for the
// prefiltered build (six, more registers).
#ifndef XNB_STAGE_ROUND
#define XNB_STAGE_ROUND 4
#endif
static constexpr int kStageRound = XNB_STAGE_ROUND;

// One group of four list entries (8 bytes), loaded past the caches: lists are read once per
// force launch, so streaming them keeps L2 and the MALL for the staged halo positions that
// neighbouring bricks read again (with the fused epilogue's streaming stores: C2 mixed force
// 0.937 -> 0.921 ms, slab multimat force 1.270 -> 1.256 ms mixed, 2.542 -> 2.516 ms FP64;
// two runs each).  The neighbour build's list stores stay cached: streamed, its scattered
// 8-byte stores took the C2 build from 2.5 to 6.9 ms.
__device__ inline uint2 load_list_group(const uint2* p) {
  const unsigned long long q = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
  return make_uint2((uint32_t)q, (uint32_t)(q >> 32));
}

// position of entry q of lane t in a brick's list slab (T lanes)
__device__ inline size_t lpos(uint32_t q, uint32_t t, uint32_t T) { return ((size_t)(q >> 2) * T + t) * 4 + (q & 3u); }

struct BrickLds {
  uint32_t *rcnt, *rpref, *roff, *bpref;
  int32_t* rdel;    // per region row: slot - LDS index of its atoms, or kNoDelta
  uint16_t* crow;   // per 64 region atoms: the row holding the first of them
  double *px, *py, *pz;
  float4* pf;   // kPosF32 staging (same LDS bytes as px): atom i's position relative to the brick
                // centre, and in w its squared norm (both FP32)
  uint8_t* pt;  // particle type (r.w) of each region atom
  uint32_t* wsum;  // scan scratch (32 words)
  uint32_t* wcnt;  // wave_group_count_add's two words
};

__device__ inline BrickLds brick_lds(const BrickParams& bp, char* lds) {
  BrickLds L;
  L.rcnt = (uint32_t*)(lds + bp.off_rcnt);
  L.rpref = (uint32_t*)(lds + bp.off_rpref);
  L.roff = (uint32_t*)(lds + bp.off_roff);
  L.bpref = (uint32_t*)(lds + bp.off_bpref);
  L.rdel = (int32_t*)(lds + bp.off_rdel);
  L.crow = (uint16_t*)(lds + bp.off_crow);
  L.px = (double*)lds;  // at off_pos == 0; interleaved x, y, z: atom i at [3 i, 3 i + 2]
  L.py = L.px + 1;
  L.pz = L.px + 2;
  L.pf = (float4*)lds;
  L.pt = (uint8_t*)(lds + bp.off_type);
  L.wsum = (uint32_t*)(lds + bp.off_wsum);
  L.wcnt = L.wsum + 32;
  return L;
}

// Dispatch position d -> brick index, in pencils: slabs of kTileZ brick layers, each cut into
// groups of kTileY brick rows; a pencil (one slab x one row group, all x) is walked x-major
// with (y, z) inside, so any kTileY * kTileZ * 4 consecutive positions are a compact 4 x 4 x 4
// block of bricks whose staged regions overlap (halo reads hit L2).  A bijection of
// [0, nbricks); only the edge slab and edge row group are thinner.
static constexpr uint32_t kTileY = 4, kTileZ = 4;
__host__ __device__ inline uint32_t brick_at(const BrickParams& bp, uint32_t d) {
  const uint32_t nx = (uint32_t)bp.nbx, ny = (uint32_t)bp.nby, nz = (uint32_t)bp.nbz;
  const uint32_t sl = d / (nx * ny * kTileZ);
  const uint32_t r1 = d - sl * nx * ny * kTileZ;
  const uint32_t tz = min(kTileZ, nz - sl * kTileZ);
  const uint32_t gy = r1 / (nx * kTileY * tz);
  const uint32_t r2 = r1 - gy * nx * kTileY * tz;
  const uint32_t ty = min(kTileY, ny - gy * kTileY);
  const uint32_t z = r2 % tz, y = (r2 / tz) % ty, x = r2 / (tz * ty);
  return ((sl * kTileZ + z) * ny + gy * kTileY + y) * nx + x;
}

// brick of this workgroup: workgroups go to the 8 XCDs round-robin, each XCD walks a
// contiguous run of dispatch positions, in pencil order (PENCIL) or in brick index order;
// >= nbricks for the few padding workgroups.  Pencils cut the mixed force kernel's HBM reads
// from 2.63 to 2.40 GB per C2 launch and the build's from 0.71 to 0.47 GB; the FP64 force
// kernel (one workgroup per CU) measured 3 % slower with them, so it keeps index order.
template <bool PENCIL>
__device__ inline uint32_t brick_of_block(const BrickParams& bp) {
  const uint32_t d = (blockIdx.x % kXcds) * bp.per_xcd + blockIdx.x / kXcds;
  return PENCIL && d < bp.nbricks ? brick_at(bp, d) : d;
}

struct BrickPos {
  int gi0, gj0, gk0;  // grid coordinates of the brick's first cell
  int ni, nj, nk;     // brick extent clipped to the inner domain
};

__device__ inline BrickPos brick_pos(const Geom& g, const BrickParams& bp, uint32_t b) {
  BrickPos P;
  const int bx = (int)(b % (uint32_t)bp.nbx);
  const int by = (int)((b / (uint32_t)bp.nbx) % (uint32_t)bp.nby);
  const int bz = (int)(b / ((uint32_t)bp.nbx * (uint32_t)bp.nby));
  P.gi0 = g.gl + bx * bp.Bi;
  P.gj0 = g.gl + by * bp.Bj;
  P.gk0 = g.gl + bz * bp.Bk;
  P.ni = min(bp.Bi, g.Di - bx * bp.Bi);
  P.nj = min(bp.Bj, g.Dj - by * bp.Bj);
  P.nk = min(bp.Bk, g.Dk - bz * bp.Bk);
  return P;
}

// What brick_setup stages besides the region tables: nothing, the FP64 positions and types
// (px/py/pz, pt), or for the neighbour build's prefilter FP32 positions relative to the brick
// centre (pf, 16 bytes an atom, in the bytes of px; kPosF32N adds each one's FP32 squared norm
// in w and the types), with the largest magnitude of a staged FP32 coordinate (as float bits)
// in wcnt[0].
enum { kPosNone = 0, kPosF64 = 1, kPosF32 = 2, kPosF32N = 3 };

// FP64 reference point of a brick's FP32 staging: the centre of the brick's cells.
__device__ inline void brick_centre(const Geom& g, const BrickPos& P, double& cx, double& cy, double& cz) {
  cx = cell_origin(g.ox, g.gl, P.gi0, g.cs) + 0.5 * (double)P.ni * g.cs;
  cy = cell_origin(g.oy, g.gl, P.gj0, g.cs) + 0.5 * (double)P.nj * g.cs;
  cz = cell_origin(g.oz, g.gl, P.gk0, g.cs) + 0.5 * (double)P.nk * g.cs;
}

// Slot of region atom idx < R, from the staging tables (row table, then cell by cell in a row
// of several deltas).
__device__ inline uint32_t region_slot(const BrickParams& bp, const BrickLds& L, uint32_t idx) {
  const uint32_t RX = (uint32_t)bp.RX;
  uint32_t q = L.crow[idx >> 6];
  while (L.rpref[(q + 1u) * RX] <= idx) q++;
  const int32_t d = L.rdel[q];
  if (d != kNoDelta) return (uint32_t)((int32_t)idx + d);
  uint32_t c = q * RX;
  while (L.rpref[c + 1u] <= idx) c++;
  return L.roff[c] + (idx - L.rpref[c]);
}

// Region tables (+ positions per pos_mode) into LDS.  Needs nrc <= blockDim.x and
// nbc <= blockDim.x.  Ends with a workgroup barrier.  Returns false (the same for every
// thread, with bp.status flagged) when the region holds more than Rmax atoms or the brick
// more than T, the sizes its LDS and lists were laid out for.
__device__ inline bool brick_setup(const Geom& g, const BrickParams& bp, const BrickPos& P, const BrickLds& L,
                                   const double4* __restrict__ r, const uint32_t* __restrict__ cell_off,
                                   const uint32_t* __restrict__ cell_n, int pos_mode) {
  const uint32_t t = threadIdx.x;
  const bool f32 = pos_mode == kPosF32 || pos_mode == kPosF32N;
  if (f32 && t == 0) L.wcnt[0] = 0u;  // ordered before the max by the scan's barrier
  uint32_t cnt = 0;
  if (t < (uint32_t)bp.nrc) {
    const int ri = (int)(t % (uint32_t)bp.RX), rj = (int)((t / (uint32_t)bp.RX) % (uint32_t)bp.RY),
              rk = (int)(t / ((uint32_t)bp.RX * (uint32_t)bp.RY));
    const int gi = P.gi0 - 1 + ri, gj = P.gj0 - 1 + rj, gk = P.gk0 - 1 + rk;
    const bool in = ri <= P.ni + 1 && rj <= P.nj + 1 && rk <= P.nk + 1 && gi < g.di && gj < g.dj && gk < g.dk;
    uint32_t off = 0;
    if (in) {
      const uint32_t c = cell_index(g, gi, gj, gk);
      cnt = cell_n[c];
      off = cell_off[c];
    }
    L.rcnt[t] = cnt;
    L.roff[t] = off;
  }
  // brick cells bc = (ck*Bj + cj)*Bi + ci (region cell ((ck+1)*RY + cj+1)*RX + ci+1), read in
  // the same memory round, and both prefix sums in one scan
  uint32_t bcnt = 0;
  if (t < (uint32_t)bp.nbc) {
    const int ci = (int)(t % (uint32_t)bp.Bi), cj = (int)((t / (uint32_t)bp.Bi) % (uint32_t)bp.Bj),
              ck = (int)(t / ((uint32_t)bp.Bi * (uint32_t)bp.Bj));
    if (ci < P.ni && cj < P.nj && ck < P.nk) bcnt = cell_n[cell_index(g, P.gi0 + ci, P.gj0 + cj, P.gk0 + ck)];
  }
  uint32_t ex, bex, rtot, btot;
  block_scan_excl2(cnt, bcnt, ex, bex, rtot, btot, L.wsum);
  if (t < (uint32_t)bp.nrc) L.rpref[t] = ex;
  if (t < (uint32_t)bp.nbc) L.bpref[t] = bex;
  if (t == 0) {
    L.rpref[bp.nrc] = rtot;
    L.bpref[bp.nbc] = btot;
  }
  __syncthreads();
  const bool fits_r = rtot <= bp.Rmax, fits_t = btot <= bp.T;
  if (!fits_r || !fits_t) {
    if (t == 0) atomicOr(bp.status, (fits_r ? 0u : 1u) | (fits_t ? 0u : 2u));
    return false;
  }
  if (pos_mode != kPosNone) {
    const uint32_t R = L.rpref[bp.nrc];
    const uint32_t RX = (uint32_t)bp.RX, nrow = (uint32_t)bp.nrc / RX;
    // Row table.  Region rows are RX cells along x; in a row of inner cells the cells' slots
    // follow each other, so every atom of the row sits at slot = LDS index + delta.  A row
    // holding ghost cells (several deltas) is looked up cell by cell.  crow[k] is the row
    // holding region atom 64 k.
    if (t < nrow) {
      const uint32_t c0 = t * RX;
      int32_t d0 = 0;
      bool any = false, uni = true;
      for (uint32_t c = c0; c < c0 + RX; c++) {
        if (!L.rcnt[c]) continue;
        const int32_t d = (int32_t)(L.roff[c] - L.rpref[c]);
        if (!any) d0 = d;
        uni = uni && (!any || d == d0);
        any = true;
      }
      L.rdel[t] = uni ? d0 : kNoDelta;
      const uint32_t s1 = L.rpref[c0 + RX];
      for (uint32_t k = (L.rpref[c0] + 63u) >> 6; (k << 6) < s1; k++) L.crow[k] = (uint16_t)t;
    }
    __syncthreads();
    double cx = 0., cy = 0., cz = 0.;
    if (f32) brick_centre(g, P, cx, cy, cz);
    uint32_t cmax = 0;  // largest |staged FP32 coordinate| of this thread, as bits
    // Up to kStageRound region atoms per thread and round, every load issued before the first
    // LDS store: one memory round trip per round (a loop that stored each load before issuing
    // the next spent ~0.36 ms of a 0.94 ms mixed force launch at 8.4M atoms staging).
    for (uint32_t base = t; base < R; base += (uint32_t)kStageRound * blockDim.x) {
      double4 p[kStageRound];
#pragma unroll
      for (int u = 0; u < kStageRound; u++) {
        const uint32_t idx = base + (uint32_t)u * blockDim.x;
        if (idx < R) p[u] = r[region_slot(bp, L, idx)];
      }
      if (f32) {
#pragma unroll
        for (int u = 0; u < kStageRound; u++) {
          const uint32_t idx = base + (uint32_t)u * blockDim.x;
          if (idx < R) {
            const float fx = (float)(p[u].x - cx), fy = (float)(p[u].y - cy), fz = (float)(p[u].z - cz);
            const float4 c = make_float4(fx, fy, fz, pos_mode == kPosF32N ? fmaf(fx, fx, fmaf(fy, fy, fz * fz)) : 0.f);
            L.pf[idx] = c;
            if (pos_mode == kPosF32N) L.pt[idx] = (uint8_t)p[u].w;
            // NaN and infinities compare above every finite magnitude as bits too
            cmax = max(cmax, max(__float_as_uint(fabsf(c.x)), max(__float_as_uint(fabsf(c.y)), __float_as_uint(fabsf(c.z)))));
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < kStageRound; u++) {
          const uint32_t idx = base + (uint32_t)u * blockDim.x;
          if (idx < R) {
            L.px[3 * idx] = p[u].x;
            L.py[3 * idx] = p[u].y;
            L.pz[3 * idx] = p[u].z;
            L.pt[idx] = (uint8_t)p[u].w;
          }
        }
      }
    }
    if (t < kSentinels) {  // sentinel atoms R .. R + kSentinels - 1: every distance to them exceeds any cutoff
      if (f32) {
        L.pf[R + t] = make_float4(1e30f, 1e30f, 1e30f, __builtin_inff());
        if (pos_mode == kPosF32N) L.pt[R + t] = 0;
      } else {
        L.px[3 * (R + t)] = 1e30;
        L.py[3 * (R + t)] = 1e30;
        L.pz[3 * (R + t)] = 1e30;
        L.pt[R + t] = 0;
      }
    }
    if (f32 && cmax) __hip_atomic_fetch_max(&L.wcnt[0], cmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
  }
  return true;
}

struct LaneAtom {
  bool live;
  uint32_t rc;    // region cell
  uint32_t p;     // index in cell
  uint32_t ia;    // LDS index
  uint32_t slot;  // global slot
};

__device__ inline LaneAtom lane_atom(const BrickParams& bp, const BrickLds& L) {
  LaneAtom A;
  const uint32_t t = threadIdx.x;
  A.live = t < L.bpref[bp.nbc];
  if (!A.live) {
    A.rc = 0;
    A.p = 0;
    A.ia = 0;
    A.slot = 0;
    return A;
  }
  const uint32_t bc = upper_index(L.bpref, (uint32_t)bp.nbc, t);
  const int ci = (int)(bc % (uint32_t)bp.Bi), cj = (int)((bc / (uint32_t)bp.Bi) % (uint32_t)bp.Bj),
            ck = (int)(bc / ((uint32_t)bp.Bi * (uint32_t)bp.Bj));
  A.rc = (uint32_t)(((ck + 1) * bp.RY + cj + 1) * bp.RX + ci + 1);
  A.p = t - L.bpref[bc];
  A.ia = L.rpref[A.rc] + A.p;
  A.slot = L.roff[A.rc] + A.p;
  return A;
}

// ------------------------------------------------------------------ statistics
// per brick: atoms in the brick and in its region -> maxima
__global__ void k_brick_stats(Geom g, BrickParams bp, const uint32_t* __restrict__ cell_n, uint32_t* __restrict__ out) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bp.nbricks) return;
  const BrickPos P = brick_pos(g, bp, b);
  uint32_t R = 0, nb = 0;
  for (int rk = 0; rk <= P.nk + 1; rk++)
    for (int rj = 0; rj <= P.nj + 1; rj++)
      for (int ri = 0; ri <= P.ni + 1; ri++) {
        const int gi = P.gi0 - 1 + ri, gj = P.gj0 - 1 + rj, gk = P.gk0 - 1 + rk;
        if (gi >= g.di || gj >= g.dj || gk >= g.dk) continue;
        const uint32_t n = cell_n[cell_index(g, gi, gj, gk)];
        R += n;
        if (ri >= 1 && ri <= P.ni && rj >= 1 && rj <= P.nj && rk >= 1 && rk <= P.nk) nb += n;
      }
  atomicMax(&out[0], nb);
  atomicMax(&out[1], R);
}

// ------------------------------------------------------------------ build
// chunk_neighbors (chunk_neighbors_execute.h:221-239): d = r_a - r_b, FP64, kept iff
// (a != b) && 0 < d2 <= nbh_dist^2, then the NeighborFilterHalfSymGhost filters.
// Candidates are tested in chunks of 32 without branches into a bit mask, then each lane
// expands its mask in candidate order, which appends entries in candidate order.  An
// overflowing list (count > cap) is rebuilt with a larger cap, so what it stored past the
// point of overflow does not matter.  FILTER enables the half_symmetric / skip_ghosts filters
// (a per-cell loop; the unfiltered build walks rows of three cells).
// Brick b of one box; the workgroup may be wider than bp.T (batched launches), lanes from
// bp.T on hold no atom and touch no list.
// PAIRWISE (multi-material runs): a candidate within nbh_dist is stored only when it is also
// within its type pair's list radius (PairCut); its d2 is recomputed as in the test.
// PREF (unfiltered builds): the tests run in FP32 on positions relative to the brick centre
// and decide every candidate whose FP32 d2 lies outside an error band around its threshold
// (or above a bound near 0); band candidates are tested again in FP64 from the global
// positions, so the lists are those of the FP64 test.  With PAIRWISE the chunk test keeps a
// superset at the atom's largest pair radius and each candidate is decided against its own
// pair's radius while the mask is expanded.
template <bool FILTER, bool PAIRWISE = false, bool PREF = false>
__device__ __forceinline__ void nbh_brick_body(const Geom& g, const BrickParams& bp, uint32_t b, const double4* __restrict__ r,
                                               const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                               double max2, int half_symmetric, int skip_ghosts, uint16_t* __restrict__ blist,
                                               uint32_t* __restrict__ bcnt, uint32_t* __restrict__ max_cnt,
                                               const PairCut* pc = nullptr) {
  static_assert(!PREF || !FILTER, "the FP32 prefilter serves the unfiltered builds");
  extern __shared__ __align__(16) char lds[];
  const BrickLds L = brick_lds(bp, lds);
  const BrickPos P = brick_pos(g, bp, b);
  if (!brick_setup(g, bp, P, L, r, cell_off, cell_n, PREF ? (PAIRWISE ? kPosF32N : kPosF32) : kPosF64)) return;
  const LaneAtom A = lane_atom(bp, L);
  const uint32_t t = threadIdx.x, T = bp.T;
  uint16_t* out = blist + (size_t)b * bp.cap * T;
  // every lane, live or not, tests against its own LDS atom when past the candidate cell's
  // end (A.ia = 0 for lanes without an atom), so d2 == 0 rejects those candidates
  double ax = 0., ay = 0., az = 0.;
  if (!PREF) {
    ax = L.px[3 * A.ia];
    ay = L.py[3 * A.ia];
    az = L.pz[3 * A.ia];
  }
  uint32_t ca = 0;
  if (A.live) {
    if (FILTER) {
      const int ri = (int)(A.rc % (uint32_t)bp.RX), rj = (int)((A.rc / (uint32_t)bp.RX) % (uint32_t)bp.RY),
                rk = (int)(A.rc / ((uint32_t)bp.RX * (uint32_t)bp.RY));
      ca = cell_index(g, P.gi0 - 1 + ri, P.gj0 - 1 + rj, P.gk0 - 1 + rk);
    }
  }
  const uint32_t last = bp.cap - 1;
  uint32_t cnt = 0;
  double thr[4] = {max2, max2, max2, max2};  // this atom's list radius^2 per neighbour type
  // the chunk test's radius^2: the largest this atom needs (its pair radii never exceed
  // nbh_dist), so candidates beyond every radius of its type are not expanded and re-tested
  double cut2 = max2;
  if (PAIRWISE) {
    const unsigned ta = min((unsigned)L.pt[A.ia], 3u);
    cut2 = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      thr[q] = q < pc->nt ? pc->thr2[ta][q] : max2;
      if (q < pc->nt) cut2 = fmax(cut2, thr[q]);
    }
    cut2 = fmin(cut2, max2);
  }
  if constexpr (PREF) {
    // Error bands of the FP32 tests.  Staged coordinates are a_f = fl32(fl64(x - centre)) with
    // |.| <= C (the brick's largest, widened by 2^-20 for the rounding of its bits), and
    // w = fl32(|a_f|^2) (a product and two fmas, off by <= 6 u C^2; u = 2^-24).
    //  * Superset chunk test of the per-pair-radius build, dot form:
    //    t = fma(-2 a_x, b_x, fma(-2 a_y, b_y, fma(-2 a_z, b_z, w_b)))
    //    is |b_f|^2 - 2 a_f.b_f + w_b - |b_f|^2 plus three roundings (<= 21 u C^2), so
    //    D = t + w_a is |a_f - b_f|^2 within 33 u C^2, and |a_f - b_f|^2 is s^2 (s the true
    //    distance) within 2 sqrt(3) s E2 + 3 E2^2, E2 = 2.01 u C (the two coordinate
    //    roundings).  The threshold T - w_a is formed in FP64 and rounded outward, so t below
    //    it proves s^2 < T and t above T + H2 - w_a proves s^2 > T, with
    //    H2 = 2 (45 u C^2 + 2 sqrt(3) sqrt(T) E2 + 3 E2^2) (both sides of the difference
    //    grow with s slower than s^2 - T for s^2 >= T).
    //  * Difference form (the plain chunk test, and the per-pair decision while expanding):
    //    dx = a_f - b_f in FP32, then d2f = fma(dz, dz, fma(dy, dy, dx dx)), off from s^2 by at
    //    most 2 sqrt(3) s E + 3 E^2 + 3.03 u s^2 with E = 4.01 u C, so d2f < T - H proves
    //    s^2 < T and d2f > T + H proves s^2 > T with H = 2 (2 sqrt(3) sqrt(T) E + 3 E^2 + 4 u T);
    //    d2f > Z = 16 E^2 proves s > 2.2 E, far from FP64 underflow, so the FP64 d2 is > 0.
    // The FP64 d2 differs from s^2 by a relative 2^-51 at most, far inside the factor 2 of
    // each band.  Non-finite positions give unbounded bands: every candidate is tested in FP64.
    const float cf = __uint_as_float(L.wcnt[0]);
    const bool finite = cf <= 3.0e38f;
    const double u = 0x1p-24;
    const double C = (double)cf * (1.0 + 0x1p-20);
    const double E = 4.01 * u * C, E2 = 2.01 * u * C;
    auto band_lo = [&](double T2) {
      const double H = 2.0 * (2.0 * 1.7320508075688772 * sqrt(T2) * E + 3.0 * E * E + 4.0 * u * T2);
      return finite ? __double2float_rd(T2 - H) : -1.f;
    };
    auto band_hi = [&](double T2) {
      const double H = 2.0 * (2.0 * 1.7320508075688772 * sqrt(T2) * E + 3.0 * E * E + 4.0 * u * T2);
      return finite ? __double2float_ru(T2 + H) : __builtin_inff();
    };
    const float zf = finite ? __double2float_ru(16.0 * E * E) : 0.f;
    const float4 af = L.pf[A.ia];
    const double wa = (double)af.w;
    const double H2 = 2.0 * (45.0 * u * C * C + 2.0 * 1.7320508075688772 * sqrt(cut2) * E2 + 3.0 * E2 * E2);
    // The plain build decides at nbh_dist in the difference form (12 bytes of LDS and 11 VALU
    // instructions a candidate; the dot form, 16 bytes and 8, measured the same: the loop
    // then waits on LDS).  A per-pair-radius build keeps a superset at this atom's largest
    // radius in the dot form (5 VALU instructions a candidate, 4.5 % faster on the slab) and
    // decides each candidate against its own pair's radius while expanding.
    const float lo = band_lo(cut2), hi = band_hi(cut2);
    const float hiA = finite ? __double2float_ru((cut2 + H2) - wa) : __builtin_inff();
    const float ax2 = -2.f * af.x, ay2 = -2.f * af.y, az2 = -2.f * af.z;
    float tlo[4], thi[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      tlo[q] = PAIRWISE ? band_lo(thr[q]) : 0.f;
      thi[q] = PAIRWISE ? band_hi(thr[q]) : 0.f;
    }
    const uint32_t R = L.rpref[bp.nrc];
    // Entries go through a 64-bit shift register (four u16, the oldest lowest) and are stored a
    // group of four at a time: group j of the lane is the 8 bytes at grp[j * T].
    uint2* const grp = reinterpret_cast<uint2*>(out) + t;
    uint32_t plo = 0, phi = 0;
    double4 pa;  // the lane's FP64 position, read for its first FP64 test
    bool have_pa = false;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++) {
        const int rc0 = (int)A.rc + (dk * bp.RY + dj) * bp.RX - 1;
        const uint32_t rbase = A.live ? L.rpref[rc0] : R;
        const uint32_t rlen = A.live ? L.rpref[rc0 + 3] - rbase : 0u;
        const uint32_t lmax = wave_max_b(rlen);
        for (uint32_t p0 = 0; p0 < lmax; p0 += 32) {
          const uint32_t np = min((lmax - p0 + 3) & ~3u, 32u);
          const uint32_t lb = p0 < rlen ? rbase + p0 : R;
          const float4* pp = L.pf + lb;
          uint32_t mhi = 0, mlo = 0;  // bit 31 - k: candidate k within the chunk radius + H2; surely kept
          for (uint32_t q = 0; q < np; q += 4) {
#pragma unroll
            for (uint32_t w = 0; w < 4; w++) {
              const float4 c = pp[q + w];
              if constexpr (PAIRWISE) {
                const float tt = fmaf(ax2, c.x, fmaf(ay2, c.y, fmaf(az2, c.z, c.w)));
                mhi = shift_in(mhi, __builtin_amdgcn_ballot_w64(tt <= hiA));
              } else {
                const float dx = af.x - c.x, dy = af.y - c.y, dz = af.z - c.z;
                const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                mhi = shift_in(mhi, __builtin_amdgcn_ballot_w64(d2 <= hi));
                mlo = shift_in(mlo, __builtin_amdgcn_ballot_w64(d2 > zf) & __builtin_amdgcn_ballot_w64(d2 < lo));
              }
            }
          }
          if (np < 32) {
            mhi <<= 32 - np;
            mlo <<= 32 - np;
          }
          const uint32_t own = p0 < rlen ? rlen - p0 : 0u;
          if (own < 32) {
            const uint32_t keep = ~(~0u >> own);
            mhi &= keep;
            mlo &= keep;
          }
          const uint32_t sp = A.ia - lb;  // the lane's own atom (d2 == 0, never kept)
          if (sp < 32u) mhi &= ~(0x80000000u >> sp);
          uint32_t m = PAIRWISE ? mhi : mlo;
          if (!PAIRWISE) {
            uint32_t band = mhi & ~mlo;
            if (__builtin_expect(__any(band != 0u), 0)) {
              if (band) {
                if (!have_pa) pa = r[A.slot];
                have_pa = true;
                while (band) {
                  const uint32_t k = (uint32_t)__builtin_clz(band);
                  band ^= 0x80000000u >> k;
                  const double4 pb = r[region_slot(bp, L, lb + k)];
                  const double dx = pa.x - pb.x;
                  const double dy = pa.y - pb.y;
                  const double dz = pa.z - pb.z;
                  const double d2 = (dx * dx + dy * dy) + dz * dz;
                  if (d2 > 0.0 && d2 <= max2) m |= 0x80000000u >> k;
                }
              }
            }
            // a list that would pass cap stores nothing more (its count still grows, so the
            // caller rebuilds with a larger cap)
            const uint32_t pcnt = (uint32_t)__popc(m);
            if (cnt + pcnt > bp.cap) {
              cnt += pcnt;
              m = 0;
            }
          }
          while (m) {
            const uint32_t k = (uint32_t)__builtin_clz(m);
            m ^= 0x80000000u >> k;
            if (PAIRWISE) {
              // kept iff 0 < d2 <= the radius^2 of this pair (types above 3 use type 0's)
              const uint32_t ib = lb + k;
              const float4 c = L.pf[ib];
              const float dx = af.x - c.x, dy = af.y - c.y, dz = af.z - c.z;
              const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
              const uint32_t tb = L.pt[ib];
              float tl = tlo[0], th = thi[0];
#pragma unroll
              for (int q = 1; q < 4; q++) {
                tl = tb == (uint32_t)q ? tlo[q] : tl;
                th = tb == (uint32_t)q ? thi[q] : th;
              }
              bool keep = d2 > zf && d2 < tl;
              if (!keep && d2 <= th) {
                if (!have_pa) pa = r[A.slot];
                have_pa = true;
                const double4 pb = r[region_slot(bp, L, ib)];
                const double ex = pa.x - pb.x;
                const double ey = pa.y - pb.y;
                const double ez = pa.z - pb.z;
                const double e2 = (ex * ex + ey * ey) + ez * ez;
                double lim = thr[0];
#pragma unroll
                for (int q = 1; q < 4; q++) lim = tb == (uint32_t)q ? thr[q] : lim;
                keep = e2 > 0.0 && e2 <= lim;
              }
              if (!keep) continue;
            }
            plo = __builtin_amdgcn_alignbit(phi, plo, 16);
            phi = __builtin_amdgcn_alignbit(lb + k, phi, 16);
            ++cnt;
            if ((cnt & 3u) == 0u && cnt <= bp.cap) grp[(size_t)((cnt >> 2) - 1u) * T] = make_uint2(plo, phi);
          }
        }
      }
    // close the lane's list: the partial group and whole groups of the sentinel up to the
    // wave's batch boundary, so force kernels need no per-lane bound inside a batch
    if (!A.live) cnt = 0;
    if (t < T) bcnt[(size_t)b * T + t] = cnt;
    const uint32_t wmax = wave_max_b(cnt);
    if ((t & 63) == 0 && wmax > 0) atomicMax(max_cnt, wmax);
    const uint32_t pad_end = min((wmax + kListBatch - 1) / kListBatch * kListBatch, bp.cap);
    const uint32_t sentinel = L.rpref[bp.nrc];
    uint32_t q = min(cnt, bp.cap);
    if (q < pad_end && (q & 3u)) {
      while (q & 3u) {
        plo = __builtin_amdgcn_alignbit(phi, plo, 16);
        phi = __builtin_amdgcn_alignbit(sentinel, phi, 16);
        ++q;
      }
      grp[(size_t)((q >> 2) - 1u) * T] = make_uint2(plo, phi);
    }
    const uint32_t s2 = sentinel | (sentinel << 16);
    for (; q < pad_end; q += 4) grp[(size_t)(q >> 2) * T] = make_uint2(s2, s2);
    return;
  } else if constexpr (!FILTER) {
    // Rows: the cells di = -1, 0, 1 of one (dk, dj) are consecutive in LDS, so a row is one
    // candidate range [rbase, rbase + rlen).  A chunk tests 4-aligned candidates p0 + k
    // (k < np, np uniform) against every lane; bit 31 - k of m is candidate k.  A lane whose
    // row is exhausted reads the sentinel block instead.  Reads stay below R + 32: a lane with
    // p0 < rlen reads up to rbase + p0 + 31 <= R + 30.  Candidates past the lane's own rlen
    // are masked off after the chunk.  A lane whose list would pass cap stores nothing more
    // (its count still grows, so the caller rebuilds with a larger cap).
    const uint32_t R = L.rpref[bp.nrc];
    // byte offset of entry cnt in the slab, 2 * lpos(cnt, t, T), advanced per stored entry
    char* const ob = reinterpret_cast<char*>(out);
    uint32_t wo = 8u * t;
    const uint32_t wstep = 8u * T - 6u;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++) {
        const int rc0 = (int)A.rc + (dk * bp.RY + dj) * bp.RX - 1;
        const uint32_t rbase = A.live ? L.rpref[rc0] : R;
        const uint32_t rlen = A.live ? L.rpref[rc0 + 3] - rbase : 0u;
        const uint32_t lmax = wave_max_b(rlen);
        for (uint32_t p0 = 0; p0 < lmax; p0 += 32) {
          const uint32_t np = min((lmax - p0 + 3) & ~3u, 32u);
          const uint32_t lb = p0 < rlen ? rbase + p0 : R;
          const double* pp = L.px + 3 * lb;
          uint32_t m = 0;
          for (uint32_t q = 0; q < np; q += 4) {
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
              const double dx = ax - pp[3 * (q + u)];
              const double dy = ay - pp[3 * (q + u) + 1];
              const double dz = az - pp[3 * (q + u) + 2];
              const double d2 = (dx * dx + dy * dy) + dz * dz;
              m = m + m + ((d2 > 0.0 && d2 <= (PAIRWISE ? cut2 : max2)) ? 1u : 0u);
            }
          }
          if (np < 32) m <<= 32 - np;
          const uint32_t own = p0 < rlen ? rlen - p0 : 0u;
          if (own < 32) m &= ~(~0u >> own);
          const uint32_t pc = (uint32_t)__popc(m);
          if (cnt + pc > bp.cap) {
            cnt += pc;
            m = 0;
          }
          while (m) {
            const uint32_t k = (uint32_t)__builtin_clz(m);
            m ^= 0x80000000u >> k;
            if (PAIRWISE) {
              const uint32_t ib = lb + k;
              const double dx = ax - L.px[3 * ib];
              const double dy = ay - L.py[3 * ib];
              const double dz = az - L.pz[3 * ib];
              const double d2 = (dx * dx + dy * dy) + dz * dz;
              const unsigned tb = L.pt[ib];
              double lim = thr[0];
#pragma unroll
              for (int q = 1; q < 4; q++) lim = tb == (unsigned)q ? thr[q] : lim;
              if (!(d2 <= lim)) continue;
            }
            *reinterpret_cast<uint16_t*>(ob + wo) = (uint16_t)(lb + k);
            ++cnt;
            wo += (cnt & 3u) ? 2u : wstep;
          }
        }
      }
  } else {
  for (int dk = -1; dk <= 1; dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        const int rcb = (int)A.rc + (dk * bp.RY + dj) * bp.RX + di;
        const uint32_t nb = A.live ? L.rcnt[rcb] : 0u;
        const uint32_t base = A.live ? L.rpref[rcb] : 0u;
        const uint32_t nbmax = wave_max_b(nb);
        bool drop_cell = false;
        uint32_t cb = 0;
        if (FILTER) {
          cb = ca + (uint32_t)((dk * g.dj + dj) * g.di + di);
          if (skip_ghosts && A.live) {
            int ii, jj, kk;
            cell_ijk(g, cb, ii, jj, kk);
            drop_cell = is_ghost(g, ii, jj, kk);
          }
          if (half_symmetric && ca < cb) drop_cell = true;
        }
        for (uint32_t p0 = 0; p0 < nbmax; p0 += 32) {
          const uint32_t pend = min(nbmax - p0, 32u);
          uint32_t m = 0;
          for (uint32_t q = 0; q < pend; q += 4) {
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
              const uint32_t p = p0 + q + u;
              const uint32_t ib = p < nb ? base + p : A.ia;  // own index: d2 == 0 is rejected
              const double dx = ax - L.px[3 * ib];
              const double dy = ay - L.py[3 * ib];
              const double dz = az - L.pz[3 * ib];
              const double d2 = (dx * dx + dy * dy) + dz * dz;
              bool keep = d2 > 0.0 && d2 <= max2;
              if (FILTER && half_symmetric && ca == cb && A.p < p) keep = false;
              m |= keep ? (1u << (q + u)) : 0u;
            }
          }
          if (FILTER && drop_cell) m = 0;
          while (m) {
            const uint32_t bit = (uint32_t)__builtin_ctz(m);
            m &= m - 1u;
            out[lpos(min(cnt, last), t, T)] = (uint16_t)(base + p0 + bit);
            ++cnt;
          }
        }
      }
  }
  if (!A.live) cnt = 0;
  if (t < T) bcnt[(size_t)b * T + t] = cnt;
  const uint32_t wmax = wave_max_b(cnt);
  if ((t & 63) == 0 && wmax > 0) atomicMax(max_cnt, wmax);
  // pad to the wave's batch boundary with the sentinel, so force kernels need no per-lane
  // bound inside a batch
  const uint32_t pad_end = min((wmax + kListBatch - 1) / kListBatch * kListBatch, bp.cap);
  const uint16_t sentinel = (uint16_t)L.rpref[bp.nrc];
  for (uint32_t q = min(cnt, bp.cap); q < pad_end; q++) out[lpos(q, t, T)] = sentinel;
}

template <bool FILTER, bool PAIRWISE, bool PREF = false>
__global__ void __launch_bounds__(1024) k_nbh_brick_mask(Geom g, BrickParams bp, const double4* __restrict__ r,
                                                         const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                                         double max2, int half_symmetric, int skip_ghosts, uint16_t* __restrict__ blist,
                                                         uint32_t* __restrict__ bcnt, uint32_t* __restrict__ max_cnt, PairCut pc) {
  const uint32_t b = brick_of_block<true>(bp);
  if (b >= bp.nbricks) return;  // whole workgroup, before any barrier
  nbh_brick_body<FILTER, PAIRWISE, PREF>(g, bp, b, r, cell_off, cell_n, max2, half_symmetric, skip_ghosts, blist, bcnt, max_cnt, &pc);
}

// ------------------------------------------------------------------ pair traversal
// compute_pair_singlemat over brick lists (compute_cell_particle_pairs_impl_default.h:37-241)
// with a force functor F (include/xnb_compute_pair.h): every lane walks its central atom's
// list in stream order.  A Batched functor gets the entries U at a time and makes the cutoff
// test itself (padding entries are the far-away sentinel); otherwise the traversal accepts a
// pair iff 0 < d2 <= rc2 (d = r_b - r_a, d2 = (dx dx + dy dy) + dz dz) and calls the
// buffer-less form.  A ComputeBufferCompatible functor sums from 0 and the sum is stored
// (f = 0 + sum, or f += sum with ACCUM); a buffer-less-only functor adds into the stored
// force (from 0 without ACCUM).  The per-atom energy is written only for functors with the
// energy field.
//
// EPI selects the epilogue after the force store:
//   EPI_NONE       nothing more;
//   EPI_KICK       v += f dth, the closing push_f_v of the step (numerical-scheme.msp);
//   EPI_NEXT       the closing push_f_v, then the next step's push_f_v_r / push_f_v
//                  (x += v dt + f dt2; v += f dth) into NextStep::r_next, and the
//                  particle_displ_over count of the new positions against the backup,
//                  all from the force in registers; f and ep are neither computed nor
//                  stored (the next force kernel overwrites both before anything reads
//                  them; the last step of a run uses EPI_KICK and stores them).
// Positions are read from r (staged in LDS) and, with EPI_NEXT, written to a second buffer,
// so no workgroup reads a position another one has already advanced.
template <class F, bool ACCUM, int EPI, int U>
__device__ __forceinline__ void compute_pair_brick_body(const Geom& g, const BrickParams& bp, uint32_t b,
                                                        const double4* __restrict__ r, const uint32_t* __restrict__ cell_off,
                                                        const uint32_t* __restrict__ cell_n, const uint16_t* __restrict__ blist,
                                                        const uint32_t* __restrict__ bcnt, const F& func, double rc2, const PState& A,
                                                        double dth, const NextStep& nx) {
  using Tr = ComputePairTraits<F>;
  static_assert(Tr::Batched || Tr::BufferLessCompatible, "the traversal calls the batched or the buffer-less form");
  static_assert(!(EPI == EPI_NEXT && ACCUM), "a fused step starts the force from 0");
  constexpr bool BUFFERED = Tr::ComputeBufferCompatible;
  constexpr bool ENERGY = Tr::CentralEnergy && EPI != EPI_NEXT;
  extern __shared__ __align__(16) char lds[];
  const BrickLds L = brick_lds(bp, lds);
  const BrickPos P = brick_pos(g, bp, b);
  if (EPI == EPI_NEXT && threadIdx.x == 0) {  // ordered before every wave's count by brick_setup's barriers
    L.wcnt[0] = 0;
    L.wcnt[1] = 0;
  }
  const uint32_t t = threadIdx.x, T = bp.T;
  // U entries per batch = U / 4 eight-byte loads per lane; the next batch's loads are issued
  // before the current batch is computed.  Batches never pass cap (a multiple of U), and the
  // build padded every list of the wave with the sentinel up to a multiple of U.
  static_assert(U % 4 == 0 && U == kListBatch, "batches are whole groups of four entries, as padded");
  constexpr int W = U / 4;
  const uint2* lst = reinterpret_cast<const uint2*>(blist + (size_t)b * bp.cap * T) + t;
  // The list length and the first batch are issued before the staging so their latency
  // overlaps it (the build wrote 0 for every lane of the brick without an atom).  The first
  // batch lies inside the lane's slab (cap >= U) and is used only by waves with a list.
  const uint32_t n_list = t < T ? bcnt[(size_t)b * T + t] : 0u;
  uint2 wn[W];
#pragma unroll
  for (int w = 0; w < W; w++) wn[w] = t < T ? load_list_group(lst + (size_t)w * T) : make_uint2(0u, 0u);
  if (!brick_setup(g, bp, P, L, r, cell_off, cell_n, kPosF64)) return;
  const LaneAtom At = lane_atom(bp, L);
  const uint32_t n = At.live ? n_list : 0u;
  const uint32_t nmax = wave_max_b(n);
  double4 ra;
  ra.x = L.px[3 * At.ia];
  ra.y = L.py[3 * At.ia];
  ra.z = L.pz[3 * At.ia];
  ra.w = 0.;
  double tx = 0.0, ty = 0.0, tz = 0.0, te = 0.0;
  if (!BUFFERED && ACCUM && At.live) {  // buffer-less form: pairs go straight into the stored force
    tx = A.fx[At.slot];
    ty = A.fy[At.slot];
    tz = A.fz[At.slot];
    if (Tr::CentralEnergy) te = A.ep[At.slot];
  }
  const int type_a = Tr::CentralType ? (int)L.pt[At.ia] : 0;
  for (uint32_t q0 = 0; q0 < nmax; q0 += U) {
    uint2 wc[W];
#pragma unroll
    for (int w = 0; w < W; w++) wc[w] = wn[w];
    if (q0 + U < nmax) {
      const uint2* nxt = lst + (size_t)((q0 + U) >> 2) * T;
#pragma unroll
      for (int w = 0; w < W; w++) wn[w] = load_list_group(nxt + (size_t)w * T);
    }
    NbhBatch<U> nb;
    nb.pos = L.px;
    nb.type = L.pt;
#pragma unroll
    for (int w = 0; w < W; w++) {
      nb.j[4 * w + 0] = wc[w].x & 0xffffu;
      nb.j[4 * w + 1] = wc[w].x >> 16;
      nb.j[4 * w + 2] = wc[w].y & 0xffffu;
      nb.j[4 * w + 3] = wc[w].y >> 16;
      // 24-bit multiplies of the 16-bit halves (one instruction each with a word select)
      nb.off[4 * w + 0] = __umul24(wc[w].x & 0xffffu, 24u);
      nb.off[4 * w + 1] = __umul24(wc[w].x >> 16, 24u);
      nb.off[4 * w + 2] = __umul24(wc[w].y & 0xffffu, 24u);
      nb.off[4 * w + 3] = __umul24(wc[w].y >> 16, 24u);
    }
    // entries past a lane's count are the sentinel (rejected by the cutoff)
    if constexpr (Tr::Batched) {
      func.template batch<U, ENERGY>(ra, nb, tx, ty, tz, te);
    } else {
      const NbhCells cells{L.pt, nullptr};
#pragma unroll
      for (int u = 0; u < U; u++) {
        const double4 rb = nb(u);
        const Vec3d dr{rb.x - ra.x, rb.y - ra.y, rb.z - ra.z};
        const double d2 = (dr.x * dr.x + dr.y * dr.y) + dr.z * dr.z;
        if (d2 > 0.0 && d2 <= rc2) {
          if constexpr (Tr::CentralType && Tr::CentralEnergy)
            func(dr, d2, type_a, tx, ty, tz, te, cells, 0, nb.j[u], 1.0);
          else if constexpr (Tr::CentralType)
            func(dr, d2, type_a, tx, ty, tz, cells, 0, nb.j[u], 1.0);
          else if constexpr (Tr::CentralEnergy)
            func(dr, d2, tx, ty, tz, te, cells, 0, nb.j[u], 1.0);
          else
            func(dr, d2, tx, ty, tz, cells, 0, nb.j[u], 1.0);
        }
      }
    }
  }
  constexpr bool NEXT = EPI == EPI_NEXT;
  bool over = false;
  if (At.live) {
    const uint32_t s = At.slot;
    if (EPI != EPI_NEXT) {
      if (BUFFERED) {  // f = 0 + sum (a zeroed force) or f += sum
        A.fx[s] = (ACCUM ? A.fx[s] : 0.0) + tx;
        A.fy[s] = (ACCUM ? A.fy[s] : 0.0) + ty;
        A.fz[s] = (ACCUM ? A.fz[s] : 0.0) + tz;
        if (Tr::CentralEnergy) A.ep[s] = (ACCUM ? A.ep[s] : 0.0) + te;
      } else {
        A.fx[s] = tx;
        A.fy[s] = ty;
        A.fz[s] = tz;
        if (Tr::CentralEnergy) A.ep[s] = te;
      }
    }
    if (EPI != EPI_NONE) {
      // the force as stored (0 + sum with ACCUM false, which EPI_NEXT requires)
      const double ax = EPI == EPI_NEXT ? 0.0 + tx : A.fx[s];
      const double ay = EPI == EPI_NEXT ? 0.0 + ty : A.fy[s];
      const double az = EPI == EPI_NEXT ? 0.0 + tz : A.fz[s];
      double ux = A.vx[s] + ax * dth;
      double uy = A.vy[s] + ay * dth;
      double uz = A.vz[s] + az * dth;
      if (NEXT) {
        double4 p = ra;  // the LDS copy of r[s].xyz
        p.w = reinterpret_cast<const double*>(r)[4 * (size_t)s + 3];
        p.x += ux * nx.dt + ax * nx.dt2;
        p.y += uy * nx.dt + ay * nx.dt2;
        p.z += uz * nx.dt + az * nx.dt2;
        // positions and velocities of the next step are streamed out (read again only
        // after the whole launch, by which time 8.4M atoms' worth has passed through L2)
        __builtin_nontemporal_store(p.x, &nx.r_next[s].x);
        __builtin_nontemporal_store(p.y, &nx.r_next[s].y);
        __builtin_nontemporal_store(p.z, &nx.r_next[s].z);
        __builtin_nontemporal_store(p.w, &nx.r_next[s].w);
        ux += ax * dth;
        uy += ay * dth;
        uz += az * dth;
        const int ri = (int)(At.rc % (uint32_t)bp.RX), rj = (int)((At.rc / (uint32_t)bp.RX) % (uint32_t)bp.RY),
                  rk = (int)(At.rc / ((uint32_t)bp.RX * (uint32_t)bp.RY));
        over = displaced_ijk(g, p, P.gi0 - 1 + ri, P.gj0 - 1 + rj, P.gk0 - 1 + rk, nx.backup, s, nx.thr2);
      }
      __builtin_nontemporal_store(ux, &A.vx[s]);
      __builtin_nontemporal_store(uy, &A.vy[s]);
      __builtin_nontemporal_store(uz, &A.vz[s]);
    }
  }
  if (NEXT) wave_group_count_add(over, nx.count, L.wcnt);
}

template <bool MIXED, bool ACCUM, int EPI, int U>
// waves_per_eu(5) caps registers at 96: measured faster for the FP64 instantiations, neutral
// for the mixed ones.  The FP64 form's spills (44-52 bytes) all sit in the pair loop's
// small-argument branch (d2 < 2^-767, the generic root and division), not in its hot path.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) k_lj_brick(Geom g, BrickParams bp, const double4* __restrict__ r,
                                                   const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                                   const uint16_t* __restrict__ blist, const uint32_t* __restrict__ bcnt,
                                                   LJConst lj, PState A, double dth, NextStep nx) {
  const uint32_t b = brick_of_block<MIXED>(bp);
  if (b >= bp.nbricks) return;  // whole workgroup, before any barrier
  compute_pair_brick_body<LennardJonesForceFunctor<MIXED>, ACCUM, EPI, U>(g, bp, b, r, cell_off, cell_n, blist, bcnt,
                                                                          LennardJonesForceFunctor<MIXED>{lj}, lj.rc2, A, dth, nx);
}

// Any other force functor on brick lists (no fused epilogue): f += pairs (ACCUM, after
// zero_particle_force).
template <class F>
__global__ void __launch_bounds__(1024) k_compute_pair_brick(Geom g, BrickParams bp, const double4* __restrict__ r,
                                                             const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                                             const uint16_t* __restrict__ blist, const uint32_t* __restrict__ bcnt, F func,
                                                             double rc2, PState A) {
  const uint32_t b = brick_of_block<true>(bp);
  if (b >= bp.nbricks) return;  // whole workgroup, before any barrier
  compute_pair_brick_body<F, true, EPI_NONE, LJB_UNROLL_B>(g, bp, b, r, cell_off, cell_n, blist, bcnt, func, rc2, A, 0.0, NextStep{});
}

// ------------------------------------------------------------------ multimat force
// compute_pair_multimat (compute_pair_multimat.h:100-253) on brick lists: a pair (a, b) is
// accepted iff d2 <= rc(pid)^2, pid = unique_pair_id(type a, type b); the per-atom sums are
// grouped by pair id, each group summed in stream order from 0 and added to f in increasing
// pair id, a group without accepted pairs not added.  For a fixed type a, pid increases with
// type b, so group t is the group of neighbours of type t.  Every lane updates all NT groups
// per pair, the non-matching ones by +0.0 (sums start at +0.0 and stay out of -0.0, so this
// is bitwise neutral).  The mixed form keeps one sum in stream order instead (it has a
// tolerance, not the reference's rounding).  ACCUM adds to the stored f (after
// zero_particle_force); otherwise f starts from 0.0 like a zeroed force.  EPI as in k_lj_brick.
template <bool MIXED, int NT, bool ACCUM, int EPI, int U>
__device__ __forceinline__ void ljmm_brick_body(const Geom& g, const BrickParams& bp, uint32_t b, const double4* __restrict__ r,
                                                const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                                const uint16_t* __restrict__ blist, const uint32_t* __restrict__ bcnt,
                                                const LJMulti& mm, const PState& A, double dth, const NextStep& nx) {
  extern __shared__ __align__(16) char lds[];
  const BrickLds L = brick_lds(bp, lds);
  const BrickPos P = brick_pos(g, bp, b);
  if (EPI == EPI_NEXT && threadIdx.x == 0) {  // ordered before every wave's count by brick_setup's barriers
    L.wcnt[0] = 0;
    L.wcnt[1] = 0;
  }
  if (!brick_setup(g, bp, P, L, r, cell_off, cell_n, kPosF64)) return;
  const LaneAtom At = lane_atom(bp, L);
  const uint32_t t = threadIdx.x, T = bp.T;
  const uint32_t n = At.live ? bcnt[(size_t)b * T + t] : 0u;
  const uint32_t nmax = wave_max_b(n);
  const double ax0 = L.px[3 * At.ia], ay0 = L.py[3 * At.ia], az0 = L.pz[3 * At.ia];
  const unsigned ta = L.pt[At.ia];
  // this lane's parameters per neighbour type
  double sg[NT], e4[NT], m24[NT], rc2[NT];
  float s6f[NT], e4f[NT], m24f[NT];
#pragma unroll
  for (int q = 0; q < NT; q++) {
    const unsigned pid = unique_pair_id(ta, (unsigned)q);
    sg[q] = mm.sigma[pid];
    e4[q] = mm.eps4[pid];
    m24[q] = mm.m24eps[pid];
    rc2[q] = mm.rc2[pid];
    s6f[q] = mm.sigma6f[pid];
    e4f[q] = mm.eps4f[pid];
    m24f[q] = mm.m24epsf[pid];
  }
  constexpr bool ENERGY = EPI != EPI_NEXT;
  constexpr int H = MIXED ? U : 2;
  double gx[NT], gy[NT], gz[NT], ge[NT];
  bool any[NT];
#pragma unroll
  for (int q = 0; q < NT; q++) {
    gx[q] = gy[q] = gz[q] = ge[q] = 0.0;
    any[q] = false;
  }
  const uint2* lst = reinterpret_cast<const uint2*>(blist + (size_t)b * bp.cap * T) + t;
  constexpr int W = U / 4;
  uint2 wn[W];
  if (nmax > 0) {
#pragma unroll
    for (int w = 0; w < W; w++) wn[w] = load_list_group(lst + (size_t)w * T);
  }
  for (uint32_t q0 = 0; q0 < nmax; q0 += U) {
    uint2 wc[W];
#pragma unroll
    for (int w = 0; w < W; w++) wc[w] = wn[w];
    if (q0 + U < nmax) {
      const uint2* nxt = lst + (size_t)((q0 + U) >> 2) * T;
#pragma unroll
      for (int w = 0; w < W; w++) wn[w] = load_list_group(nxt + (size_t)w * T);
    }
    uint32_t j[U];
#pragma unroll
    for (int w = 0; w < W; w++) {
      j[4 * w + 0] = wc[w].x & 0xffffu;
      j[4 * w + 1] = wc[w].x >> 16;
      j[4 * w + 2] = wc[w].y & 0xffffu;
      j[4 * w + 3] = wc[w].y >> 16;
    }
    // pairs in sub-batches of H: one range decision per sub-batch for the FP64 roots.  FP64
    // takes pairs by two: its per-type tables and sums leave too few registers for more in
    // flight (timed on the two-species slab: force 3.39 ms at H = 2, 3.44 at 4, 3.54 at 8)
#pragma unroll
    for (int h = 0; h < U; h += H) {
      // per pair: displacement, neighbour type, inside its pair's cutoff
      double dx[H], dy[H], dz[H], d2[H];
      unsigned tb[H];
      bool in[H];
      bool small = false;
#pragma unroll
      for (int u = 0; u < H; u++) {
        dx[u] = L.px[3 * j[h + u]] - ax0;
        dy[u] = L.py[3 * j[h + u]] - ay0;
        dz[u] = L.pz[3 * j[h + u]] - az0;
        tb[u] = min((unsigned)L.pt[j[h + u]], (unsigned)(NT - 1));
        double prc2 = rc2[0];
#pragma unroll
        for (int q = 1; q < NT; q++) prc2 = tb[u] == (unsigned)q ? rc2[q] : prc2;
        d2[u] = MIXED ? fma(dz[u], dz[u], fma(dy[u], dy[u], dx[u] * dx[u])) : (dx[u] * dx[u] + dy[u] * dy[u]) + dz[u] * dz[u];
        in[u] = d2[u] <= prc2;
        small = small || (in[u] && d2[u] < 0x1p-767);
      }
      double inv_r[H];  // plain form: 1.0 / sqrt(d2), each correctly rounded (1.0 outside the cutoff)
      if (!MIXED) {
        if (__builtin_expect(__any(small), 0)) {
#pragma unroll
          for (int u = 0; u < H; u++) inv_r[u] = 1.0 / sqrt(in[u] ? d2[u] : 1.0);
        } else {
#pragma unroll
          for (int u = 0; u < H; u++) inv_r[u] = recip_normal(sqrt_normal(in[u] ? d2[u] : 1.0));
        }
      }
#pragma unroll
      for (int u = 0; u < H; u++) {
        double psg = sg[0], pe4 = e4[0], pm24 = m24[0];
        float ps6f = s6f[0], pe4f = e4f[0], pm24f = m24f[0];
#pragma unroll
        for (int q = 1; q < NT; q++)
          if (tb[u] == (unsigned)q) {
            psg = sg[q];
            pe4 = e4[q];
            pm24 = m24[q];
            ps6f = s6f[q];
            pe4f = e4f[q];
            pm24f = m24f[q];
          }
        double coef, ce;  // force coefficient (f += coef * d) and pair energy
        if (MIXED) {
          const float inv_r2 = in[u] ? __builtin_amdgcn_rcpf((float)d2[u]) : 0.0f;  // 0 outside: coefficient and energy 0
          float s6;
          coef = (double)lj_coef_f32(inv_r2, ps6f, pm24f, s6);
          ce = ENERGY ? (double)lj_energy_f32(s6, pe4f) : 0.0;
        } else {
          const double ratio = psg * inv_r[u];
          const double ratio2 = ratio * ratio;
          const double ratio6 = ratio2 * ratio2 * ratio2;
          const double ratio12 = ratio6 * ratio6;
          double de = (pm24 * (2. * ratio12 - ratio6)) * inv_r[u];
          de *= inv_r[u];
          coef = in[u] ? de : 0.0;
          ce = (ENERGY && in[u]) ? pe4 * (ratio12 - ratio6) : 0.0;
        }
        if (MIXED) {
          // one sum in stream order, f = fma(c, d, f) as in the singlemat mixed form (c and the
          // energy are +-0 outside the cutoff): the tolerance form needs no per-type grouping
          gx[0] = fma(coef, dx[u], gx[0]);
          gy[0] = fma(coef, dy[u], gy[0]);
          gz[0] = fma(coef, dz[u], gz[0]);
          if (ENERGY) ge[0] += ce;
        } else {
#pragma unroll
          for (int q = 0; q < NT; q++) {  // f += (de / r) * d per neighbour-type group, the plain form
            const bool mine = in[u] && tb[u] == (unsigned)q;
            const double cq = mine ? coef : 0.0;
            gx[q] += cq * dx[u];
            gy[q] += cq * dy[u];
            gz[q] += cq * dz[u];
            if (ENERGY) ge[q] += mine ? ce : 0.0;
            any[q] = any[q] || mine;
          }
        }
      }
    }
  }
  constexpr bool NEXT = EPI == EPI_NEXT;
  bool over = false;
  if (At.live) {
    const uint32_t s = At.slot;
    double fx = ACCUM ? A.fx[s] : 0.0, fy = ACCUM ? A.fy[s] : 0.0, fz = ACCUM ? A.fz[s] : 0.0, fe = ACCUM ? A.ep[s] : 0.0;
#pragma unroll
    for (int q = 0; q < NT; q++)
      if (MIXED ? q == 0 : any[q]) {
        fx += gx[q];
        fy += gy[q];
        fz += gz[q];
        fe += ge[q];
      }
    if (EPI != EPI_NEXT) {
      A.fx[s] = fx;
      A.fy[s] = fy;
      A.fz[s] = fz;
      A.ep[s] = fe;
    }
    if (EPI != EPI_NONE) {
      double ux = A.vx[s] + fx * dth;
      double uy = A.vy[s] + fy * dth;
      double uz = A.vz[s] + fz * dth;
      if (NEXT) {
        double4 p;
        p.x = ax0;
        p.y = ay0;
        p.z = az0;
        p.w = reinterpret_cast<const double*>(r)[4 * (size_t)s + 3];
        p.x += ux * nx.dt + fx * nx.dt2;
        p.y += uy * nx.dt + fy * nx.dt2;
        p.z += uz * nx.dt + fz * nx.dt2;
        __builtin_nontemporal_store(p.x, &nx.r_next[s].x);
        __builtin_nontemporal_store(p.y, &nx.r_next[s].y);
        __builtin_nontemporal_store(p.z, &nx.r_next[s].z);
        __builtin_nontemporal_store(p.w, &nx.r_next[s].w);
        ux += fx * dth;
        uy += fy * dth;
        uz += fz * dth;
        const int ri = (int)(At.rc % (uint32_t)bp.RX), rj = (int)((At.rc / (uint32_t)bp.RX) % (uint32_t)bp.RY),
                  rk = (int)(At.rc / ((uint32_t)bp.RX * (uint32_t)bp.RY));
        over = displaced_ijk(g, p, P.gi0 - 1 + ri, P.gj0 - 1 + rj, P.gk0 - 1 + rk, nx.backup, s, nx.thr2);
      }
      __builtin_nontemporal_store(ux, &A.vx[s]);
      __builtin_nontemporal_store(uy, &A.vy[s]);
      __builtin_nontemporal_store(uz, &A.vz[s]);
    }
  }
  if (NEXT) wave_group_count_add(over, nx.count, L.wcnt);
}

template <bool MIXED, int NT, bool ACCUM, int EPI>
__global__ void __launch_bounds__(1024) k_ljmm_brick(Geom g, BrickParams bp, const double4* __restrict__ r,
                                                     const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                                                     const uint16_t* __restrict__ blist, const uint32_t* __restrict__ bcnt, LJMulti mm,
                                                     PState A, double dth, NextStep nx) {
  const uint32_t b = brick_of_block<MIXED>(bp);
  if (b >= bp.nbricks) return;  // whole workgroup, before any barrier
  ljmm_brick_body<MIXED, NT, ACCUM, EPI, LJB_UNROLL_B>(g, bp, b, r, cell_off, cell_n, blist, bcnt, mm, A, dth, nx);
}

// ------------------------------------------------------------------ conversion
// brick list -> generic per-slot u32 list (slot indices), for stream export and the generic
// traversals.
__global__ void __launch_bounds__(1024) k_brick_to_slots(Geom g, BrickParams bp, const uint32_t* __restrict__ cell_off,
                                                         const uint32_t* __restrict__ cell_n, const uint16_t* __restrict__ blist,
                                                         const uint32_t* __restrict__ bcnt, uint32_t* __restrict__ nbr,
                                                         uint32_t* __restrict__ nbr_cnt, uint32_t cap) {
  extern __shared__ __align__(16) char lds[];
  const uint32_t b = brick_of_block<true>(bp);
  if (b >= bp.nbricks) return;  // whole workgroup, before any barrier
  const BrickLds L = brick_lds(bp, lds);
  const BrickPos P = brick_pos(g, bp, b);
  if (!brick_setup(g, bp, P, L, nullptr, cell_off, cell_n, kPosNone)) return;
  const LaneAtom A = lane_atom(bp, L);
  if (!A.live) return;
  const uint32_t t = threadIdx.x, T = bp.T;
  const uint32_t n = min(bcnt[(size_t)b * T + t], bp.cap);
  const uint16_t* lst = blist + (size_t)b * bp.cap * T;
  uint32_t* outp = nbr + nbr_index(A.slot, 0, cap);
  for (uint32_t q = 0; q < n; q++) {
    const uint32_t idx = lst[lpos(q, t, T)];
    const uint32_t rc = upper_index(L.rpref, (uint32_t)bp.nrc, idx);
    outp[(size_t)q << 6] = L.roff[rc] + (idx - L.rpref[rc]);
  }
  nbr_cnt[A.slot] = n;
}

// ------------------------------------------------------------------ launchers
BrickParams brick_params(const Geom& g, int Bi, int Bj, int Bk) {
  BrickParams bp{};
  bp.Bi = Bi;
  bp.Bj = Bj;
  bp.Bk = Bk;
  bp.nbx = (g.Di + Bi - 1) / Bi;
  bp.nby = (g.Dj + Bj - 1) / Bj;
  bp.nbz = (g.Dk + Bk - 1) / Bk;
  bp.nbricks = (uint32_t)bp.nbx * bp.nby * bp.nbz;
  bp.per_xcd = (bp.nbricks + kXcds - 1) / kXcds;
  bp.RX = Bi + 2;
  bp.RY = Bj + 2;
  bp.RZ = Bk + 2;
  bp.nrc = bp.RX * bp.RY * bp.RZ;
  bp.nbc = Bi * Bj * Bk;
  return bp;
}

static inline uint32_t align8(uint32_t x) { return (x + 7u) & ~7u; }

// LDS layout of the build, force and conversion kernels; returns its bytes
uint32_t brick_layout(BrickParams& bp, bool f32_build, bool f32_types) {
  // positions first, at the start of dynamic LDS, so their addresses need no runtime base
  uint32_t o = 0;
  bp.off_pos = o;
  // region atoms + the sentinel block: FP64 x, y, z and a type byte, or the prefiltered
  // build's FP32 float4 (a type byte only for the per-pair-radius build)
  o += (f32_build ? 16u : 24u) * (bp.Rmax + kSentinels);
  bp.off_type = o;
  if (!f32_build || f32_types) o = align8(o + bp.Rmax + kSentinels);
  bp.off_rcnt = o;
  o = align8(o + 4u * bp.nrc);
  bp.off_rpref = o;
  o = align8(o + 4u * (bp.nrc + 1));
  bp.off_roff = o;
  o = align8(o + 4u * bp.nrc);
  bp.off_bpref = o;
  o = align8(o + 4u * (bp.nbc + 1));
  bp.off_rdel = o;
  o = align8(o + 4u * (uint32_t)(bp.nrc / bp.RX));
  bp.off_crow = o;
  o = align8(o + 2u * ((bp.Rmax + kSentinels + 63u) / 64u + 1u));
  bp.off_wsum = o;
  o += 4u * (32u + 2u);
  return align8(o);
}

void launch_brick_stats(const Geom& g, const BrickParams& bp, const uint32_t* cell_n, uint32_t* out, hipStream_t st) {
  k_brick_stats<<<(bp.nbricks + 255) / 256, 256, 0, st>>>(g, bp, cell_n, out);
}

// Brick kernels keep all their LDS in the dynamic block, which then starts at LDS address 0:
// the force traversal addresses staged atoms by absolute LDS address (NbhBatch::off).  Checked
// once per kernel before its first launch.
static hipError_t check_no_static_lds(const void* fn) {
  static std::mutex m;
  static std::vector<const void*> ok;
  std::lock_guard<std::mutex> g(m);
  if (std::find(ok.begin(), ok.end(), fn) != ok.end()) return hipSuccess;
  hipFuncAttributes a{};
  const hipError_t e = hipFuncGetAttributes(&a, fn);
  if (e != hipSuccess) return e;
  if (a.sharedSizeBytes != 0) return hipErrorInvalidConfiguration;
  ok.push_back(fn);
  return hipSuccess;
}

static hipError_t set_lds(const void* fn, uint32_t bytes) {
  const hipError_t e = check_no_static_lds(fn);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// pc.nt > 0: per type-pair list radii (not with the half_symmetric / skip_ghosts filters)
hipError_t launch_nbh_brick(const Geom& g, BrickParams bp, const double4* r, const uint32_t* cell_off, const uint32_t* cell_n,
                            double max2, const PairCut& pc, int half_symmetric, int skip_ghosts, bool prefilter, uint16_t* blist,
                            uint32_t* bcnt, uint32_t* max_cnt, hipStream_t st) {
  const uint32_t lds = brick_layout(bp);
  hipError_t e;
  if (half_symmetric || skip_ghosts) {
    if (pc.nt > 0) return hipErrorInvalidValue;
    e = set_lds((const void*)k_nbh_brick_mask<true, false>, lds);
    if (e != hipSuccess) return e;
    k_nbh_brick_mask<true, false><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, r, cell_off, cell_n, max2, half_symmetric, skip_ghosts,
                                                                        blist, bcnt, max_cnt, pc);
  } else if (pc.nt > 0 && prefilter) {
    BrickParams bq = bp;
    const uint32_t lds_b = brick_layout(bq, true, true);
    e = set_lds((const void*)k_nbh_brick_mask<false, true, true>, lds_b);
    if (e != hipSuccess) return e;
    k_nbh_brick_mask<false, true, true><<<kXcds * bq.per_xcd, bq.T, lds_b, st>>>(g, bq, r, cell_off, cell_n, max2, 0, 0, blist, bcnt,
                                                                                max_cnt, pc);
  } else if (pc.nt > 0) {
    e = set_lds((const void*)k_nbh_brick_mask<false, true>, lds);
    if (e != hipSuccess) return e;
    k_nbh_brick_mask<false, true><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, r, cell_off, cell_n, max2, 0, 0, blist, bcnt, max_cnt,
                                                                        pc);
  } else if (prefilter) {
    BrickParams bq = bp;  // the same bricks and lists, the prefilter's smaller LDS layout
    const uint32_t lds_b = brick_layout(bq, true);
    e = set_lds((const void*)k_nbh_brick_mask<false, false, true>, lds_b);
    if (e != hipSuccess) return e;
    k_nbh_brick_mask<false, false, true><<<kXcds * bq.per_xcd, bq.T, lds_b, st>>>(g, bq, r, cell_off, cell_n, max2, 0, 0, blist, bcnt,
                                                                                 max_cnt, pc);
  } else {
    e = set_lds((const void*)k_nbh_brick_mask<false, false>, lds);
    if (e != hipSuccess) return e;
    k_nbh_brick_mask<false, false><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, r, cell_off, cell_n, max2, 0, 0, blist, bcnt, max_cnt,
                                                                         pc);
  }
  return hipGetLastError();
}

static constexpr int LJB_UNROLL = LJB_UNROLL_B;

template <bool MIXED, bool ACCUM, int EPI>
static hipError_t ljb_launch(const Geom& g, const BrickParams& bp, uint32_t lds, const double4* r, const uint32_t* cell_off,
                             const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, const LJConst& lj, const PState& A,
                             double dth, const NextStep& nx, hipStream_t st) {
  hipError_t e = set_lds((const void*)k_lj_brick<MIXED, ACCUM, EPI, LJB_UNROLL>, lds);
  if (e != hipSuccess) return e;
  k_lj_brick<MIXED, ACCUM, EPI, LJB_UNROLL><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, r, cell_off, cell_n, blist, bcnt, lj, A, dth, nx);
  return hipGetLastError();
}

uint32_t brick_list_granule() { return LJB_UNROLL; }

hipError_t launch_lj_brick(int precision, bool accum, int epi, const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off,
                           const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, double eps, double sigma, double rcut,
                           double dth, const NextStep* next, hipStream_t st) {
  const LJConst lj = make_lj_const(eps, sigma, rcut);
  const uint32_t lds = brick_layout(bp);
  const NextStep nx = next ? *next : NextStep{};
  const bool next_epi = epi == EPI_NEXT;
  if (next_epi && (accum || !next)) return hipErrorInvalidValue;
#define XNB_LJB_ARGS g, bp, lds, A.r, cell_off, cell_n, blist, bcnt, lj, A, dth, nx, st
  if (precision) {
    if (epi == EPI_NEXT) return ljb_launch<true, false, EPI_NEXT>(XNB_LJB_ARGS);
    if (accum) return epi ? ljb_launch<true, true, EPI_KICK>(XNB_LJB_ARGS) : ljb_launch<true, true, EPI_NONE>(XNB_LJB_ARGS);
    return epi ? ljb_launch<true, false, EPI_KICK>(XNB_LJB_ARGS) : ljb_launch<true, false, EPI_NONE>(XNB_LJB_ARGS);
  }
  if (epi == EPI_NEXT) return ljb_launch<false, false, EPI_NEXT>(XNB_LJB_ARGS);
  if (accum) return epi ? ljb_launch<false, true, EPI_KICK>(XNB_LJB_ARGS) : ljb_launch<false, true, EPI_NONE>(XNB_LJB_ARGS);
  return epi ? ljb_launch<false, false, EPI_KICK>(XNB_LJB_ARGS) : ljb_launch<false, false, EPI_NONE>(XNB_LJB_ARGS);
#undef XNB_LJB_ARGS
}

template <bool MIXED, int NT, bool ACCUM, int EPI>
static hipError_t ljmm_launch(const Geom& g, const BrickParams& bp, uint32_t lds, const uint32_t* cell_off, const uint32_t* cell_n,
                              const uint16_t* blist, const uint32_t* bcnt, const LJMulti& mm, const PState& A, double dth,
                              const NextStep& nx, hipStream_t st) {
  hipError_t e = set_lds((const void*)k_ljmm_brick<MIXED, NT, ACCUM, EPI>, lds);
  if (e != hipSuccess) return e;
  k_ljmm_brick<MIXED, NT, ACCUM, EPI><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, A.r, cell_off, cell_n, blist, bcnt, mm, A, dth, nx);
  return hipGetLastError();
}

template <bool MIXED, int NT>
static hipError_t ljmm_dispatch(int epi, const Geom& g, const BrickParams& bp, uint32_t lds, const uint32_t* cell_off,
                                const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, const LJMulti& mm, const PState& A,
                                double dth, const NextStep& nx, hipStream_t st) {
  switch (epi) {
    case EPI_NONE: return ljmm_launch<MIXED, NT, true, EPI_NONE>(g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
    case EPI_KICK: return ljmm_launch<MIXED, NT, false, EPI_KICK>(g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
    default: return ljmm_launch<MIXED, NT, false, EPI_NEXT>(g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
  }
}

// Multimat force on brick lists.  epi EPI_NONE accumulates into f (after zero_particle_force);
// the other epilogues start f from 0, as the run loop's zero + force.  ntypes 1..4.
hipError_t launch_ljmm_brick(int precision, int ntypes, int epi, const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off,
                             const uint32_t* cell_n, const uint16_t* blist, const uint32_t* bcnt, const LJMulti& mm, double dth,
                             const NextStep* next, hipStream_t st) {
  if (ntypes < 1 || ntypes > 4) return hipErrorInvalidValue;
  if (epi == EPI_NEXT && !next) return hipErrorInvalidValue;
  const uint32_t lds = brick_layout(bp);
  const NextStep nx = next ? *next : NextStep{};
  if (precision) {
    if (ntypes <= 2) return ljmm_dispatch<true, 2>(epi, g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
    return ljmm_dispatch<true, 4>(epi, g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
  }
  if (ntypes <= 2) return ljmm_dispatch<false, 2>(epi, g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
  return ljmm_dispatch<false, 4>(epi, g, bp, lds, cell_off, cell_n, blist, bcnt, mm, A, dth, nx, st);
}

hipError_t launch_gravity_brick(const Geom& g, BrickParams bp, const PState& A, const uint32_t* cell_off, const uint32_t* cell_n,
                                const uint16_t* blist, const uint32_t* bcnt, const GravitationalForceFunctor& gf, double rc2,
                                hipStream_t st) {
  const uint32_t lds = brick_layout(bp);
  hipError_t e = set_lds((const void*)k_compute_pair_brick<GravitationalForceFunctor>, lds);
  if (e != hipSuccess) return e;
  k_compute_pair_brick<GravitationalForceFunctor><<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, A.r, cell_off, cell_n, blist, bcnt, gf,
                                                                                         rc2, A);
  return hipGetLastError();
}

hipError_t launch_brick_to_slots(const Geom& g, BrickParams bp, const uint32_t* cell_off, const uint32_t* cell_n, const uint16_t* blist,
                                 const uint32_t* bcnt, uint32_t* nbr, uint32_t* nbr_cnt, uint32_t cap, hipStream_t st) {
  const uint32_t lds = brick_layout(bp);
  hipError_t e = set_lds((const void*)k_brick_to_slots, lds);
  if (e != hipSuccess) return e;
  k_brick_to_slots<<<kXcds * bp.per_xcd, bp.T, lds, st>>>(g, bp, cell_off, cell_n, blist, bcnt, nbr, nbr_cnt, cap);
  return hipGetLastError();
}

// ------------------------------------------------------------------ batched boxes
// work[blockIdx.x] = box << 20 | brick of the box (NONE: idle).  The workgroup has the batch's
// widest T; a box with a smaller T leaves the extra waves without atoms.

__device__ inline bool batch_brick(const uint32_t* __restrict__ work, uint32_t& box, uint32_t& b) {
  const uint32_t w = work[blockIdx.x];
  if (w == NONE) return false;
  box = w >> kWorkBoxShift;
  b = w & kWorkBlockMask;
  return true;
}

__global__ void __launch_bounds__(1024) kb_nbh_brick(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work, uint32_t par) {
  uint32_t box, b;
  if (!batch_brick(work, box, b)) return;  // whole workgroup
  const BoxFrame& f = F[box];
  // ghost slots overflowed in this rebuild (bit 2, set by the ghost fill): the cell ranges
  // reach past the slot arrays, so the box is left to the one-box recovery
  if (f.words[0] & 4u) return;
  const double4* r = f.rbuf[(f.rcur ^ par) & 1u];
  nbh_brick_body<false, false, true>(f.g, f.bpn, b, r, f.cell_off, f.cell_n, f.max2, 0, 0, f.blist, f.bcnt, f.words + 1);
}

template <bool MIXED, int EPI>
__device__ __forceinline__ void lj_batch_case(const BoxFrame& f, uint32_t b, uint32_t par, uint32_t gstep) {
  PState A = f.ab ? f.B : f.A;
  const uint32_t cur = (f.rcur ^ par) & 1u;
  A.r = f.rbuf[cur];
  // the next step's count: local step k = gstep - gstep0 + 1, count of step k + 1 at [k]
  const NextStep nx{f.rbuf[cur ^ 1u], f.backup, f.dt, f.dt2, f.thr2, f.counts + (gstep - f.gstep0 + 1)};
  compute_pair_brick_body<LennardJonesForceFunctor<MIXED>, false, EPI, LJB_UNROLL_B>(
      f.g, f.bp, b, A.r, f.cell_off, f.cell_n, f.blist, f.bcnt, LennardJonesForceFunctor<MIXED>{f.lj}, f.lj.rc2, A, f.dth, nx);
}

// Force of every brick of the listed boxes, all of one precision (one kernel per precision,
// so the mixed form runs with its own, smaller register budget); the epilogue is per box
// (uniform per workgroup).  EPI_KICK closes a box's run; EPI_NEXT fuses its next step;
// EPI_NONE stores the force and energy only.  The
// register cap of 5 waves per SIMD (the FP64 form otherwise takes 110 VGPRs, 4 waves) measured
// 3% faster on 256 FP64 boxes; the mixed form fits under it anyway.
template <bool MIXED>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5))) kb_lj_brick(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                    const uint32_t* __restrict__ clock) {
  uint32_t box, b;
  if (!batch_brick(work, box, b)) return;  // whole workgroup
  const uint32_t par = clock[0], gstep = clock[1];
  const uint32_t epi = work_entry_epi(b);
  b = work_entry_brick(b);
  const BoxFrame& f = F[box];
  // A rebuild in this step that failed or outgrew the box's frame (status bits, a list past
  // its capacity, an atom outside a non-periodic axis) leaves the box untouched: the host
  // reads the words after the step and redoes it with the one-box operators (or fails it).
  // The words are cleared at attach, before each batched rebuild and after a recovery.
  if (f.words[0] != 0u || f.words[1] > f.bp.cap || f.words[2] != 0u) return;
  if (epi == EPI_NEXT) lj_batch_case<MIXED, EPI_NEXT>(f, b, par, gstep);
  else if (epi == EPI_KICK) lj_batch_case<MIXED, EPI_KICK>(f, b, par, gstep);
  else lj_batch_case<MIXED, EPI_NONE>(f, b, par, gstep);  // force and energy only (the batch autotuner's timings)
}

static const void* batch_kernel(int kind) {
  switch (kind) {
    case kBatchLjFp64: return (const void*)kb_lj_brick<false>;
    case kBatchLjMixed: return (const void*)kb_lj_brick<true>;
    default: return (const void*)kb_nbh_brick;
  }
}

hipError_t launch_batch_nbh(const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t T, uint32_t lds, uint32_t par,
                            hipStream_t st, bool set) {
  if (!nwork) return hipSuccess;
  if (set) {
    hipError_t e = set_lds((const void*)kb_nbh_brick, lds);
    if (e != hipSuccess) return e;
  }
  kb_nbh_brick<<<nwork, T, lds, st>>>(F, work, par);
  return hipGetLastError();
}

hipError_t batch_set_lds(int kind, uint32_t lds) { return set_lds(batch_kernel(kind), lds); }

// Static LDS (wave counters) and the waves per CU its registers allow, for counting how
// many workgroups of a batched brick kernel fit one CU.
BatchKernelInfo batch_kernel_info(int kind) {
  BatchKernelInfo info{256u, 16u};
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, batch_kernel(kind)) == hipSuccess) {
    info.static_lds = (uint32_t)a.sharedSizeBytes;
    if (a.numRegs > 0) {
      const uint32_t vg = ((uint32_t)a.numRegs + 7u) & ~7u;  // allocation granule of 8 VGPRs
      info.waves_per_cu = 4u * std::min(8u, std::max(1u, 512u / vg));
    }
  }
  return info;
}

hipError_t launch_batch_lj(bool mixed, const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t T, uint32_t lds,
                           const uint32_t* clock, hipStream_t st, bool set) {
  if (!nwork) return hipSuccess;
  if (set) {
    hipError_t e = batch_set_lds(mixed ? kBatchLjMixed : kBatchLjFp64, lds);
    if (e != hipSuccess) return e;
  }
  if (mixed) kb_lj_brick<true><<<nwork, T, lds, st>>>(F, work, clock);
  else kb_lj_brick<false><<<nwork, T, lds, st>>>(F, work, clock);
  return hipGetLastError();
}

}  // namespace xnb
