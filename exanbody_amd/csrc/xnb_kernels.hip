// xnb_kernels.hip -- gfx950 kernels of the short-range LJ hot path.
//
// Storage (per box, in HBM): particles sorted by cell, inner cells first in cell-index
// order in slots [0, N), ghost cells after in [N, N+G).  cell_off/cell_n give each cell's
// slot range.  Positions are double4 (x, y, z, type) so one 32-byte load fetches a
// neighbour; velocities, forces and energies are SoA doubles.
//
// Compiled with -ffp-contract=off: FP64 arithmetic is evaluated in source order without
// contraction.  Explicit fma() is used only in the mixed-precision force kernel.
#include "xnb_kernels.h"

namespace xnb {

// ----------------------------------------------------------------------------- scans
// Exclusive scan of n u32 values: per-block sums, one-block scan of block sums, then
// per-block scan with offsets.  SCAN_ITEMS items per thread, 256 threads per block.
static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_THREADS / 64; i++) {
    if (i < w) woff += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return x - v + woff;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile_sums(const uint32_t* in, uint32_t n, uint32_t* bsum) {
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++)
    if (base + q < n) s += in[base + q];
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// one block of SCAN_THREADS threads scans nb <= SCAN_TILE block sums in place
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_bsums(uint32_t* bsum, uint32_t nb, uint32_t* total) {
  const uint32_t base = threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    v[q] = (base + q < nb) ? bsum[base + q] : 0u;
    s += v[q];
  }
  uint32_t tot;
  uint32_t off = block_excl_scan(s, &tot);
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    if (base + q < nb) bsum[base + q] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) *total = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* in, uint32_t n, const uint32_t* bsum, uint32_t add,
                                                             uint32_t* out) {
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    v[q] = (base + q < n) ? in[base + q] : 0u;
    s += v[q];
  }
  uint32_t tot;
  uint32_t off = block_excl_scan(s, &tot) + bsum[blockIdx.x] + add;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; q++) {
    if (base + q < n) out[base + q] = off;
    off += v[q];
  }
}

// in and out may alias.  bsum needs ceil(n/SCAN_TILE) <= SCAN_TILE entries.
void scan_exclusive(const uint32_t* in, uint32_t n, uint32_t add, uint32_t* out, uint32_t* bsum, uint32_t* total,
                    hipStream_t st) {
  uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0) nb = 1;
  k_scan_tile_sums<<<nb, SCAN_THREADS, 0, st>>>(in, n, bsum);
  k_scan_bsums<<<1, SCAN_THREADS, 0, st>>>(bsum, nb, total);
  k_scan_apply<<<nb, SCAN_THREADS, 0, st>>>(in, n, bsum, add, out);
}

uint32_t scan_max_items() { return (uint32_t)SCAN_TILE * (uint32_t)SCAN_TILE; }

// --------------------------------------------------------------- move_particles
// Re-binning (move_particles_across_cells.h:78-235) with the order one OpenMP thread
// produces: stayers packed by swap-with-last, then incoming movers in source-slot order
// (source cells are visited in index order, particles in cell order, and inner slots are
// laid out in exactly that order).

__device__ inline void move_flag_item(Geom g, uint32_t N, const double4* __restrict__ r, const uint32_t* __restrict__ slot_cell,
                            uint32_t* __restrict__ mv_dst, uint32_t* __restrict__ cnt_out, uint32_t* __restrict__ cnt_in,
                            uint32_t* __restrict__ err, uint32_t s) {
  const uint32_t c = slot_cell[s];
  int si, sj, sk;
  cell_ijk(g, c, si, sj, sk);
  const double4 p = r[s];
  double x = p.x, y = p.y, z = p.z;
  int li, lj, lk;
  periodic_location(g, x, y, z, li, lj, lk);
  const int di = li + g.gl, dj = lj + g.gl, dk = lk + g.gl;
  bool moved = (di != si || dj != sj || dk != sk);
  if (!moved) {
    // min_distance2_between(ro, cell_bounds(dst)) >= epsilon_cell_size2 (grid.h:60,89-90)
    const double eps = g.cs * (1.0 / 281474976710656.0);
    const double eps2 = eps * eps;
    const double b0[3] = {cell_origin(g.ox, g.gl, di, g.cs), cell_origin(g.oy, g.gl, dj, g.cs), cell_origin(g.oz, g.gl, dk, g.cs)};
    const double b1[3] = {cell_origin(g.ox, g.gl, di + 1, g.cs), cell_origin(g.oy, g.gl, dj + 1, g.cs),
                          cell_origin(g.oz, g.gl, dk + 1, g.cs)};
    const double ro[3] = {p.x, p.y, p.z};
    double d[3];
    for (int a = 0; a < 3; a++) {
      d[a] = 0.;
      if (ro[a] < b0[a]) d[a] = b0[a] - ro[a];
      else if (ro[a] > b1[a]) d[a] = ro[a] - b1[a];
    }
    moved = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) >= eps2;
  }
  if (!moved) {
    mv_dst[s] = NONE;
    return;
  }
  if (di < g.gl || dj < g.gl || dk < g.gl || di >= g.di - g.gl || dj >= g.dj - g.gl || dk >= g.dk - g.gl) {
    atomicOr(err, 1u);  // leaves the domain on a non-periodic axis (otb_particles)
    mv_dst[s] = NONE;
    return;
  }
  const uint32_t dc = cell_index(g, di, dj, dk);
  mv_dst[s] = dc;
  atomicAdd(&cnt_out[c], 1u);
  atomicAdd(&cnt_in[dc], 1u);
}

__global__ void k_move_flag(Geom g, uint32_t N, const double4* __restrict__ r, const uint32_t* __restrict__ slot_cell,
                            uint32_t* __restrict__ mv_dst, uint32_t* __restrict__ cnt_out, uint32_t* __restrict__ cnt_in,
                            uint32_t* __restrict__ err) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  move_flag_item(g, N, r, slot_cell, mv_dst, cnt_out, cnt_in, err, s);
}

// one thread per cell: swap-with-last packing (move_particles_across_cells.h:186-209)
__device__ inline void move_pack_item(Geom g, const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                            const uint32_t* __restrict__ cnt_out, const uint32_t* __restrict__ cnt_in,
                            const uint32_t* __restrict__ mv_dst, uint32_t* __restrict__ pk, uint32_t* __restrict__ newloc,
                            uint32_t* __restrict__ new_n, uint32_t c) {
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  if (is_ghost(g, i, j, k)) {
    new_n[c] = 0;
    return;
  }
  const uint32_t n = cell_n[c], off = cell_off[c], nout = cnt_out[c];
  new_n[c] = n - nout + cnt_in[c];
  if (nout == 0) return;  // every particle keeps its index (the scatter derives it)
  for (uint32_t q = 0; q < n; q++) pk[off + q] = (mv_dst[off + q] == NONE) ? q : NONE;
  int m = (int)n;
  for (int q = (int)n - 1; q >= 0; q--) {
    if (pk[off + q] == NONE) {
      if (q < m - 1) pk[off + q] = pk[off + m - 1];
      --m;
    }
  }
  for (int q = 0; q < m; q++) newloc[off + pk[off + q]] = (uint32_t)q;
}

__global__ void k_move_pack(Geom g, const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n,
                            const uint32_t* __restrict__ cnt_out, const uint32_t* __restrict__ cnt_in,
                            const uint32_t* __restrict__ mv_dst, uint32_t* __restrict__ pk, uint32_t* __restrict__ newloc,
                            uint32_t* __restrict__ new_n) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells) return;
  move_pack_item(g, cell_off, cell_n, cnt_out, cnt_in, mv_dst, pk, newloc, new_n, c);
}

// forces false: f and ep are not carried (inside a run loop, where the next force kernel
// overwrites both before anything reads them)
__device__ inline void copy_particle(const PState& A, uint32_t s, const PState& B, uint32_t t, bool forces) {
  B.r[t] = A.r[s];
  B.vx[t] = A.vx[s]; B.vy[t] = A.vy[s]; B.vz[t] = A.vz[s];
  if (forces) {
    B.fx[t] = A.fx[s]; B.fy[t] = A.fy[s]; B.fz[t] = A.fz[s];
    B.ep[t] = A.ep[s];
  }
  B.id[t] = A.id[s];
  // constant indices keep the pointer arrays in registers (a runtime index would move the
  // whole state to scratch)
#pragma unroll
  for (uint32_t k = 0; k < kMaxExtraFields; k++)
    if (k < A.nx) B.x[k][t] = A.x[k][s];
}

// newloc is written only for cells that lose particles; in the others a particle keeps its
// index, s - cell_off[c].
__device__ inline void move_scatter_item(uint32_t N, PState A, PState B, const uint32_t* __restrict__ slot_cell,
                               const uint32_t* __restrict__ mv_dst, const uint32_t* __restrict__ newloc,
                               const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ new_off,
                               const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ cnt_out,
                               uint32_t* __restrict__ in_cur, uint32_t* __restrict__ in_list, uint32_t* __restrict__ slot_cell_B,
                               uint32_t s, bool forces) {
  const uint32_t c = slot_cell[s];
  const uint32_t dc = mv_dst[s];
  if (dc == NONE) {
    const uint32_t t = new_off[c] + (cnt_out[c] ? newloc[s] : s - cell_off[c]);
    copy_particle(A, s, B, t, forces);
    slot_cell_B[t] = c;
  } else {
    const uint32_t pos = atomicAdd(&in_cur[dc], 1u);
    in_list[new_off[dc] + (cell_n[dc] - cnt_out[dc]) + pos] = s;
  }
}

__global__ void k_move_scatter(uint32_t N, PState A, PState B, const uint32_t* __restrict__ slot_cell,
                               const uint32_t* __restrict__ mv_dst, const uint32_t* __restrict__ newloc,
                               const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ new_off,
                               const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ cnt_out,
                               uint32_t* __restrict__ in_cur, uint32_t* __restrict__ in_list, uint32_t* __restrict__ slot_cell_B,
                               int forces) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  move_scatter_item(N, A, B, slot_cell, mv_dst, newloc, cell_off, new_off, cell_n, cnt_out, in_cur, in_list, slot_cell_B, s,
                    forces != 0);
}

// one thread per cell: order incoming movers by source slot, copy them with wrapped positions
__device__ inline void move_incoming_item(Geom g, PState A, PState B, const uint32_t* __restrict__ new_off,
                                const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ cnt_out,
                                const uint32_t* __restrict__ cnt_in, uint32_t* __restrict__ in_list,
                                uint32_t* __restrict__ slot_cell_B, uint32_t c, bool forces) {
  const uint32_t nin = cnt_in[c];
  if (nin == 0) return;
  const uint32_t b = new_off[c] + (cell_n[c] - cnt_out[c]);
  for (uint32_t q = 1; q < nin; q++) {  // insertion sort (few movers per cell)
    uint32_t v = in_list[b + q];
    int p = (int)q - 1;
    while (p >= 0 && in_list[b + p] > v) {
      in_list[b + p + 1] = in_list[b + p];
      --p;
    }
    in_list[b + p + 1] = v;
  }
  for (uint32_t q = 0; q < nin; q++) {
    const uint32_t s = in_list[b + q], t = b + q;
    copy_particle(A, s, B, t, forces);
    double4 p = A.r[s];
    int li, lj, lk;
    periodic_location(g, p.x, p.y, p.z, li, lj, lk);
    B.r[t] = p;
    slot_cell_B[t] = c;
  }
}

__global__ void k_move_incoming(Geom g, PState A, PState B, const uint32_t* __restrict__ new_off,
                                const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ cnt_out,
                                const uint32_t* __restrict__ cnt_in, uint32_t* __restrict__ in_list,
                                uint32_t* __restrict__ slot_cell_B, int forces) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells) return;
  move_incoming_item(g, A, B, new_off, cell_n, cnt_out, cnt_in, in_list, slot_cell_B, c, forces != 0);
}

__device__ inline void set_inner_cells_item(uint32_t ncells, const uint32_t* __restrict__ new_n, const uint32_t* __restrict__ new_off,
                                  Geom g, uint32_t* __restrict__ cell_n, uint32_t* __restrict__ cell_off, uint32_t c) {
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  if (is_ghost(g, i, j, k)) return;
  cell_n[c] = new_n[c];
  cell_off[c] = new_off[c];
}

__global__ void k_set_inner_cells(uint32_t ncells, const uint32_t* __restrict__ new_n, const uint32_t* __restrict__ new_off,
                                  Geom g, uint32_t* __restrict__ cell_n, uint32_t* __restrict__ cell_off) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  set_inner_cells_item(ncells, new_n, new_off, g, cell_n, cell_off, c);
}

// -------------------------------------------------------------------- rebuild_amr
// rebuild_amr (rebuild_amr.cpp:34-69, enforced_ordering 1, no Z order) =
// project_particles_in_sub_grids (amr_grid_algorithm.h:112-374).  A cell of n particles has a
// sub-grid of side sub_grid_size(n, density) (amr_grid_algorithm.h:66-77); a cell whose side
// exceeds 1 gets its particles stably sorted by sub-cell (k s + j) s + i, each index
// clamp((long)(pcoord * s), 0, s - 1) with pcoord = (r - cell_low) / cell_size
// (Grid::particle_pcoord, grid.h:184-190).  Runs after move_particles, when ghost cells are
// empty, so only inner cells are visited.

// floor(cbrt(n / density)), 1 below 2, at most 16, as the largest s with s^3 <= n / density:
// the same integer wherever n / density is not within an ulp below a perfect cube (never for
// density 6.5: 2n / 13 is an integer or at least 1/13 from one)
__host__ __device__ inline int amr_side(uint32_t n, double density) {
  if (n == 0) return 0;
  const double x = (double)n / density;
  if (!(x >= 8.0)) return 1;
  int s = 2;
  while (s < 16 && (double)((s + 1) * (s + 1) * (s + 1)) <= x) s++;
  return s;
}

// Dense cells are listed first (one thread per cell, rare atomics), then each listed cell is
// sorted by one wave: sub-cell keys in LDS, each lane's particles ranked against all keys
// (broadcast LDS reads; stable: keys below plus equal keys earlier in the cell) and copied
// from A into B at their rank, and in a last pass copied back from B into A.  The copy-back
// pass is its own launch, so every wave's writes to B are complete before it reads them.
constexpr uint32_t kAmrMaxKeys = 4096;  // particles of one cell whose keys fit the wave's LDS
constexpr uint32_t kAmrGrid = 1024;     // waves of the sort and copy-back launches

// inner cell c: appended to list (entry tag | c) if its sub-grid side exceeds 1; a cell of
// more than kAmrMaxKeys particles sets too_dense instead
__device__ inline void amr_mark_item(const Geom& g, double density, const uint32_t* __restrict__ cell_n, uint32_t c, uint32_t tag,
                                     uint32_t* __restrict__ list, uint32_t* __restrict__ count, uint32_t* __restrict__ too_dense,
                                     uint32_t flag) {
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  if (is_ghost(g, i, j, k)) return;
  const uint32_t n = cell_n[c];
  if (amr_side(n, density) <= 1) return;
  if (n > kAmrMaxKeys) {
    atomicOr(too_dense, flag);
    return;
  }
  list[atomicAdd(count, 1u)] = tag | c;
}

// one wave (the whole 64-thread workgroup) sorts cell c of A into B
__device__ inline void amr_sort_wave(const Geom& g, double density, const PState& A, const PState& B,
                                     const uint32_t* __restrict__ cell_off, const uint32_t* __restrict__ cell_n, uint32_t c,
                                     uint32_t* keys, bool forces) {
  const uint32_t lane = threadIdx.x;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  const uint32_t n = cell_n[c];
  const int sg = amr_side(n, density);
  const double lo[3] = {g.ox + (double)(i - g.gl) * g.cs, g.oy + (double)(j - g.gl) * g.cs, g.oz + (double)(k - g.gl) * g.cs};
  const uint32_t off = cell_off[c];
  for (uint32_t p = lane; p < n; p += 64) {
    const double4 r = A.r[off + p];
    const double x[3] = {r.x, r.y, r.z};
    int q[3];
    for (int a = 0; a < 3; a++) {
      const long long v = (long long)(((x[a] - lo[a]) / g.cs) * (double)sg);
      q[a] = v < 0 ? 0 : (v > sg - 1 ? sg - 1 : (int)v);
    }
    keys[p] = (uint32_t)((q[2] * sg + q[1]) * sg + q[0]);
  }
  __syncthreads();
  for (uint32_t p = lane; p < n; p += 64) {
    const uint32_t kp = keys[p];
    uint32_t rank = 0;
    for (uint32_t t = 0; t < n; t++) {
      const uint32_t kt = keys[t];
      rank += (kt < kp || (kt == kp && t < p)) ? 1u : 0u;
    }
    copy_particle(A, off + p, B, off + rank, forces);
  }
  __syncthreads();  // keys are rewritten by the workgroup's next cell
}

__device__ inline void amr_back_wave(const PState& A, const PState& B, const uint32_t* __restrict__ cell_off,
                                     const uint32_t* __restrict__ cell_n, uint32_t c, bool forces) {
  const uint32_t n = cell_n[c], off = cell_off[c];
  for (uint32_t p = threadIdx.x; p < n; p += 64) copy_particle(B, off + p, A, off + p, forces);
}

__global__ void k_amr_mark(Geom g, double density, const uint32_t* __restrict__ cell_n, uint32_t* __restrict__ list,
                           uint32_t* __restrict__ count, uint32_t* __restrict__ too_dense) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells) return;
  amr_mark_item(g, density, cell_n, c, 0u, list, count, too_dense, 2u);
}

__global__ void __launch_bounds__(64) k_amr_sort(Geom g, double density, PState A, PState B, const uint32_t* __restrict__ cell_off,
                                                 const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ count, int forces) {
  __shared__ uint32_t keys[kAmrMaxKeys];
  const uint32_t m = *count;
  for (uint32_t q = blockIdx.x; q < m; q += gridDim.x)
    amr_sort_wave(g, density, A, B, cell_off, cell_n, list[q], keys, forces != 0);
}

__global__ void __launch_bounds__(64) k_amr_back(PState A, PState B, const uint32_t* __restrict__ cell_off,
                                                 const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ count, int forces) {
  const uint32_t m = *count;
  for (uint32_t q = blockIdx.x; q < m; q += gridDim.x) amr_back_wave(A, B, cell_off, cell_n, list[q], forces != 0);
}

// ------------------------------------------------------------------------ ghosts
// ghost_comm_scheme + ghost_update_all on one rank (update_ghosts_comm_scheme.cpp:400-470):
// ghost cell at domain location gd takes inner cell gd - shift*D; a source particle is kept
// iff its shifted image is inside enlarge(domain bounds, ghost_dist), bounds inclusive.

__device__ inline bool ghost_source(const Geom& g, uint32_t c, uint32_t& src, int& si, int& sj, int& sk) {
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  if (!is_ghost(g, i, j, k)) return false;
  const int gi = i - g.gl, gj = j - g.gl, gk = k - g.gl;
  si = gi < 0 ? -1 : (gi >= g.Di ? 1 : 0);
  sj = gj < 0 ? -1 : (gj >= g.Dj ? 1 : 0);
  sk = gk < 0 ? -1 : (gk >= g.Dk ? 1 : 0);
  if ((si && !g.pi) || (sj && !g.pj) || (sk && !g.pk)) return false;
  src = cell_index(g, gi - si * g.Di + g.gl, gj - sj * g.Dj + g.gl, gk - sk * g.Dk + g.gl);
  return true;
}

__device__ inline bool ghost_keep(const Geom& g, double4 p, int si, int sj, int sk, double4& q) {
  const double bx0 = ((double)0 * g.cs + g.ox) - g.ghost_dist, bx1 = ((double)g.Di * g.cs + g.ox) + g.ghost_dist;
  const double by0 = ((double)0 * g.cs + g.oy) - g.ghost_dist, by1 = ((double)g.Dj * g.cs + g.oy) + g.ghost_dist;
  const double bz0 = ((double)0 * g.cs + g.oz) - g.ghost_dist, bz1 = ((double)g.Dk * g.cs + g.oz) + g.ghost_dist;
  q.x = shift_coord(p.x, si, g.ox, g.ex);
  q.y = shift_coord(p.y, sj, g.oy, g.ey);
  q.z = shift_coord(p.z, sk, g.oz, g.ez);
  q.w = p.w;
  return q.x >= bx0 && q.x <= bx1 && q.y >= by0 && q.y <= by1 && q.z >= bz0 && q.z <= bz1;
}

// Ghost cells are handled one wave per cell (only the ghost cells are visited; gcnt of inner
// cells stays 0 from allocation): lanes take the source cell's particles 64 at a time and the
// kept ones are compacted in source order by a ballot prefix.
__device__ inline void ghost_count_wave(const Geom& g, const double4* __restrict__ r, const uint32_t* __restrict__ cell_off,
                                        const uint32_t* __restrict__ cell_n, uint32_t* __restrict__ gcnt, uint32_t c) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t src;
  int si, sj, sk;
  uint32_t cnt = 0;
  if (ghost_source(g, c, src, si, sj, sk)) {
    const uint32_t off = cell_off[src], n = cell_n[src];
    for (uint32_t q0 = 0; q0 < n; q0 += 64) {
      double4 gp;
      const bool keep = q0 + lane < n && ghost_keep(g, r[off + q0 + lane], si, sj, sk, gp);
      cnt += (uint32_t)__popcll(__ballot(keep));
    }
  }
  if (lane == 0) gcnt[c] = cnt;
}

// wave w of the launch -> ghost cell w (wave-uniform; blockDim.x is a multiple of 64)
__device__ inline bool ghost_wave_cell(const Geom& g, uint32_t thread, uint32_t& c) {
  const uint32_t w = thread >> 6;
  if (w >= ghost_cell_count(g)) return false;
  c = ghost_cell_at(g, w);
  return true;
}

__global__ void __launch_bounds__(256) k_ghost_count(Geom g, const double4* __restrict__ r, const uint32_t* __restrict__ cell_off,
                                                     const uint32_t* __restrict__ cell_n, uint32_t* __restrict__ gcnt) {
  uint32_t c;
  if (!ghost_wave_cell(g, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
  ghost_count_wave(g, r, cell_off, cell_n, gcnt, c);
}

// Ghost cell c of one box (one wave).  Slots past cap are not written: the cell count is still
// stored and *status gets bit 2, so the caller sizes the buffers and runs the fill again.
// Each image copies every field of its source, forces and energy included, like
// ghost_update_all (update_ghosts.cu:45, AddDefaultFields of the grid's fields); forces false
// (run loops, whose re-binning does not carry f and ep) writes 0 to f and ep instead.
__device__ inline void ghost_fill_wave(const Geom& g, uint32_t N, const PState& B, const uint32_t* __restrict__ goff,
                                       uint32_t* __restrict__ cell_off, uint32_t* __restrict__ cell_n, uint32_t* __restrict__ slot_cell,
                                       uint32_t* __restrict__ ghost_src, uint32_t* __restrict__ ghost_shift, uint32_t cap,
                                       uint32_t* __restrict__ status, uint32_t c, bool forces) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t src;
  int si, sj, sk;
  const uint32_t base = N + goff[c];
  uint32_t cnt = 0;
  if (ghost_source(g, c, src, si, sj, sk)) {
    const uint32_t off = cell_off[src], n = cell_n[src];
    const uint32_t code = pack_shift(si, sj, sk);
    for (uint32_t q0 = 0; q0 < n; q0 += 64) {
      const uint32_t s = off + q0 + lane;
      double4 gp;
      const bool keep = q0 + lane < n && ghost_keep(g, B.r[s], si, sj, sk, gp);
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const uint32_t t = base + cnt + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (t >= cap) {
          atomicOr(status, 4u);
        } else {
          B.r[t] = gp;
          B.vx[t] = B.vx[s]; B.vy[t] = B.vy[s]; B.vz[t] = B.vz[s];
          if (forces) {
            B.fx[t] = B.fx[s]; B.fy[t] = B.fy[s]; B.fz[t] = B.fz[s]; B.ep[t] = B.ep[s];
          } else {
            B.fx[t] = 0.; B.fy[t] = 0.; B.fz[t] = 0.; B.ep[t] = 0.;
          }
          B.id[t] = B.id[s];
#pragma unroll
          for (uint32_t k = 0; k < kMaxExtraFields; k++)
            if (k < B.nx) B.x[k][t] = B.x[k][s];
          slot_cell[t] = c;
          ghost_src[t - N] = s;
          ghost_shift[t - N] = code;
        }
      }
      cnt += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0) {
    cell_off[c] = base;
    cell_n[c] = cnt;
  }
}

__global__ void __launch_bounds__(256) k_ghost_fill(Geom g, uint32_t N, PState B, const uint32_t* __restrict__ goff,
                                                    uint32_t* __restrict__ cell_off, uint32_t* __restrict__ cell_n,
                                                    uint32_t* __restrict__ slot_cell, uint32_t* __restrict__ ghost_src,
                                                    uint32_t* __restrict__ ghost_shift, uint32_t cap, uint32_t* __restrict__ status,
                                                    int forces) {
  uint32_t c;
  if (!ghost_wave_cell(g, blockIdx.x * blockDim.x + threadIdx.x, c)) return;
  ghost_fill_wave(g, N, B, goff, cell_off, cell_n, slot_cell, ghost_src, ghost_shift, cap, status, c, forces != 0);
}

// ghost_update_r (update_ghosts.cu:58): re-copy shifted positions along the stored scheme
__device__ inline void ghost_update_r_item(Geom g, uint32_t N, uint32_t G, double4* __restrict__ r, const uint32_t* __restrict__ ghost_src,
                                 const uint32_t* __restrict__ ghost_shift, uint32_t q) {
  const uint32_t code = ghost_shift[q];
  const uint32_t src = ghost_src[q];
  if (src >= N) return;  // a ghost's source is an inner slot; anything else is a stale scheme
  const double4 p = r[src];
  double4 o;
  o.x = shift_coord(p.x, (int)(code & 3u) - 1, g.ox, g.ex);
  o.y = shift_coord(p.y, (int)((code >> 2) & 3u) - 1, g.oy, g.ey);
  o.z = shift_coord(p.z, (int)((code >> 4) & 3u) - 1, g.oz, g.ez);
  o.w = p.w;
  r[N + q] = o;
}

__global__ void k_ghost_update_r(Geom g, uint32_t N, uint32_t G, double4* __restrict__ r, const uint32_t* __restrict__ ghost_src,
                                 const uint32_t* __restrict__ ghost_shift) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= G) return;
  ghost_update_r_item(g, N, G, r, ghost_src, ghost_shift, q);
}

// --------------------------------------------------------------- backup / trigger

__device__ inline void backup_r_item(Geom g, uint32_t N, const double4* __restrict__ r, const uint32_t* __restrict__ slot_cell,
                           uint32_t* __restrict__ backup, uint32_t s) {
  int i, j, k;
  cell_ijk(g, slot_cell[s], i, j, k);
  const double4 p = r[s];
  backup[3 * (size_t)s + 0] = encode_double_u32(p.x, cell_origin(g.ox, g.gl, i, g.cs), g.cs);
  backup[3 * (size_t)s + 1] = encode_double_u32(p.y, cell_origin(g.oy, g.gl, j, g.cs), g.cs);
  backup[3 * (size_t)s + 2] = encode_double_u32(p.z, cell_origin(g.oz, g.gl, k, g.cs), g.cs);
}

__global__ void k_backup_r(Geom g, uint32_t N, const double4* __restrict__ r, const uint32_t* __restrict__ slot_cell,
                           uint32_t* __restrict__ backup) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  backup_r_item(g, N, r, slot_cell, backup, s);
}

// particle_displ_over (particle_displ_over.cu:38-77)
__global__ void __launch_bounds__(1024) k_displ_over(Geom g, uint32_t N, const double4* __restrict__ r, const uint32_t* __restrict__ slot_cell,
                             const uint32_t* __restrict__ backup, double thr2, unsigned long long* __restrict__ count) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  bool over = false;
  if (s < N) over = displaced(g, r[s], slot_cell[s], backup, s, thr2);
  block_count_add(over, count);
}

// -------------------------------------------------------------------- integrators

// push_f_v_r (push_vec3_2nd_order.h:33-38): x += v*dt + f*dt2
__global__ void k_push_f_v_r(uint32_t N, double4* __restrict__ r, const double* __restrict__ vx, const double* __restrict__ vy,
                             const double* __restrict__ vz, const double* __restrict__ fx, const double* __restrict__ fy,
                             const double* __restrict__ fz, double dt, double dt2) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  double4 p = r[s];
  p.x += vx[s] * dt + fx[s] * dt2;
  p.y += vy[s] * dt + fy[s] * dt2;
  p.z += vz[s] * dt + fz[s] * dt2;
  r[s] = p;
}

// push_f_v (push_vec3_1st_order.h:33-38): v += f*dt
__global__ void k_push_f_v(uint32_t N, double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                           const double* __restrict__ fx, const double* __restrict__ fy, const double* __restrict__ fz, double dt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  vx[s] += fx[s] * dt;
  vy[s] += fy[s] * dt;
  vz[s] += fz[s] * dt;
}

// verlet_first_half fused: push_f_v_r(1.0) then push_f_v(0.5), then the
// particle_displ_over test of the new position (same arithmetic as the separate kernels).
__device__ inline bool verlet_first_half_item(const Geom& g, double4* __restrict__ r, double* __restrict__ vx,
                                              double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ fx,
                                              const double* __restrict__ fy, const double* __restrict__ fz, double dt, double dt2,
                                              double dth, const uint32_t* __restrict__ slot_cell, const uint32_t* __restrict__ backup,
                                              double thr2, uint32_t s) {
  double4 p = r[s];
  double ux = vx[s], uy = vy[s], uz = vz[s];
  const double ax = fx[s], ay = fy[s], az = fz[s];
  p.x += ux * dt + ax * dt2;
  p.y += uy * dt + ay * dt2;
  p.z += uz * dt + az * dt2;
  r[s] = p;
  ux += ax * dth;
  uy += ay * dth;
  uz += az * dth;
  vx[s] = ux;
  vy[s] = uy;
  vz[s] = uz;
  return displaced(g, p, slot_cell[s], backup, s, thr2);
}

__global__ void __launch_bounds__(1024) k_verlet_first_half(Geom g, uint32_t N, double4* __restrict__ r, double* __restrict__ vx,
                                    double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ fx,
                                    const double* __restrict__ fy, const double* __restrict__ fz, double dt, double dt2,
                                    double dth, const uint32_t* __restrict__ slot_cell, const uint32_t* __restrict__ backup,
                                    double thr2, unsigned long long* __restrict__ count) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  bool over = false;
  if (s < N) over = verlet_first_half_item(g, r, vx, vy, vz, fx, fy, fz, dt, dt2, dth, slot_cell, backup, thr2, s);
  block_count_add(over, count);
}

__global__ void k_zero_force(uint32_t n, double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
                             double* __restrict__ ep) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  fx[s] = 0.;
  fy[s] = 0.;
  fz[s] = 0.;
  ep[s] = 0.;
}

// ------------------------------------------------------------- chunk_neighbors
// One thread per atom.  Candidate cells are visited in (k, j, i) order and particles in
// cell order, which is the ascending (cell_enc, p_b) order the reference obtains by
// std::sort (chunk_neighbors_execute.h:308), so entries come out already sorted.
// The cutoff+skin test is FP64 on d = r_a - r_b (chunk_neighbors_execute.h:225-239).

__global__ void __launch_bounds__(256) k_nbh_build(Geom g, uint32_t n_atoms, const double4* __restrict__ r,
                                                   const uint32_t* __restrict__ slot_cell, const uint32_t* __restrict__ cell_off,
                                                   const uint32_t* __restrict__ cell_n, double max2, int lay, int half_symmetric,
                                                   int skip_ghosts, uint32_t* __restrict__ nbr, uint32_t* __restrict__ nbr_cnt,
                                                   uint32_t cap, uint32_t* __restrict__ max_cnt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_atoms) return;
  const uint32_t ca = slot_cell[s];
  int ai, aj, ak;
  cell_ijk(g, ca, ai, aj, ak);
  const uint32_t pa = s - cell_off[ca];
  const double4 ra = r[s];
  // candidate cells [a - lay, a + lay] clamped to the grid (chunk_neighbors_execute.h:134-139)
  const int i0 = max(ai - lay, 0), i1 = min(ai + lay, g.di - 1);
  const int j0 = max(aj - lay, 0), j1 = min(aj + lay, g.dj - 1);
  const int k0 = max(ak - lay, 0), k1 = min(ak + lay, g.dk - 1);
  uint32_t cnt = 0;
  uint32_t* L = nbr + nbr_index(s, 0, cap);
  for (int bk = k0; bk <= k1; bk++)
    for (int bj = j0; bj <= j1; bj++)
      for (int bi = i0; bi <= i1; bi++) {
        const uint32_t cb = cell_index(g, bi, bj, bk);
        const uint32_t off = cell_off[cb], nb = cell_n[cb];
        const bool bghost = is_ghost(g, bi, bj, bk);
        for (uint32_t pb = 0; pb < nb; pb++) {
          const double4 rb = r[off + pb];
          const double dx = ra.x - rb.x;
          const double dy = ra.y - rb.y;
          const double dz = ra.z - rb.z;
          const double d2 = (dx * dx + dy * dy) + dz * dz;
          if ((ca != cb || pa != pb) && d2 > 0.0 && d2 <= max2) {
            if (half_symmetric && (ca < cb || (ca == cb && pa < pb))) continue;
            if (skip_ghosts && bghost) continue;
            if (cnt < cap) L[(size_t)cnt << 6] = off + pb;
            ++cnt;
          }
        }
      }
  nbr_cnt[s] = cnt;
  atomicMax(max_cnt, cnt);
}

// ------------------------------------------------------------------- LJ forces
// Pair arithmetic (lj_pair_fp64 / lj_pair_mixed) lives in xnb_pair.h.

// compute_pair_singlemat over per-slot lists (all cells when the lists cover the ghosts) with a
// force functor F (include/xnb_compute_pair.h), calling its buffer-less form.  Thread per atom,
// neighbours in list (= stream) order.  The trip count is the wave's largest list; lanes past
// their own count read their own slot, which the d2 > 0 test rejects exactly as the reference
// rejects self pairs.  U list rows are loaded, then U positions, then U pairs are evaluated, so
// each wave keeps 3U loads in flight.  Requires cap % U == 0.  Summation as in the brick
// traversal: a ComputeBufferCompatible functor sums from 0 and stores f = 0 + sum (f += sum with
// ACCUM, after zero_particle_force); a buffer-less-only functor adds into the stored force
// (from 0 without ACCUM).  PUSH: fused push_f_v(0.5) on the new force.
template <class F, bool ACCUM, bool PUSH, int U>
__global__ void __launch_bounds__(256) k_compute_pair_slots(uint32_t n_atoms, const double4* __restrict__ r,
                                                            const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ nbr_cnt,
                                                            uint32_t cap, F func, double rc2, double* __restrict__ fx,
                                                            double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ ep,
                                                            double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                                                            double dth) {
  using Tr = ComputePairTraits<F>;
  static_assert(Tr::BufferLessCompatible, "the per-slot traversal calls the buffer-less form");
  constexpr bool BUFFERED = Tr::ComputeBufferCompatible;
  const uint32_t s0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s0 < n_atoms;
  const uint32_t s = live ? s0 : n_atoms - 1;
  const uint32_t n = live ? nbr_cnt[s] : 0u;
  const uint32_t nmax = wave_max_u32(n);
  const double4 ra = r[s];
  const int type_a = Tr::CentralType ? (int)ra.w : 0;
  const uint32_t* L = nbr + nbr_index(s, 0, cap);
  double tx = 0.0, ty = 0.0, tz = 0.0, te = 0.0;
  if (!BUFFERED && ACCUM) {
    tx = fx[s];
    ty = fy[s];
    tz = fz[s];
    if (Tr::CentralEnergy) te = ep[s];
  }
  const NbhCells cells{nullptr, r};
  for (uint32_t q0 = 0; q0 < nmax; q0 += U) {
    uint32_t j[U];
#pragma unroll
    for (int u = 0; u < U; u++) j[u] = L[(size_t)(q0 + u) << 6];  // rows < cap: always in bounds
#pragma unroll
    for (int u = 0; u < U; u++) j[u] = (q0 + u < n) ? j[u] : s;
    double4 rb[U];
#pragma unroll
    for (int u = 0; u < U; u++) rb[u] = r[j[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const Vec3d dr{rb[u].x - ra.x, rb[u].y - ra.y, rb[u].z - ra.z};
      const double d2 = (dr.x * dr.x + dr.y * dr.y) + dr.z * dr.z;
      if (d2 > 0.0 && d2 <= rc2) {
        if constexpr (Tr::CentralType && Tr::CentralEnergy)
          func(dr, d2, type_a, tx, ty, tz, te, cells, 0, j[u], 1.0);
        else if constexpr (Tr::CentralType)
          func(dr, d2, type_a, tx, ty, tz, cells, 0, j[u], 1.0);
        else if constexpr (Tr::CentralEnergy)
          func(dr, d2, tx, ty, tz, te, cells, 0, j[u], 1.0);
        else
          func(dr, d2, tx, ty, tz, cells, 0, j[u], 1.0);
      }
    }
  }
  if (!live) return;
  if (BUFFERED) {
    fx[s] = (ACCUM ? fx[s] : 0.0) + tx;
    fy[s] = (ACCUM ? fy[s] : 0.0) + ty;
    fz[s] = (ACCUM ? fz[s] : 0.0) + tz;
    if (Tr::CentralEnergy) ep[s] = (ACCUM ? ep[s] : 0.0) + te;
  } else {
    fx[s] = tx;
    fy[s] = ty;
    fz[s] = tz;
    if (Tr::CentralEnergy) ep[s] = te;
  }
  if (PUSH) {
    vx[s] += fx[s] * dth;
    vy[s] += fy[s] * dth;
    vz[s] += fz[s] * dth;
  }
}

// Multi-material LJ (compute_pair_multimat.h:100-253): accepted iff d2 <= rc(pair)^2, with
// the per-atom sums grouped by pair id (each group summed in stream order from 0, then added
// to f in increasing pair id).  At most MM_PAIRS pair ids (up to 4 types).
template <bool MIXED>
__global__ void __launch_bounds__(256) k_lj_multimat(uint32_t n_atoms, const double4* __restrict__ r, const uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ nbr_cnt, uint32_t cap, LJMulti lj,
                                                     double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
                                                     double* __restrict__ ep) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_atoms) return;
  const double4 ra = r[s];
  const unsigned ta = (unsigned)ra.w;
  const uint32_t n = nbr_cnt[s];
  const uint32_t* L = nbr + nbr_index(s, 0, cap);
  double ax = fx[s], ay = fy[s], az = fz[s], ae = ep[s];
  // group sums in increasing pair id: only pair ids involving type ta can occur
  for (int pid = 0; pid < lj.npairs; pid++) {
    double tx = 0.0, ty = 0.0, tz = 0.0, te = 0.0;
    bool any = false;
    for (uint32_t q = 0; q < n; q++) {
      const uint32_t j = L[(size_t)q << 6];
      if (j == s) continue;
      const double4 rb = r[j];
      const unsigned tb = (unsigned)rb.w;
      if ((int)unique_pair_id(ta, tb) != pid) continue;
      const double dx = rb.x - ra.x;
      const double dy = rb.y - ra.y;
      const double dz = rb.z - ra.z;
      if (MIXED) {
        const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
        if (d2 <= lj.rc2[pid]) {
          any = true;
          float s6;
          const double cd = (double)lj_coef_f32(__builtin_amdgcn_rcpf((float)d2), lj.sigma6f[pid], lj.m24epsf[pid], s6);
          tx = fma(cd, dx, tx);
          ty = fma(cd, dy, ty);
          tz = fma(cd, dz, tz);
          te += (double)lj_energy_f32(s6, lj.eps4f[pid]);
        }
      } else {
        const double d2 = (dx * dx + dy * dy) + dz * dz;
        if (d2 <= lj.rc2[pid]) {
          any = true;
          const double rr = sqrt(d2);
          const double inv_r = 1.0 / rr;
          const double ratio = lj.sigma[pid] * inv_r;
          const double ratio2 = ratio * ratio;
          const double ratio6 = ratio2 * ratio2 * ratio2;
          const double ratio12 = ratio6 * ratio6;
          const double e = lj.eps4[pid] * (ratio12 - ratio6);
          double de = (lj.m24eps[pid] * (2. * ratio12 - ratio6)) * inv_r;
          de *= inv_r;
          tx += de * dx;
          ty += de * dy;
          tz += de * dz;
          te += e;
        }
      }
    }
    if (any) {
      ax += tx;
      ay += ty;
      az += tz;
      ae += te;
    }
  }
  fx[s] = ax;
  fy[s] = ay;
  fz[s] = az;
  ep[s] = ae;
}

// pair statistics of the current list: list entries and in-cutoff pairs (FP64 test)
__global__ void __launch_bounds__(256) k_pair_counts(uint32_t n_atoms, const double4* __restrict__ r, const uint32_t* __restrict__ nbr,
                              const uint32_t* __restrict__ nbr_cnt, uint32_t cap, double rc2,
                              unsigned long long* __restrict__ out) {
  __shared__ unsigned long long acc[2];
  if (threadIdx.x == 0) acc[0] = acc[1] = 0;
  __syncthreads();
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n = 0, m = 0;
  if (s < n_atoms) {
    const double4 ra = r[s];
    n = nbr_cnt[s];
    const uint32_t* L = nbr + nbr_index(s, 0, cap);
    for (uint32_t q = 0; q < n; q++) {
      const double4 rb = r[L[(size_t)q << 6]];
      const double dx = rb.x - ra.x, dy = rb.y - ra.y, dz = rb.z - ra.z;
      const double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 > 0.0 && d2 <= rc2) ++m;
    }
  }
  atomicAdd(&acc[0], (unsigned long long)n);
  atomicAdd(&acc[1], (unsigned long long)m);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], acc[0]);
    atomicAdd(&out[1], acc[1]);
  }
}

// ----------------------------------------------------------------- stream export
// GridChunkNeighbors CS=1 stream (chunk_neighbors_execute.h:265-398), derived from the
// device neighbour list: per particle [ncg][cell_enc, n, p_b ...]..., optional offset tables.

__device__ inline void rel_of(const Geom& g, uint32_t ca, uint32_t cb, int& ri, int& rj, int& rk) {
  int ai, aj, ak, bi, bj, bk;
  cell_ijk(g, ca, ai, aj, ak);
  cell_ijk(g, cb, bi, bj, bk);
  ri = bi - ai;
  rj = bj - aj;
  rk = bk - ak;
}

// words of one particle's stream (excluding offset tables)
__global__ void k_stream_plen(uint32_t n_atoms, const uint32_t* __restrict__ slot_cell, const uint32_t* __restrict__ nbr,
                              const uint32_t* __restrict__ nbr_cnt, uint32_t cap, int random_access,
                              uint32_t* __restrict__ plen) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_atoms) return;
  const uint32_t n = nbr_cnt[s];
  const uint32_t* L = nbr + nbr_index(s, 0, cap);
  if (random_access) {
    plen[s] = 2 * n;
    return;
  }
  uint32_t groups = 0, last = NONE;
  for (uint32_t q = 0; q < n; q++) {
    const uint32_t cb = slot_cell[L[(size_t)q << 6]];
    if (cb != last) {
      ++groups;
      last = cb;
    }
  }
  plen[s] = 1 + 2 * groups + n;
}

// per cell: stream length (tables + particle streams) and per-particle start within cell
__global__ void k_stream_cell_len(Geom g, uint32_t n_atoms, const uint32_t* __restrict__ cell_off,
                                  const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ plen, int ntables,
                                  uint32_t* __restrict__ pstart, uint32_t* __restrict__ clen) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells) return;
  const uint32_t off = cell_off[c], n = cell_n[c];
  if (n == 0 || off + n > n_atoms) {
    clen[c] = 0;
    return;
  }
  const uint32_t table = ntables >= 1 ? (n + 1) * 2 + (ntables >= 2 ? n * 2 : 0) : 0;
  uint32_t acc = 0;
  for (uint32_t q = 0; q < n; q++) {
    pstart[off + q] = acc;
    acc += plen[off + q];
  }
  clen[c] = table + acc;
}

__global__ void k_stream_fill(Geom g, uint32_t n_atoms, const uint32_t* __restrict__ slot_cell, const uint32_t* __restrict__ cell_off,
                              const uint32_t* __restrict__ cell_n, const uint32_t* __restrict__ nbr,
                              const uint32_t* __restrict__ nbr_cnt, uint32_t cap, const uint32_t* __restrict__ pstart,
                              const uint32_t* __restrict__ plen, const uint32_t* __restrict__ coff, int ntables, int random_access,
                              uint16_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_atoms) return;
  const uint32_t ca = slot_cell[s];
  const uint32_t off = cell_off[ca], na = cell_n[ca], pa = s - off;
  const uint32_t table = ntables >= 1 ? (na + 1) * 2 + (ntables >= 2 ? na * 2 : 0) : 0;
  uint16_t* cs = out + coff[ca];
  uint16_t* st = cs + table + pstart[s];
  const uint32_t n = nbr_cnt[s];
  const uint32_t* L = nbr + nbr_index(s, 0, cap);
  uint32_t w = 0, n_sym = 0;
  uint32_t cg_idx = 0, ch_idx = 0, last = NONE;
  if (!random_access) st[w++] = 0;
  for (uint32_t q = 0; q < n; q++) {
    const uint32_t j = L[(size_t)q << 6];
    const uint32_t cb = slot_cell[j];
    const uint32_t pb = j - cell_off[cb];
    int ri, rj, rk;
    rel_of(g, ca, cb, ri, rj, rk);
    const uint16_t enc = encode_cell_rel(ri, rj, rk);
    if (cb < ca || (cb == ca && pb < pa)) ++n_sym;
    if (random_access) {
      st[w++] = enc;
      st[w++] = (uint16_t)pb;
    } else {
      if (cb != last) {
        ++cg_idx;
        last = cb;
        st[w++] = enc;
        ch_idx = w;
        st[w++] = 0;
      }
      st[ch_idx] = (uint16_t)(st[ch_idx] + 1);
      st[w++] = (uint16_t)pb;
    }
  }
  if (!random_access) st[0] = (uint16_t)cg_idx;
  if (ntables >= 1) {
    const uint32_t o = pstart[s] + (uint32_t)ntables;
    cs[pa * 2 + 0] = (uint16_t)o;
    cs[pa * 2 + 1] = (uint16_t)(o >> 16);
    if (pa == na - 1) {
      const uint32_t e = pstart[s] + plen[s] + (uint32_t)ntables;
      cs[na * 2 + 0] = (uint16_t)e;
      cs[na * 2 + 1] = (uint16_t)(e >> 16);
    }
    if (ntables >= 2) {
      cs[(na + 1 + pa) * 2 + 0] = (uint16_t)n_sym;
      cs[(na + 1 + pa) * 2 + 1] = (uint16_t)(n_sym >> 16);
    }
  }
}

// --------------------------------------------------------------------- energies
// Deterministic two-pass reduction: fixed per-block shuffle tree, then one block sums the
// block partials in index order.
static constexpr int RED_THREADS = 256;

__device__ inline double block_sum(double v) {
  __shared__ double ws[RED_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < RED_THREADS / 64; i++) t += ws[i];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(RED_THREADS) k_energy_partials(uint32_t N, const double* __restrict__ vx, const double* __restrict__ vy,
                                                                const double* __restrict__ vz, const double* __restrict__ ep,
                                                                double* __restrict__ part) {
  double ke = 0.0, pe = 0.0;
  for (uint32_t s = blockIdx.x * RED_THREADS + threadIdx.x; s < N; s += gridDim.x * RED_THREADS) {
    ke += 0.5 * ((vx[s] * vx[s] + vy[s] * vy[s]) + vz[s] * vz[s]);
    pe += 0.5 * ep[s];
  }
  const double k = block_sum(ke);
  const double p = block_sum(pe);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = k;
    part[2 * blockIdx.x + 1] = p;
  }
}

__global__ void k_energy_final(uint32_t nb, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double k = 0.0, p = 0.0;
    for (uint32_t b = 0; b < nb; b++) {
      k += part[2 * b];
      p += part[2 * b + 1];
    }
    out[0] = k;
    out[1] = p;
  }
}

// --------------------------------------------------------------------- copy probe
// Streams n 16-byte words from src to dst: the reference copy kernel used to calibrate HBM
// traffic counters (bytes read = bytes written = 16 n).
__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) dst[q] = src[q];
}

// device evaluation of the FP64 sqrt / reciprocal sequences of the LJ kernels
__global__ void k_fp64_core(const double* __restrict__ x, size_t n, double* __restrict__ sq, double* __restrict__ rc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sq[i] = sqrt_normal(x[i]);
  rc[i] = recip_normal(x[i]);
}

void launch_fp64_core(const double* x, size_t n, double* sq, double* rc, hipStream_t st) {
  if (n == 0) return;
  k_fp64_core<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x, n, sq, rc);
}

void launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  k_copy_probe<<<4096, 256, 0, st>>>((const uint4*)src, (uint4*)dst, n);
}

// ---------------------------------------------------------------------- launchers

static inline uint32_t nblk(uint32_t n, uint32_t t) { return n == 0 ? 1 : (n + t - 1) / t; }

void launch_move_flag(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, uint32_t* mv_dst, uint32_t* cnt_out,
                      uint32_t* cnt_in, uint32_t* err, hipStream_t st) {
  k_move_flag<<<nblk(N, 256), 256, 0, st>>>(g, N, r, slot_cell, mv_dst, cnt_out, cnt_in, err);
}
void launch_move_pack(const Geom& g, const uint32_t* cell_off, const uint32_t* cell_n, const uint32_t* cnt_out,
                      const uint32_t* cnt_in, const uint32_t* mv_dst, uint32_t* pk, uint32_t* newloc, uint32_t* new_n,
                      hipStream_t st) {
  k_move_pack<<<nblk(g.ncells, 128), 128, 0, st>>>(g, cell_off, cell_n, cnt_out, cnt_in, mv_dst, pk, newloc, new_n);
}
void launch_move_scatter(uint32_t N, const PState& A, const PState& B, const uint32_t* slot_cell, const uint32_t* mv_dst,
                         const uint32_t* newloc, const uint32_t* cell_off, const uint32_t* new_off, const uint32_t* cell_n,
                         const uint32_t* cnt_out, uint32_t* in_cur, uint32_t* in_list, uint32_t* slot_cell_B, bool forces,
                         hipStream_t st) {
  k_move_scatter<<<nblk(N, 256), 256, 0, st>>>(N, A, B, slot_cell, mv_dst, newloc, cell_off, new_off, cell_n, cnt_out, in_cur,
                                               in_list, slot_cell_B, forces ? 1 : 0);
}
void launch_move_incoming(const Geom& g, const PState& A, const PState& B, const uint32_t* new_off, const uint32_t* cell_n,
                          const uint32_t* cnt_out, const uint32_t* cnt_in, uint32_t* in_list, uint32_t* slot_cell_B,
                          bool forces, hipStream_t st) {
  k_move_incoming<<<nblk(g.ncells, 128), 128, 0, st>>>(g, A, B, new_off, cell_n, cnt_out, cnt_in, in_list, slot_cell_B,
                                                        forces ? 1 : 0);
}
void launch_set_inner_cells(const Geom& g, const uint32_t* new_n, const uint32_t* new_off, uint32_t* cell_n, uint32_t* cell_off,
                            hipStream_t st) {
  k_set_inner_cells<<<nblk(g.ncells, 256), 256, 0, st>>>(g.ncells, new_n, new_off, g, cell_n, cell_off);
}
void launch_rebuild_amr(const Geom& g, double density, const PState& A, const PState& B, const uint32_t* cell_off,
                        const uint32_t* cell_n, uint32_t* list, uint32_t* count, uint32_t* too_dense, bool forces, hipStream_t st) {
  k_amr_mark<<<nblk(g.ncells, 256), 256, 0, st>>>(g, density, cell_n, list, count, too_dense);
  k_amr_sort<<<kAmrGrid, 64, 0, st>>>(g, density, A, B, cell_off, cell_n, list, count, forces ? 1 : 0);
  k_amr_back<<<kAmrGrid, 64, 0, st>>>(A, B, cell_off, cell_n, list, count, forces ? 1 : 0);
}
void launch_ghost_count(const Geom& g, const double4* r, const uint32_t* cell_off, const uint32_t* cell_n, uint32_t* gcnt,
                        hipStream_t st) {
  k_ghost_count<<<nblk(ghost_cell_count(g) * 64u, 256), 256, 0, st>>>(g, r, cell_off, cell_n, gcnt);
}
void launch_ghost_fill(const Geom& g, uint32_t N, const PState& B, const uint32_t* goff, uint32_t* cell_off, uint32_t* cell_n,
                       uint32_t* slot_cell, uint32_t* ghost_src, uint32_t* ghost_shift, uint32_t cap, uint32_t* status, bool forces,
                       hipStream_t st) {
  k_ghost_fill<<<nblk(ghost_cell_count(g) * 64u, 256), 256, 0, st>>>(g, N, B, goff, cell_off, cell_n, slot_cell, ghost_src,
                                                                       ghost_shift, cap, status, forces ? 1 : 0);
}
void launch_ghost_update_r(const Geom& g, uint32_t N, uint32_t G, double4* r, const uint32_t* ghost_src,
                           const uint32_t* ghost_shift, hipStream_t st) {
  if (G == 0) return;
  k_ghost_update_r<<<nblk(G, 256), 256, 0, st>>>(g, N, G, r, ghost_src, ghost_shift);
}
void launch_backup_r(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, uint32_t* backup, hipStream_t st) {
  k_backup_r<<<nblk(N, 256), 256, 0, st>>>(g, N, r, slot_cell, backup);
}
void launch_displ_over(const Geom& g, uint32_t N, const double4* r, const uint32_t* slot_cell, const uint32_t* backup, double thr2,
                       unsigned long long* count, hipStream_t st) {
  k_displ_over<<<nblk(N, 1024), 1024, 0, st>>>(g, N, r, slot_cell, backup, thr2, count);
}
void launch_push_f_v_r(uint32_t N, const PState& A, double dt, double dt2, hipStream_t st) {
  k_push_f_v_r<<<nblk(N, 256), 256, 0, st>>>(N, A.r, A.vx, A.vy, A.vz, A.fx, A.fy, A.fz, dt, dt2);
}
void launch_push_f_v(uint32_t N, const PState& A, double dt, hipStream_t st) {
  k_push_f_v<<<nblk(N, 256), 256, 0, st>>>(N, A.vx, A.vy, A.vz, A.fx, A.fy, A.fz, dt);
}
void launch_verlet_first_half(const Geom& g, uint32_t N, const PState& A, double dt, double dt2, double dth,
                              const uint32_t* slot_cell, const uint32_t* backup, double thr2, unsigned long long* count,
                              hipStream_t st) {
  k_verlet_first_half<<<nblk(N, 1024), 1024, 0, st>>>(g, N, A.r, A.vx, A.vy, A.vz, A.fx, A.fy, A.fz, dt, dt2, dth, slot_cell, backup,
                                                    thr2, count);
}
// divide_force_by_type_scalar (vec3_typescalar_op.cu:52-120, 161): f /= value[type] per
// particle, the type taken from r.w; values of up to 256 types in a kernel argument
struct TypeScalars {
  double v[256];
};
__global__ void k_divide_force_by_type(uint32_t n, const double4* __restrict__ r, double* __restrict__ fx, double* __restrict__ fy,
                                       double* __restrict__ fz, TypeScalars ts) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double m = ts.v[(uint32_t)r[s].w & 255u];
  fx[s] = fx[s] / m;
  fy[s] = fy[s] / m;
  fz[s] = fz[s] / m;
}

void launch_divide_force_by_type(uint32_t n, const PState& A, const double* values, int ntypes, hipStream_t st) {
  TypeScalars ts{};
  for (int t = 0; t < 256; t++) ts.v[t] = t < ntypes ? values[t] : 1.0;
  if (n) k_divide_force_by_type<<<nblk(n, 256), 256, 0, st>>>(n, A.r, A.fx, A.fy, A.fz, ts);
}

void launch_zero_force(uint32_t n, const PState& A, hipStream_t st) {
  k_zero_force<<<nblk(n, 256), 256, 0, st>>>(n, A.fx, A.fy, A.fz, A.ep);
}
void launch_nbh_build(const Geom& g, uint32_t n_atoms, const double4* r, const uint32_t* slot_cell, const uint32_t* cell_off,
                      const uint32_t* cell_n, double max2, int lay, int half_symmetric, int skip_ghosts, uint32_t* nbr,
                      uint32_t* nbr_cnt, uint32_t cap, uint32_t* max_cnt, hipStream_t st) {
  k_nbh_build<<<nblk(n_atoms, 256), 256, 0, st>>>(g, n_atoms, r, slot_cell, cell_off, cell_n, max2, lay, half_symmetric, skip_ghosts,
                                                  nbr, nbr_cnt, cap, max_cnt);
}


static constexpr int LJ_UNROLL = 8;

template <bool MIXED, bool ACCUM, bool PUSH>
static void lj_launch(uint32_t n_atoms, const PState& A, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap, const LJConst& lj,
                      double dth, hipStream_t st) {
  k_compute_pair_slots<LennardJonesForceFunctor<MIXED>, ACCUM, PUSH, LJ_UNROLL><<<nblk(n_atoms, 256), 256, 0, st>>>(
      n_atoms, A.r, nbr, nbr_cnt, cap, LennardJonesForceFunctor<MIXED>{lj}, lj.rc2, A.fx, A.fy, A.fz, A.ep, A.vx, A.vy, A.vz, dth);
}

uint32_t lj_list_granule() { return LJ_UNROLL; }

void launch_lj(int precision, bool accum, bool push, uint32_t n_atoms, const PState& A, const uint32_t* nbr,
               const uint32_t* nbr_cnt, uint32_t cap, double eps, double sigma, double rcut, double dth, hipStream_t st) {
  if (n_atoms == 0) return;
  const LJConst lj = make_lj_const(eps, sigma, rcut);
  const int key = (precision ? 4 : 0) | (accum ? 2 : 0) | (push ? 1 : 0);
  switch (key) {
    case 0: lj_launch<false, false, false>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 1: lj_launch<false, false, true>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 2: lj_launch<false, true, false>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 3: lj_launch<false, true, true>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 4: lj_launch<true, false, false>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 5: lj_launch<true, false, true>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    case 6: lj_launch<true, true, false>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
    default: lj_launch<true, true, true>(n_atoms, A, nbr, nbr_cnt, cap, lj, dth, st); break;
  }
}

// gravitational_force on per-slot lists (f += pairs, in stream order, into the stored force)
void launch_gravity_slots(uint32_t n_atoms, const PState& A, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                          const GravitationalForceFunctor& gf, double rc2, hipStream_t st) {
  if (n_atoms == 0) return;
  k_compute_pair_slots<GravitationalForceFunctor, true, false, LJ_UNROLL><<<nblk(n_atoms, 256), 256, 0, st>>>(
      n_atoms, A.r, nbr, nbr_cnt, cap, gf, rc2, A.fx, A.fy, A.fz, A.ep, A.vx, A.vy, A.vz, 0.0);
}

// largest particle type (r.w) of slots [0, n) -> atomicMax into *out
__global__ void __launch_bounds__(256) k_max_type(uint32_t n, const double4* __restrict__ r, uint32_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t t = s < n ? (uint32_t)r[s].w : 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t = max(t, (uint32_t)__shfl_xor((int)t, o, 64));
  if ((threadIdx.x & 63) == 0 && t) atomicMax(out, t);
}
void launch_max_type(uint32_t n, const double4* r, uint32_t* out, hipStream_t st) {
  if (n) k_max_type<<<nblk(n, 256), 256, 0, st>>>(n, r, out);
}

int launch_lj_multimat(int precision, uint32_t n_atoms, const PState& A, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                       int ntypes, const LJMulti& m, hipStream_t st) {
  if (ntypes < 1 || m.npairs != ntypes * (ntypes + 1) / 2) return -1;
  const uint32_t nb = nblk(n_atoms, 256);
  if (precision == 0) k_lj_multimat<false><<<nb, 256, 0, st>>>(n_atoms, A.r, nbr, nbr_cnt, cap, m, A.fx, A.fy, A.fz, A.ep);
  else k_lj_multimat<true><<<nb, 256, 0, st>>>(n_atoms, A.r, nbr, nbr_cnt, cap, m, A.fx, A.fy, A.fz, A.ep);
  return 0;
}

void launch_pair_counts(uint32_t n_atoms, const double4* r, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap, double rc2,
                        unsigned long long* out, hipStream_t st) {
  k_pair_counts<<<nblk(n_atoms, 256), 256, 0, st>>>(n_atoms, r, nbr, nbr_cnt, cap, rc2, out);
}

void launch_stream_plen(uint32_t n_atoms, const uint32_t* slot_cell, const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap,
                        int random_access, uint32_t* plen, hipStream_t st) {
  k_stream_plen<<<nblk(n_atoms, 256), 256, 0, st>>>(n_atoms, slot_cell, nbr, nbr_cnt, cap, random_access, plen);
}
void launch_stream_cell_len(const Geom& g, uint32_t n_atoms, const uint32_t* cell_off, const uint32_t* cell_n, const uint32_t* plen,
                            int ntables, uint32_t* pstart, uint32_t* clen, hipStream_t st) {
  k_stream_cell_len<<<nblk(g.ncells, 128), 128, 0, st>>>(g, n_atoms, cell_off, cell_n, plen, ntables, pstart, clen);
}
void launch_stream_fill(const Geom& g, uint32_t n_atoms, const uint32_t* slot_cell, const uint32_t* cell_off, const uint32_t* cell_n,
                        const uint32_t* nbr, const uint32_t* nbr_cnt, uint32_t cap, const uint32_t* pstart, const uint32_t* plen,
                        const uint32_t* coff, int ntables, int random_access, uint16_t* out, hipStream_t st) {
  k_stream_fill<<<nblk(n_atoms, 256), 256, 0, st>>>(g, n_atoms, slot_cell, cell_off, cell_n, nbr, nbr_cnt, cap, pstart, plen, coff,
                                                    ntables, random_access, out);
}

void launch_energies(uint32_t N, const PState& A, double* part, uint32_t nparts, double* out, hipStream_t st) {
  k_energy_partials<<<nparts, RED_THREADS, 0, st>>>(N, A.vx, A.vy, A.vz, A.ep, part);
  k_energy_final<<<1, 64, 0, st>>>(nparts, part, out);
}

// ================================================================== batched boxes
// Every batched launch covers many boxes: work[blockIdx.x] names the box and the block
// within it (NONE: idle workgroup).  Each workgroup serves one box, so a box's frame is read
// with wave-uniform (scalar) loads, and per-item bodies are those of the one-box kernels.

__device__ inline bool batch_item(const uint32_t* __restrict__ work, uint32_t& box, uint32_t& idx) {
  const uint32_t w = work[blockIdx.x];
  if (w == NONE) return false;
  box = w >> kWorkBoxShift;
  idx = (w & kWorkBlockMask) * blockDim.x + threadIdx.x;
  return true;
}

struct FrameState {
  PState A, B;
  uint32_t *scA, *scB;
};

__device__ inline FrameState frame_state(const BoxFrame& f, uint32_t par) {
  FrameState s;
  s.A = f.ab ? f.B : f.A;
  s.B = f.ab ? f.A : f.B;
  s.scA = f.slot_cell[f.ab & 1u];
  s.scB = f.slot_cell[(f.ab & 1u) ^ 1u];
  const uint32_t cur = (f.rcur ^ par) & 1u;
  s.A.r = f.rbuf[cur];
  s.B.r = f.rbuf[cur ^ 1u];
  return s;
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_zero_cells(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work) {
  uint32_t box, c;
  if (!batch_item(work, box, c)) return;
  const BoxFrame& f = F[box];
  if (c >= f.g.ncells) return;
  f.cnt_out[c] = 0;
  f.cnt_in[c] = 0;
  f.in_cur[c] = 0;
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_move_flag(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                 uint32_t par) {
  uint32_t box, s;
  if (!batch_item(work, box, s)) return;
  const BoxFrame& f = F[box];
  if (s >= f.N) return;
  const FrameState X = frame_state(f, par);
  move_flag_item(f.g, f.N, X.A.r, X.scA, f.mv_dst, f.cnt_out, f.cnt_in, f.words + 2, s);
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_move_pack(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work) {
  uint32_t box, c;
  if (!batch_item(work, box, c)) return;
  const BoxFrame& f = F[box];
  if (c >= f.g.ncells) return;
  move_pack_item(f.g, f.cell_off, f.cell_n, f.cnt_out, f.cnt_in, f.mv_dst, f.pk, f.newloc, f.new_n, c);
}

// One workgroup of 1024 threads per listed box: exclusive scan of the box's per-cell counts.
// which 0: new_n -> new_off; which 1: gcnt -> goff and the total into the frame's G.
__global__ void __launch_bounds__(1024) kb_scan_cells(BoxFrame* __restrict__ F, const uint32_t* __restrict__ boxes, int which) {
  __shared__ uint32_t wsum[16];
  const uint32_t box = boxes[blockIdx.x];
  BoxFrame& f = F[box];
  const uint32_t* in = which ? f.gcnt : f.new_n;
  uint32_t* out = which ? f.goff : f.new_off;
  const uint32_t n = f.g.ncells;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < 16; q++) {
      if (q < w) woff += wsum[q];
      tot += wsum[q];
    }
    __syncthreads();
    if (i < n) out[i] = carry + woff + x - v;
    carry += tot;
  }
  if (which && threadIdx.x == 0) {
    f.G = carry;
    f.words[3] = carry;
  }
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_move_scatter(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                    uint32_t par) {
  uint32_t box, s;
  if (!batch_item(work, box, s)) return;
  const BoxFrame& f = F[box];
  if (s >= f.N) return;
  const FrameState X = frame_state(f, par);
  move_scatter_item(f.N, X.A, X.B, X.scA, f.mv_dst, f.newloc, f.cell_off, f.new_off, f.cell_n, f.cnt_out, f.in_cur, f.in_list,
                    X.scB, s, false);  // inside a run: forces not carried
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_move_incoming(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                     uint32_t par) {
  uint32_t box, c;
  if (!batch_item(work, box, c)) return;
  const BoxFrame& f = F[box];
  if (c >= f.g.ncells) return;
  const FrameState X = frame_state(f, par);
  move_incoming_item(f.g, X.A, X.B, f.new_off, f.cell_n, f.cnt_out, f.cnt_in, f.in_list, X.scB, c, false);
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_set_inner_cells(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work) {
  uint32_t box, c;
  if (!batch_item(work, box, c)) return;
  const BoxFrame& f = F[box];
  if (c >= f.g.ncells) return;
  set_inner_cells_item(f.g.ncells, f.new_n, f.new_off, f.g, f.cell_n, f.cell_off, c);
}

// rebuild_amr of the rebuilding boxes (after kb_flip: X.A current, X.B spare): dense cells of
// every box listed (entry box << kWorkBoxShift | cell) over the cells work table, then sorted and
// copied back by one wave each
__global__ void __launch_bounds__(kBatchSlotBlock) kb_amr_mark(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
  uint32_t box, c;
  if (!batch_item(work, box, c)) return;
  const BoxFrame& f = F[box];
  if (c >= f.g.ncells) return;
  amr_mark_item(f.g, kSubGridDensity, f.cell_n, c, box << kWorkBoxShift, list, count, f.words + 0, 8u);
}

__global__ void __launch_bounds__(64) kb_amr_sort(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count, uint32_t par) {
  __shared__ uint32_t keys[kAmrMaxKeys];
  const uint32_t m = *count;
  for (uint32_t q = blockIdx.x; q < m; q += gridDim.x) {
    const uint32_t e = list[q];
    const BoxFrame& f = F[e >> kWorkBoxShift];
    const FrameState X = frame_state(f, par);
    amr_sort_wave(f.g, kSubGridDensity, X.A, X.B, f.cell_off, f.cell_n, e & kWorkBlockMask, keys, false);
  }
}

__global__ void __launch_bounds__(64) kb_amr_back(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ count, uint32_t par) {
  const uint32_t m = *count;
  for (uint32_t q = blockIdx.x; q < m; q += gridDim.x) {
    const uint32_t e = list[q];
    const BoxFrame& f = F[e >> kWorkBoxShift];
    const FrameState X = frame_state(f, par);
    amr_back_wave(X.A, X.B, f.cell_off, f.cell_n, e & kWorkBlockMask, false);
  }
}

// after move_particles: the B arrays become the current ones (the one-box swap_buffers)
__global__ void kb_flip(BoxFrame* __restrict__ F, const uint32_t* __restrict__ boxes, uint32_t nboxes) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nboxes) return;
  BoxFrame& f = F[boxes[q]];
  f.ab ^= 1u;
  f.rcur ^= 1u;
  f.G = 0;
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_backup_r(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                uint32_t par) {
  uint32_t box, s;
  if (!batch_item(work, box, s)) return;
  const BoxFrame& f = F[box];
  if (s >= f.N) return;
  const FrameState X = frame_state(f, par);
  backup_r_item(f.g, f.N, X.A.r, X.scA, f.backup, s);
}

// work: blocks of kBatchSlotBlock threads over ghost_cell_count * 64 threads per box
__global__ void __launch_bounds__(kBatchSlotBlock) kb_ghost_count(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                   uint32_t par) {
  uint32_t box, idx, c;
  if (!batch_item(work, box, idx)) return;
  const BoxFrame& f = F[box];
  if (!ghost_wave_cell(f.g, idx, c)) return;
  const FrameState X = frame_state(f, par);
  ghost_count_wave(f.g, X.A.r, f.cell_off, f.cell_n, f.gcnt, c);
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_ghost_fill(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                  uint32_t par) {
  uint32_t box, idx, c;
  if (!batch_item(work, box, idx)) return;
  const BoxFrame& f = F[box];
  if (!ghost_wave_cell(f.g, idx, c)) return;
  const FrameState X = frame_state(f, par);
  ghost_fill_wave(f.g, f.N, X.A, f.goff, f.cell_off, f.cell_n, X.scA, f.ghost_src, f.ghost_shift, f.cap, f.words + 0, c, false);
}

// clock: the batch's step state {parity, global step}, advanced by kb_advance after every step
__global__ void __launch_bounds__(kBatchSlotBlock) kb_ghost_update_r(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                      const uint32_t* __restrict__ clock) {
  uint32_t box, q;
  if (!batch_item(work, box, q)) return;
  const BoxFrame& f = F[box];
  if (q >= f.G) return;  // blocks cover the box's ghost capacity
  const uint32_t par = clock[0];
  ghost_update_r_item(f.g, f.N, f.G, f.rbuf[(f.rcur ^ par) & 1u], f.ghost_src, f.ghost_shift, q);
}

__global__ void __launch_bounds__(kBatchSlotBlock) kb_first_half(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ work,
                                                                  uint32_t par, uint32_t gstep) {
  uint32_t box, s;
  if (!batch_item(work, box, s)) return;  // whole workgroup
  const BoxFrame& f = F[box];
  const FrameState X = frame_state(f, par);
  bool over = false;
  if (s < f.N)
    over = verlet_first_half_item(f.g, X.A.r, X.A.vx, X.A.vy, X.A.vz, X.A.fx, X.A.fy, X.A.fz, f.dt, f.dt2, f.dth, X.scA, f.backup,
                                  f.thr2, s);
  block_count_add(over, f.counts + (gstep - f.gstep0));  // local step k = gstep - gstep0 + 1 -> [k - 1]
}

__global__ void kb_zero_words(uint32_t* __restrict__ words, const uint32_t* __restrict__ boxes, uint32_t n) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 4 * n) return;
  words[4 * (size_t)boxes[q >> 2] + (q & 3u)] = 0;
}

void launch_batch_zero_words(uint32_t* words, const uint32_t* boxes, uint32_t n, hipStream_t st) {
  if (n) kb_zero_words<<<(4 * n + 255) / 256, 256, 0, st>>>(words, boxes, n);
}

void launch_batch_zero_cells(const BoxFrame* F, const uint32_t* work, uint32_t nwork, hipStream_t st) {
  if (nwork) kb_zero_cells<<<nwork, kBatchSlotBlock, 0, st>>>(F, work);
}

void launch_batch_move(BoxFrame* F, const uint32_t* wslots, uint32_t nslots, const uint32_t* wcells, uint32_t ncells,
                       const uint32_t* boxes, uint32_t nboxes, uint32_t par, hipStream_t st) {
  if (!nboxes) return;
  launch_batch_zero_cells(F, wcells, ncells, st);
  if (nslots) kb_move_flag<<<nslots, kBatchSlotBlock, 0, st>>>(F, wslots, par);
  if (ncells) kb_move_pack<<<ncells, kBatchSlotBlock, 0, st>>>(F, wcells);
  kb_scan_cells<<<nboxes, 1024, 0, st>>>(F, boxes, 0);
  if (nslots) kb_move_scatter<<<nslots, kBatchSlotBlock, 0, st>>>(F, wslots, par);
  if (ncells) kb_move_incoming<<<ncells, kBatchSlotBlock, 0, st>>>(F, wcells, par);
  if (ncells) kb_set_inner_cells<<<ncells, kBatchSlotBlock, 0, st>>>(F, wcells);
  kb_flip<<<(nboxes + 255) / 256, 256, 0, st>>>(F, boxes, nboxes);
}

void launch_batch_ghosts(BoxFrame* F, const uint32_t* wslots, uint32_t nslots, const uint32_t* wcells, uint32_t ncells,
                         const uint32_t* wghost, uint32_t nghost, const uint32_t* boxes, uint32_t nboxes, uint32_t* amr_list,
                         uint32_t* amr_count, uint32_t par, hipStream_t st) {
  if (!nboxes) return;
  if (ncells) {
    kb_amr_mark<<<ncells, kBatchSlotBlock, 0, st>>>(F, wcells, amr_list, amr_count);
    kb_amr_sort<<<kAmrGrid, 64, 0, st>>>(F, amr_list, amr_count, par);
    kb_amr_back<<<kAmrGrid, 64, 0, st>>>(F, amr_list, amr_count, par);
  }
  if (nslots) kb_backup_r<<<nslots, kBatchSlotBlock, 0, st>>>(F, wslots, par);
  if (nghost) kb_ghost_count<<<nghost, kBatchSlotBlock, 0, st>>>(F, wghost, par);
  kb_scan_cells<<<nboxes, 1024, 0, st>>>(F, boxes, 1);
  if (nghost) kb_ghost_fill<<<nghost, kBatchSlotBlock, 0, st>>>(F, wghost, par);
}

void launch_batch_ghost_update_r(const BoxFrame* F, const uint32_t* work, uint32_t nwork, const uint32_t* clock, hipStream_t st) {
  if (nwork) kb_ghost_update_r<<<nwork, kBatchSlotBlock, 0, st>>>(F, work, clock);
}

// end of a batched step: parity flips, global step + 1 (one thread)
__global__ void kb_advance(uint32_t* __restrict__ clock) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    clock[2] = clock[0];  // the step just run, for force launches the host adds after it
    clock[3] = clock[1];
    clock[0] ^= 1u;
    clock[1] += 1u;
  }
}

void launch_batch_advance(uint32_t* clock, hipStream_t st) { kb_advance<<<1, 64, 0, st>>>(clock); }

// displacement trigger of each listed box for the step at gstep: its particle_displ_over
// count (local step k = gstep - gstep0 + 1 keeps it at counts[k - 1]) is positive
__global__ void kb_trigger(const BoxFrame* __restrict__ F, const uint32_t* __restrict__ boxes, uint32_t n, uint32_t gstep,
                           uint32_t* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t box = boxes[q];
  const BoxFrame& f = F[box];
  out[box] = f.counts[gstep - f.gstep0] > 0 ? 1u : 0u;
}

void launch_batch_trigger(const BoxFrame* F, const uint32_t* boxes, uint32_t n, uint32_t gstep, uint32_t* out, hipStream_t st) {
  if (n) kb_trigger<<<(n + 255) / 256, 256, 0, st>>>(F, boxes, n, gstep, out);
}

void launch_batch_first_half(const BoxFrame* F, const uint32_t* work, uint32_t nwork, uint32_t par, uint32_t gstep, hipStream_t st) {
  if (nwork) kb_first_half<<<nwork, kBatchSlotBlock, 0, st>>>(F, work, par, gstep);
}

}  // namespace xnb
