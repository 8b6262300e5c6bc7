// xnb_device.h -- device-side geometry helpers shared by the gfx950 kernels.
//
// All FP64 expressions follow the evaluation order of the exaNBody sources they restate
// and the translation unit is compiled with -ffp-contract=off, so results are bitwise
// those of a non-contracted IEEE FP64 evaluation (sqrt and division correctly rounded).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xnb {

static constexpr uint32_t NONE = 0xFFFFFFFFu;

// Grid geometry of one box: domain cells D, ghost layers gl, grid dims d = D + 2 gl,
// grid offset -gl.  Cells are indexed k-major ((k*dj)+j)*di+i
// (src/core/include/exanb/core/grid_algorithm.h:84-91).
struct Geom {
  int di, dj, dk;  // grid dims including ghost layers
  int gl;
  int Di, Dj, Dk;  // domain dims
  int pi, pj, pk;  // periodic flags
  double ox, oy, oz;  // domain origin (bounds min)
  double ex, ey, ez;  // domain extent (bounds max)
  double cs;          // cell size
  double ghost_dist;  // grid.max_neighbor_distance()
  uint32_t ncells;
};

__host__ __device__ inline uint32_t cell_index(const Geom& g, int i, int j, int k) {
  return (uint32_t)((k * g.dj + j) * g.di + i);
}

__host__ __device__ inline void cell_ijk(const Geom& g, uint32_t c, int& i, int& j, int& k) {
  i = (int)(c % (uint32_t)g.di);
  uint32_t t = c / (uint32_t)g.di;
  j = (int)(t % (uint32_t)g.dj);
  k = (int)(t / (uint32_t)g.dj);
}

__host__ __device__ inline bool is_ghost(const Geom& g, int i, int j, int k) {
  return i < g.gl || j < g.gl || k < g.gl || i >= g.di - g.gl || j >= g.dj - g.gl || k >= g.dk - g.gl;
}

__host__ __device__ inline int pmod(int i, int n) { return ((i % n) + n) % n; }

// Ghost cells (the gl-thick shell of the grid) enumerated as: the two k slabs (all i, j),
// then the two j slabs of the inner k range (all i), then the two i slabs of the inner j and
// k ranges.
__host__ __device__ inline uint32_t ghost_cell_count(const Geom& g) {
  const uint32_t gl = (uint32_t)g.gl, di = (uint32_t)g.di, dj = (uint32_t)g.dj, dk = (uint32_t)g.dk;
  const uint32_t km = dk - 2 * gl, jm = dj - 2 * gl;
  return 2 * gl * di * dj + 2 * gl * di * km + 2 * gl * jm * km;
}
__host__ __device__ inline uint32_t ghost_cell_at(const Geom& g, uint32_t q) {
  const uint32_t gl = (uint32_t)g.gl, di = (uint32_t)g.di, dj = (uint32_t)g.dj, dk = (uint32_t)g.dk;
  const uint32_t km = dk - 2 * gl, jm = dj - 2 * gl;
  uint32_t i, j, k;
  const uint32_t A = gl * di * dj, B = gl * di * km, C = gl * jm * km;
  if (q < 2 * A) {
    const uint32_t hi = q >= A ? 1u : 0u, r = q - hi * A;
    i = r % di;
    j = (r / di) % dj;
    k = r / (di * dj) + hi * (dk - gl);
  } else if (q < 2 * A + 2 * B) {
    q -= 2 * A;
    const uint32_t hi = q >= B ? 1u : 0u, r = q - hi * B;
    i = r % di;
    j = (r / di) % gl + hi * (dj - gl);
    k = gl + r / (di * gl);
  } else {
    q -= 2 * A + 2 * B;
    const uint32_t hi = q >= C ? 1u : 0u, r = q - hi * C;
    i = r % gl + hi * (di - gl);
    j = gl + (r / gl) % jm;
    k = gl + r / (gl * jm);
  }
  return cell_index(g, (int)i, (int)j, (int)k);
}


// domain_periodic_location (src/core/include/exanb/core/domain.h:160-191).
// Returns the domain cell location and wraps x on periodic axes.
__host__ __device__ inline void periodic_location(const Geom& g, double& x, double& y, double& z,
                                                  int& li, int& lj, int& lk) {
  li = (int)floor((x - g.ox) / g.cs);
  lj = (int)floor((y - g.oy) / g.cs);
  lk = (int)floor((z - g.oz) / g.cs);
  if ((li < 0 || li >= g.Di) && g.pi) {
    int o = li;
    li = pmod(li, g.Di);
    x += (double)(li - o) * g.cs;
  }
  if ((lj < 0 || lj >= g.Dj) && g.pj) {
    int o = lj;
    lj = pmod(lj, g.Dj);
    y += (double)(lj - o) * g.cs;
  }
  if ((lk < 0 || lk >= g.Dk) && g.pk) {
    int o = lk;
    lk = pmod(lk, g.Dk);
    z += (double)(lk - o) * g.cs;
  }
}

// Grid::cell_position (src/core/include/exanb/core/grid.h:113-117) along one axis,
// for a grid location (ghost layers included): origin + (offset + loc) * cell_size.
__host__ __device__ inline double cell_origin(double o, int gl, int loc, double cs) {
  return o + (double)(loc - gl) * cs;
}

// GhostBoundaryModifier periodic shift (src/mpi/include/exanb/mpi/ghosts_comm_scheme.h:76-81)
__host__ __device__ inline double shift_coord(double x, int sh, double rmin, double rmax) {
  if (sh == 0) return x;
  return x + (sh > 0 ? 1.0 : -1.0) * (rmax - rmin);
}

// shift code: 2 bits per axis, value sh+1 in {0,1,2}
__host__ __device__ inline uint32_t pack_shift(int si, int sj, int sk) {
  return (uint32_t)((si + 1) | ((sj + 1) << 2) | ((sk + 1) << 4));
}

// restore_u32_double / encode_double_u32 (src/core/include/exanb/core/backup_r.h:31-51)
__host__ __device__ inline double restore_u32_double(uint32_t x, double o, double r) {
  return o + ((double)x * r) / 4294967296.0;
}

__host__ __device__ inline uint32_t encode_double_u32(double x, double o, double r) {
  double xo = x - o;
  long long v = (long long)((xo * 4294967296.0) / r);
  if (v < 0) v = 0;
  if (v > 4294967295LL) v = 4294967295LL;
  uint32_t a = (uint32_t)v;
  uint32_t b = a - 1u;
  uint32_t c = a + 1u;
  double ea = fabs(restore_u32_double(a, o, r) - x);
  double eb = fabs(restore_u32_double(b, o, r) - x);
  double ec = fabs(restore_u32_double(c, o, r) - x);
  if (eb < ea) return b;
  if (ec < ea) return c;
  return a;
}

// encode_cell_index (src/particle_neighbors/include/exanb/particle_neighbors/chunk_neighbors.h:137-150)
__host__ __device__ inline uint16_t encode_cell_rel(int ri, int rj, int rk) {
  return (uint16_t)((((rk + 16) << 10) | ((rj + 16) << 5) | (ri + 16)));
}

// Neighbour list layout: per wave of 64 consecutive atoms, cap rows of 64 u32 slot indices,
// so entry k of atom s sits at (s/64)*64*cap + k*64 + s%64 and lanes read coalesced rows.
__host__ __device__ inline size_t nbr_index(uint32_t s, uint32_t k, uint32_t cap) {
  return (size_t)(s >> 6) * ((size_t)cap << 6) + ((size_t)k << 6) + (s & 63u);
}

#ifdef __HIPCC__
// |r - backup|^2 >= thr2, backup decoded relative to the origin of the cell at grid
// coordinates (i, j, k) (particle_displ_over.cu:62-65)
__device__ inline bool displaced_ijk(const Geom& g, double4 p, int i, int j, int k, const uint32_t* backup, uint32_t s, double thr2) {
  const double dx = p.x - restore_u32_double(backup[3 * (size_t)s + 0], cell_origin(g.ox, g.gl, i, g.cs), g.cs);
  const double dy = p.y - restore_u32_double(backup[3 * (size_t)s + 1], cell_origin(g.oy, g.gl, j, g.cs), g.cs);
  const double dz = p.z - restore_u32_double(backup[3 * (size_t)s + 2], cell_origin(g.oz, g.gl, k, g.cs), g.cs);
  return ((dx * dx + dy * dy) + dz * dz) >= thr2;
}

__device__ inline bool displaced(const Geom& g, double4 p, uint32_t c, const uint32_t* backup, uint32_t s, double thr2) {
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  return displaced_ijk(g, p, i, j, k, backup, s, thr2);
}

// counter += number of threads of the workgroup with pred set.  Every thread of the workgroup
// must call it (it holds two barriers).  The waves add into LDS and one thread adds the
// workgroup's total to the device counter: device-scope atomics on one address serialise
// across the XCDs, and one per wave (131,072 per launch at 8.4M atoms) added 0.8 ms to a
// force launch once most waves held a displaced atom.
__device__ inline void block_count_add(bool pred, unsigned long long* counter) {
  __shared__ uint32_t wg_count;
  if (threadIdx.x == 0) wg_count = 0;
  __syncthreads();
  const unsigned long long m = __ballot(pred);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&wg_count, (uint32_t)__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0 && wg_count) atomicAdd(counter, (unsigned long long)wg_count);
}

// block_count_add without a barrier, for kernels whose waves should leave as soon as they
// finish (an end-of-kernel barrier kept every wave of an 11-wave workgroup resident until the
// slowest one finished, and the next workgroup could not start in their place: the fused FP64
// force launch went from 1.78 to 2.1 ms).  Each wave adds its count into LDS; the last wave to
// finish adds the workgroup's total to the device counter.  Every wave of the workgroup calls
// it exactly once.  w: two LDS words of the workgroup, [0] threads counted, [1] waves done,
// zeroed by one thread ahead of a barrier that precedes every wave's call.
__device__ inline void wave_group_count_add(bool pred, unsigned long long* counter, uint32_t* w) {
  const unsigned long long m = __ballot(pred);
  if ((threadIdx.x & 63) == 0) {
    if (m) __hip_atomic_fetch_add(&w[0], (uint32_t)__popcll(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t nw = (blockDim.x + 63u) >> 6;
    if (__hip_atomic_fetch_add(&w[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == nw - 1u) {
      const uint32_t tot = __hip_atomic_load(&w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (tot) atomicAdd(counter, (unsigned long long)tot);
    }
  }
}
#endif

}  // namespace xnb
