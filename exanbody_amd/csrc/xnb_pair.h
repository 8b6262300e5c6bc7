// xnb_pair.h -- Lennard-Jones pair arithmetic shared by the force kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xnb_compute_pair.h"

namespace xnb {

// Plain FP64 form: the reference traversal and functor arithmetic, neighbours in stream
// order, accepted iff 0 < d2 <= rc^2 with d = r_b - r_a
// (compute_cell_particle_pairs_impl_default.h:183-186), then lj_compute_energy
// (lennard_jones.cu:46-56) and f += (de * (1/r)) * d (lennard_jones.cu:117-123).
// The per-atom sum starts from 0 and is added once to f, like the buffer functor
// (lennard_jones.cu:80-101).

struct LJConst {
  double sigma, eps4, m24eps, rc2;
  float sigma6f, eps4f, m24epsf;  // sigma^6, 4 eps, -24 eps for the mixed form
};

// FP32 part of the mixed form from x = 1/r^2 (0 outside the cutoff): s6 = sigma^6 x^3, the
// coefficient de/r = -24 eps (2 s6 - 1) s6 x and the energy 4 eps (s6 - 1) s6, with fused
// multiply-adds (the mixed form has a tolerance, not the reference's rounding).
__device__ inline float lj_coef_f32(float x, float sigma6, float m24eps, float& s6) {
  s6 = sigma6 * (x * x * x);
  return (s6 * x) * fmaf(s6, 2.0f * m24eps, -m24eps);
}
__device__ inline float lj_energy_f32(float s6, float eps4) { return eps4 * fmaf(s6, s6, -s6); }

template <bool ACCUM>
__device__ inline void store_force(uint32_t s, double tx, double ty, double tz, double te, double* __restrict__ fx,
                                   double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ ep) {
  if (ACCUM) {
    fx[s] += tx;
    fy[s] += ty;
    fz[s] += tz;
    ep[s] += te;
  } else {
    fx[s] = 0.0 + tx;
    fy[s] = 0.0 + ty;
    fz[s] = 0.0 + tz;
    ep[s] = 0.0 + te;
  }
}

// Wave-wide maximum (every lane of the wave must be active), returned as a wave-uniform (scalar) value so loops bounded by it stay
// scalar branches
__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Correctly rounded FP64 square root and reciprocal for normal-range arguments.
//
// sqrt_normal: Goldschmidt iteration from the hardware reciprocal square root estimate
// (v_rsq_f64), then two fma-corrected Newton steps on the root (P. Markstein, "IA-64 and
// Elementary Functions", 2000, ch. 9).  recip_normal: two Newton-Raphson steps on the
// hardware reciprocal estimate (v_rcp_f64) and one fma-corrected quotient step.  For
// sqrt_normal(x), x in [2^-767, 2^1000], and recip_normal(y), y in [2^-383, 2^383], these
// are the sequences hipcc emits for sqrt(x) and 1.0 / y without the range scaling and the
// special-value fix-ups, which are identities there; the results are the IEEE correctly
// rounded ones (checked against the host's IEEE sqrt and division by the GPU tests).
__device__ inline double sqrt_normal(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g;
  double h = g * 0.5;
  const double r = fma(-h, s, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  double d = fma(-s, s, x);
  s = fma(d, h, s);
  d = fma(-s, s, x);
  return fma(d, h, s);
}

__device__ inline double recip_normal(double y) {
  double a = __builtin_amdgcn_rcp(y);
  double e = fma(-y, a, 1.0);
  a = fma(a, e, a);
  e = fma(-y, a, 1.0);
  a = fma(a, e, a);
  const double rem = fma(-y, a, 1.0);
  return fma(rem, a, a);
}

// 1 / sqrt(x) as two correctly rounded steps, r = sqrt(x) then 1.0 / r; the normal-range
// sequences unless some lane of the wave has x below their range (never for physical pair
// distances), in which case the whole wave takes the library operations.
__device__ inline double inv_sqrt_rounded(double x) {
  if (__builtin_expect(__any(x < 0x1p-767), 0)) return 1.0 / sqrt(x);
  return recip_normal(sqrt_normal(x));
}

// One pair of the plain FP64 form.  Pairs outside 0 < d2 <= rc^2 contribute through a
// select of 0.0: adding +-0 leaves the running sums bitwise unchanged (they start at +0.0 and
// round-to-nearest never produces -0 from a sum with a +0 operand), so the result equals the
// reference's branch.
template <bool ENERGY = true>
__device__ inline void lj_pair_fp64(const LJConst& lj, const double4& ra, const double4& rb, double& tx, double& ty, double& tz,
                                    double& te) {
  const double dx = rb.x - ra.x;
  const double dy = rb.y - ra.y;
  const double dz = rb.z - ra.z;
  const double d2 = (dx * dx + dy * dy) + dz * dz;
  const bool in = d2 > 0.0 && d2 <= lj.rc2;
  const double inv_r = inv_sqrt_rounded(in ? d2 : 1.0);  // 1.0 / sqrt(d2), each correctly rounded
  const double ratio = lj.sigma * inv_r;
  const double ratio2 = ratio * ratio;
  const double ratio6 = ratio2 * ratio2 * ratio2;
  const double ratio12 = ratio6 * ratio6;
  const double e = lj.eps4 * (ratio12 - ratio6);
  double de = (lj.m24eps * (2. * ratio12 - ratio6)) * inv_r;
  de *= inv_r;  // interaction_weight / r with weight 1.0: the same correctly rounded 1.0/r
  de = in ? de : 0.0;
  tx += de * dx;
  ty += de * dy;
  tz += de * dz;
  if (ENERGY) te += in ? e : 0.0;
}

// One pair of the mixed form: FP64 displacement and cutoff test, FP32 coefficient
// de/r = -24 eps (2 s^12 - s^6) / r^2 (lj_coef_f32) from (float)d2 with a hardware
// reciprocal, FP64 accumulation f = fma(c, d, f).
template <bool ENERGY = true>
__device__ inline void lj_pair_mixed(const LJConst& lj, const double4& ra, const double4& rb, double& tx, double& ty, double& tz,
                                     double& te) {
  const double dx = rb.x - ra.x;
  const double dy = rb.y - ra.y;
  const double dz = rb.z - ra.z;
  const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const bool in = d2 > 0.0 && d2 <= lj.rc2;
  // outside the cutoff 1/r^2 = 0 makes the coefficient and the energy +-0
  const float inv_r2 = in ? __builtin_amdgcn_rcpf((float)d2) : 0.0f;
  float s6;
  const double cd = (double)lj_coef_f32(inv_r2, lj.sigma6f, lj.m24epsf, s6);
  tx = fma(cd, dx, tx);
  ty = fma(cd, dy, ty);
  tz = fma(cd, dz, tz);
  if (ENERGY) te += (double)lj_energy_f32(s6, lj.eps4f);
}

// Multi-material LJ parameters per unique pair id (particle_type_pair.h: id of types a <= b
// is b (b + 1) / 2 + a); up to 4 types.
static constexpr int MM_PAIRS = 10;
struct LJMulti {
  int npairs;
  double sigma[MM_PAIRS], eps4[MM_PAIRS], m24eps[MM_PAIRS], rc2[MM_PAIRS];
  float sigma6f[MM_PAIRS], eps4f[MM_PAIRS], m24epsf[MM_PAIRS];
};

__host__ __device__ inline unsigned unique_pair_id(unsigned a, unsigned b) {
  if (a > b) {
    unsigned t = a;
    a = b;
    b = t;
  }
  return (b * (b + 1)) / 2 + a;
}

inline int make_lj_multi(int ntypes, const double* eps, const double* sigma, const double* rcut, LJMulti& m) {
  const int np = ntypes * (ntypes + 1) / 2;
  if (ntypes < 1 || np > MM_PAIRS) return -1;
  m.npairs = np;
  for (int q = 0; q < np; q++) {
    m.sigma[q] = sigma[q];
    m.eps4[q] = 4. * eps[q];
    m.m24eps[q] = -24. * eps[q];
    m.rc2[q] = rcut[q] * rcut[q];
    m.sigma6f[q] = (float)(sigma[q] * sigma[q] * sigma[q] * sigma[q] * sigma[q] * sigma[q]);
    m.eps4f[q] = (float)(4. * eps[q]);
    m.m24epsf[q] = (float)(-24. * eps[q]);
  }
  return 0;
}

// U pairs of the plain FP64 form at once, summed in u order (the same sums as U calls of
// lj_pair_fp64): the distance tests first, one wave-uniform range decision for all U roots, then
// the U independent root / reciprocal chains back to back, which the compiler interleaves.
template <int U, bool ENERGY = true>
__device__ inline void lj_batch_fp64(const LJConst& lj, const double4& ra, const double4 (&rb)[U], double& tx, double& ty,
                                     double& tz, double& te) {
  double dx[U], dy[U], dz[U], x[U];
  bool in[U];
  bool small = false;
#pragma unroll
  for (int u = 0; u < U; u++) {
    dx[u] = rb[u].x - ra.x;
    dy[u] = rb[u].y - ra.y;
    dz[u] = rb[u].z - ra.z;
    const double d2 = (dx[u] * dx[u] + dy[u] * dy[u]) + dz[u] * dz[u];
    in[u] = d2 > 0.0 && d2 <= lj.rc2;
    x[u] = in[u] ? d2 : 1.0;
    small = small || x[u] < 0x1p-767;
  }
  double inv_r[U];
  if (__builtin_expect(__any(small), 0)) {
#pragma unroll
    for (int u = 0; u < U; u++) inv_r[u] = 1.0 / sqrt(x[u]);
  } else {
#pragma unroll
    for (int u = 0; u < U; u++) inv_r[u] = recip_normal(sqrt_normal(x[u]));
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const double ratio = lj.sigma * inv_r[u];
    const double ratio2 = ratio * ratio;
    const double ratio6 = ratio2 * ratio2 * ratio2;
    const double ratio12 = ratio6 * ratio6;
    double de = (lj.m24eps * (2. * ratio12 - ratio6)) * inv_r[u];
    de *= inv_r[u];
    de = in[u] ? de : 0.0;
    tx += de * dx[u];
    ty += de * dy[u];
    tz += de * dz[u];
    if (ENERGY) {
      const double e = lj.eps4 * (ratio12 - ratio6);
      te += in[u] ? e : 0.0;
    }
  }
}

// ------------------------------------------------------------------ force functors
// LennardJonesForceFunctor (lennard_jones.cu:59-125) for the pair traversal: the buffer-less
// form (lennard_jones.cu:105-124, plus the per-atom energy field) and the batched form the
// brick traversal uses.  MIXED selects the mixed form (FP32 pair coefficient, FP64
// displacement, cutoff test and accumulation).
template <bool MIXED>
struct LennardJonesForceFunctor {
  LJConst p;

  __device__ inline void operator()(Vec3d dr, double d2, double& fx, double& fy, double& fz, double& ep, const NbhCells&, size_t,
                                    size_t, double w) const {
    if (MIXED) {
      float s6;
      const double c = (double)lj_coef_f32(__builtin_amdgcn_rcpf((float)d2), p.sigma6f, p.m24epsf, s6);
      fx = fma(c, dr.x, fx);
      fy = fma(c, dr.y, fy);
      fz = fma(c, dr.z, fz);
      ep += (double)lj_energy_f32(s6, p.eps4f);
    } else {
      // lj_compute_energy (lennard_jones.cu:46-56), then de *= w / r, f += de * dr
      const double r = sqrt(d2);
      const double inv_r = 1.0 / r;
      const double ratio = p.sigma * inv_r;
      const double ratio2 = ratio * ratio;
      const double ratio6 = ratio2 * ratio2 * ratio2;
      const double ratio12 = ratio6 * ratio6;
      const double e = p.eps4 * (ratio12 - ratio6);
      double de = (p.m24eps * (2. * ratio12 - ratio6)) * inv_r;
      de *= w / r;
      fx += de * dr.x;
      fy += de * dr.y;
      fz += de * dr.z;
      ep += e;
    }
  }

  // U list entries in list order (sentinel padding included); pairs outside 0 < d2 <= rc^2
  // add +-0.  FP64: pairs by two, each pair's positions read just before its arithmetic and
  // the scheduler kept from hoisting the next pair's reads (with the force kernel's register
  // cap: 8.4M-atom FP64 force 1.91 -> 1.78 ms against eight pairs read up front).
  template <int U, bool ENERGY>
  __device__ __forceinline__ void batch(const double4& ra, const NbhBatch<U>& nb, double& fx, double& fy, double& fz,
                                        double& ep) const {
    if (MIXED) {
      double4 rb[U];
#pragma unroll
      for (int u = 0; u < U; u++) rb[u] = nb(u);
#pragma unroll
      for (int u = 0; u < U; u++) lj_pair_mixed<ENERGY>(p, ra, rb[u], fx, fy, fz, ep);
    } else {
#pragma unroll
      for (int h = 0; h < U; h += 2) {
        double4 rb[2];
#pragma unroll
        for (int u = 0; u < 2; u++) rb[u] = nb(h + u);
        lj_batch_fp64<2, ENERGY>(p, ra, rb, fx, fy, fz, ep);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
};

template <bool MIXED>
struct ComputePairTraits<LennardJonesForceFunctor<MIXED>> {
  static constexpr bool ComputeBufferCompatible = true;  // lennard_jones.cu:132-138
  static constexpr bool BufferLessCompatible = true;
  static constexpr bool CentralType = false;
  static constexpr bool CentralEnergy = true;
  static constexpr bool Batched = true;
};

// GravitationalForceFunctor (contribs/pi/gravitational_force.cu:44-82): Newtonian attraction
// between point masses given per particle type (the particle_type_properties "mass" scalar),
// e = -G m_a m_b / r, de = G m_a m_b / r^2, then de *= w / r and f += de * dr.  Types are at
// most kGravityTypes; the operator checks every particle's type against the table first.
static constexpr int kGravityTypes = 16;
struct GravitationalForceFunctor {
  double G;
  double mass[kGravityTypes];

  __device__ inline void operator()(Vec3d dr, double d2, int type_a, double& fx, double& fy, double& fz, const NbhCells& cells,
                                    size_t neighbor_cell, size_t neighbor_particle, double w) const {
    const double mass_a = mass[type_a];
    const int type_b = cells[neighbor_cell][field::type][neighbor_particle];
    const double mass_b = mass[type_b];
    const double r = sqrt(d2);
    const double inv_r = 1.0 / r;
    double de = G * mass_a * mass_b * inv_r * inv_r;  // gravitational_compute_energy (e is not used)
    de *= w / r;
    fx += de * dr.x;
    fy += de * dr.y;
    fz += de * dr.z;
  }
};

template <>
struct ComputePairTraits<GravitationalForceFunctor> {
  static constexpr bool ComputeBufferCompatible = false;  // gravitational_force.cu:136-143
  static constexpr bool BufferLessCompatible = true;
  static constexpr bool CentralType = true;
  static constexpr bool CentralEnergy = false;
  static constexpr bool Batched = false;
};

inline LJConst make_lj_const(double eps, double sigma, double rcut) {
  LJConst c;
  c.sigma = sigma;
  c.eps4 = 4. * eps;
  c.m24eps = -24. * eps;
  c.rc2 = rcut * rcut;
  c.sigma6f = (float)(sigma * sigma * sigma * sigma * sigma * sigma);
  c.eps4f = (float)(4. * eps);
  c.m24epsf = (float)(-24. * eps);
  return c;
}

}  // namespace xnb
