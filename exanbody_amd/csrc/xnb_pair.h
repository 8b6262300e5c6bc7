// xnb_pair.h -- Lennard-Jones pair arithmetic shared by the force kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xnb {

// Plain FP64 form: the reference traversal and functor arithmetic, neighbours in stream
// order, accepted iff 0 < d2 <= rc^2 with d = r_b - r_a
// (compute_cell_particle_pairs_impl_default.h:183-186), then lj_compute_energy
// (lennard_jones.cu:46-56) and f += (de * (1/r)) * d (lennard_jones.cu:117-123).
// The per-atom sum starts from 0 and is added once to f, like the buffer functor
// (lennard_jones.cu:80-101).

struct LJConst {
  double sigma, eps4, m24eps, rc2;
  float sigma2f, eps4f, m24epsf;
};

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

// One pair of the plain FP64 form.  Pairs outside 0 < d2 <= rc^2 contribute through a
// select of 0.0: adding +-0 leaves the running sums bitwise unchanged (they start at +0.0 and
// round-to-nearest never produces -0 from a sum with a +0 operand), so the result equals the
// reference's branch.
__device__ inline void lj_pair_fp64(const LJConst& lj, const double4& ra, const double4& rb, double& tx, double& ty, double& tz,
                                    double& te) {
  const double dx = rb.x - ra.x;
  const double dy = rb.y - ra.y;
  const double dz = rb.z - ra.z;
  const double d2 = (dx * dx + dy * dy) + dz * dz;
  const bool in = d2 > 0.0 && d2 <= lj.rc2;
  const double rr = sqrt(in ? d2 : 1.0);
  const double inv_r = 1.0 / rr;
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
  te += in ? e : 0.0;
}

// One pair of the mixed form: FP64 displacement and cutoff test, FP32 coefficient
// de/r = -24 eps (2 s^12 - s^6) / r^2 from (float)d2 with a hardware reciprocal, FP64
// accumulation f = fma(c, d, f).
__device__ inline void lj_pair_mixed(const LJConst& lj, const double4& ra, const double4& rb, double& tx, double& ty, double& tz,
                                     double& te) {
  const double dx = rb.x - ra.x;
  const double dy = rb.y - ra.y;
  const double dz = rb.z - ra.z;
  const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const bool in = d2 > 0.0 && d2 <= lj.rc2;
  const float inv_r2 = __builtin_amdgcn_rcpf(in ? (float)d2 : 1.0f);
  const float s2 = lj.sigma2f * inv_r2;
  const float s6 = s2 * s2 * s2;
  const float s12 = s6 * s6;
  const float c = in ? lj.m24epsf * (2.0f * s12 - s6) * inv_r2 : 0.0f;
  const double cd = (double)c;
  tx = fma(cd, dx, tx);
  ty = fma(cd, dy, ty);
  tz = fma(cd, dz, tz);
  te += (double)(in ? lj.eps4f * (s12 - s6) : 0.0f);
}

inline LJConst make_lj_const(double eps, double sigma, double rcut) {
  LJConst c;
  c.sigma = sigma;
  c.eps4 = 4. * eps;
  c.m24eps = -24. * eps;
  c.rc2 = rcut * rcut;
  c.sigma2f = (float)(sigma * sigma);
  c.eps4f = (float)(4. * eps);
  c.m24epsf = (float)(-24. * eps);
  return c;
}

}  // namespace xnb
