// xnb_generate.cpp -- seeded FCC lattice input generator (host only, no GPU call).
//
// Counter-based SplitMix64 stream: value(site, stream) = mix(key + 8*site + stream), with
// key = mix(seed ^ golden).  Uniform doubles use the top 53 bits.  Velocities are
// Box-Muller normals, made momentum free and rescaled to the target temperature with
// 3N-3 degrees of freedom (mass 1, k_B = 1).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/xnb_amd.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }       // [0,1)
inline double u01_open(uint64_t x) { return ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0); }  // (0,1]

}  // namespace

extern "C" int xnb_generate_fcc(int64_t ncell, double rho, int64_t n_atoms, double jitter, double temperature, uint64_t seed,
                                double* r, double* v, uint64_t* id, double* box_length) {
  if (ncell <= 0 || !(rho > 0.) || !(jitter >= 0.) || !(temperature >= 0.)) return XNB_E_ARG;
  const int64_t full = 4 * ncell * ncell * ncell;
  if (n_atoms <= 0) n_atoms = full;
  if (n_atoms > full) return XNB_E_ARG;
  // full lattice: a = (4/rho)^(1/3), L = ncell*a; partial fill keeps density rho: L = (n/rho)^(1/3)
  const double a_full = std::cbrt(4.0 / rho);
  const double L = (n_atoms == full) ? (double)ncell * a_full : std::cbrt((double)n_atoms / rho);
  const double cell_a = (n_atoms == full) ? a_full : L / (double)ncell;
  if (box_length) *box_length = L;
  static const double basis[4][3] = {{0., 0., 0.}, {0.5, 0.5, 0.}, {0.5, 0., 0.5}, {0., 0.5, 0.5}};
  const uint64_t key = mix64(seed ^ 0x9E3779B97F4A7C15ull);
  for (int64_t m = 0; m < n_atoms; m++) {
    const uint64_t site = (uint64_t)((n_atoms == full) ? m : (int64_t)(((unsigned __int128)m * (uint64_t)full) / (uint64_t)n_atoms));
    const int64_t b = (int64_t)(site & 3u);
    const int64_t cidx = (int64_t)(site >> 2);
    const int64_t i = cidx % ncell, j = (cidx / ncell) % ncell, k = cidx / (ncell * ncell);
    const int64_t ijk[3] = {i, j, k};
    for (int q = 0; q < 3; q++) {
      double x = ((double)ijk[q] + basis[b][q]) * cell_a;
      const double u = u01(mix64(key + 8 * site + (uint64_t)q));
      x += jitter * (2.0 * u - 1.0);
      if (x < 0.) x += L;
      if (x >= L) x -= L;
      if (r) r[3 * m + q] = x;
    }
    if (id) id[m] = (uint64_t)m;
  }
  if (v) {
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t m = 0; m < n_atoms; m++) {
      const uint64_t site = (uint64_t)((n_atoms == full) ? m : (int64_t)(((unsigned __int128)m * (uint64_t)full) / (uint64_t)n_atoms));
      const double u1 = u01_open(mix64(key + 8 * site + 3)), u2 = u01(mix64(key + 8 * site + 4));
      const double u3 = u01_open(mix64(key + 8 * site + 5)), u4 = u01(mix64(key + 8 * site + 6));
      const double r1 = std::sqrt(-2.0 * std::log(u1)), r2 = std::sqrt(-2.0 * std::log(u3));
      v[3 * m + 0] = r1 * std::cos(two_pi * u2);
      v[3 * m + 1] = r1 * std::sin(two_pi * u2);
      v[3 * m + 2] = r2 * std::cos(two_pi * u4);
    }
    if (n_atoms > 1) {
      double mean[3] = {0., 0., 0.};
      for (int64_t m = 0; m < n_atoms; m++)
        for (int q = 0; q < 3; q++) mean[q] += v[3 * m + q];
      for (int q = 0; q < 3; q++) mean[q] /= (double)n_atoms;
      double ke2 = 0.;
      for (int64_t m = 0; m < n_atoms; m++)
        for (int q = 0; q < 3; q++) {
          v[3 * m + q] -= mean[q];
          ke2 += v[3 * m + q] * v[3 * m + q];
        }
      const double dof = 3.0 * (double)n_atoms - 3.0;
      const double tcur = ke2 / dof;
      const double scale = tcur > 0. ? std::sqrt(temperature / tcur) : 0.;
      for (int64_t m = 0; m < 3 * n_atoms; m++) v[m] *= scale;
    } else {
      for (int q = 0; q < 3; q++) v[q] = 0.;
    }
  }
  return XNB_OK;
}
