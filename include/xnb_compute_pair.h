/*
 * xnb_compute_pair.h -- force-functor contract of the pair traversal (C++ / HIP, gfx950).
 *
 * The traversal visits, for every central particle, its neighbours in the order of the
 * chunk_neighbors stream (compute_cell_particle_pairs_impl_default.h:148-221), accepts a pair
 * iff 0 < d2 <= rcut^2 with dr = r_b - r_a and d2 = (dr.x dr.x + dr.y dr.y) + dr.z dr.z, and
 * calls the functor once per accepted pair.  A functor is written against this header exactly
 * as against the reference's buffer-less form (lennard_jones.cu:105-124,
 * gravitational_force.cu:56-82):
 *
 *   __device__ void operator()(Vec3d dr, double d2, [int type_a,] double& fx, double& fy,
 *                              double& fz, [double& ep,] const NbhCells& cells,
 *                              size_t neighbor_cell, size_t neighbor_particle, double w) const;
 *
 * with the central-particle fields in field-set order (compute_cell_particle_pairs_impl_default.h:
 * 200-203): the particle type when ComputePairTraits<F>::CentralType, the per-atom energy after
 * the force when ComputePairTraits<F>::CentralEnergy.  cells[neighbor_cell][field::type]
 * [neighbor_particle] is the neighbour's type, as in the reference's cell arrays; the
 * (neighbor_cell, neighbor_particle) pair is a handle into `cells` (the neighbour's place in
 * the workgroup's staged region), not a grid cell index.  The interaction weight w is 1.
 *
 * Summation order, by the reference's traits (compute_pair_traits.h:28-59): a functor with
 * ComputeBufferCompatible sums its pairs from 0 and adds the sum to the stored force once (the
 * buffer form, lennard_jones.cu:67-102); a buffer-less-only functor adds every pair straight
 * into the stored force (gravitational_force.cu:56-82).
 *
 * An MI355X extension: a functor with ComputePairTraits<F>::Batched provides
 *
 *   template <int U, bool ENERGY> __device__ void batch(const double4& ra, const NbhBatch<U>& nb,
 *                                          double& fx, double& fy, double& fz, double& ep) const;
 *
 * which evaluates U list entries at once in list order (nb(u) reads neighbour u's position
 * from the staged region, when the functor chooses), makes the cutoff test itself and adds
 * +-0 for a rejected pair (the list padding is a far-away sentinel atom); ENERGY false means
 * the energy is not wanted.  The traversal then calls it once per batch of U entries instead
 * of the per-pair form, so a functor can interleave the long-latency parts of U pairs.
 *
 * The traversals are instantiated inside the library: compute_pair_brick_body (brick lists,
 * exanbody_amd/csrc/xnb_brick.hip) and k_compute_pair_slots (per-slot lists, all cells,
 * exanbody_amd/csrc/xnb_kernels.hip).  A new functor is written beside
 * LennardJonesForceFunctor and GravitationalForceFunctor (exanbody_amd/csrc/xnb_pair.h) and
 * gets an operator entry point in xnb_amd.h.
 */
#ifndef XNB_COMPUTE_PAIR_H
#define XNB_COMPUTE_PAIR_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace xnb {

struct Vec3d {
  double x, y, z;
};

/* Functor capabilities; the defaults are the reference's ComputePairTraits<void>
 * (compute_pair_traits.h:28-40), plus this traversal's central-field flags. */
template <class F>
struct ComputePairTraits {
  static constexpr bool ComputeBufferCompatible = true;
  static constexpr bool BufferLessCompatible = false;
  static constexpr bool CentralType = false;
  static constexpr bool CentralEnergy = false;
  static constexpr bool Batched = false;
};

namespace field {
struct type_t {};
constexpr type_t type{};
}  // namespace field

/* Neighbour-side fields: cells[cell][field::type][p] is the neighbour's type.  The brick
 * traversal reads it from the staged region (type8), the per-slot traversal from the
 * particle records (x, y, z, type) in slot order (rw). */
struct NbhCells {
  const uint8_t* type8;
  const double4* rw;
  struct TypeArray {
    const uint8_t* t8;
    const double4* rw;
    __device__ unsigned operator[](size_t p) const { return t8 ? (unsigned)t8[p] : (unsigned)rw[p].w; }
  };
  struct Cell {
    TypeArray ta;
    __device__ TypeArray operator[](field::type_t) const { return ta; }
  };
  __device__ Cell operator[](size_t) const { return Cell{TypeArray{type8, rw}}; }
};

/* U neighbours of one central particle, in list order, in the workgroup's staged region. */
template <int U>
struct NbhBatch {
  uint32_t j[U];       /* staged-region index of each neighbour */
  uint32_t off[U];     /* LDS byte address of neighbour u's x (24 j[u]: the staged positions
                          start at LDS address 0, the traversal kernels hold no static LDS) */
  const double* pos;   /* x, y, z of region atom i at pos[3 i], pos[3 i + 1], pos[3 i + 2] */
  const uint8_t* type; /* type of region atom i at type[i] */
  __device__ double4 operator()(int u) const {
    /* read through an LDS-space pointer made from the address, so that forming it is one
       24-bit multiply of the list entry's half-word */
    typedef const double __attribute__((address_space(3))) lds_f64;
    const lds_f64* q = (const lds_f64*)(uintptr_t)off[u];
    double4 p;
    p.x = q[0];
    p.y = q[1];
    p.z = q[2];
    p.w = 0.0;
    return p;
  }
  __device__ unsigned type_of(int u) const { return type[j[u]]; }
};

}  // namespace xnb

#endif
