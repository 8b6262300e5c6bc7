"""CPU restatement of the reference's deterministic initial state: the `lattice` operator
(FCC and the other unit cells are replicated in one fixed loop order, ids numbered per domain
cell) followed by `gaussian_noise_r` with `deterministic_noise: true` (every domain cell
seeds its own 64-bit Mersenne twister and draws normal deviates for rx, ry, rz of each of its
particles in cell order), and the unit constants of the reference's physics library.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker that turns the reference's own
sample deck (contribs/microStamp/samples/benchmark_lj_snap/input_lj_Ni.msp) into particle
arrays, so that its golden file check_values_lj_Ni.dat can be compared with a run.  The
product package (exanbody_amd) never imports it.

Third-party algorithms restated here (not in the reference tree): the C++ standard library's
std::mt19937_64 (the standard fixes its parameters and its 10000th output), and GCC
libstdc++'s std::generate_canonical<double, 53> over a 64-bit engine and
std::normal_distribution<double> (the Marsaglia polar method, returning the second deviate
of a pair first and keeping the first for the next call).  tests/golden/noise_fixture.json
holds deviates produced by the compiled libstdc++ (tests/golden/make_noise_fixture.py), and
the tests compare this module with them bitwise.
"""
import math

import numpy as np

_M64 = (1 << 64) - 1

# Unit constants of the reference's physics library (onika, not in the reference tree): the
# values under which the oracle reproduces check_values_lj_Ni.dat to 1e-10 (elementary charge
# of CODATA 1973, Dalton of CODATA 2014; CODATA 2018 values miss it by 8e-6 relative in eps).
ELEMENTARY_CHARGE_C = 1.6021892e-19
DALTON_KG = 1.660539040e-27
# internal units of the reference: length angstrom, mass Dalton, time picosecond
INTERNAL_ENERGY_J = DALTON_KG * 1.0e-20 / 1.0e-24


def ev_to_internal(e_ev):
    """Energy in eV -> internal units (Dalton angstrom^2 / ps^2)."""
    return e_ev * ELEMENTARY_CHARGE_C / INTERNAL_ENERGY_J


class MT19937_64:
    """std::mt19937_64 (n = 312, m = 156, the standard's tempering); seed() as the standard's
    seed(value)."""

    def __init__(self, seed=5489):
        self.seed(seed)

    def seed(self, s):
        mt = [0] * 312
        mt[0] = s & _M64
        for i in range(1, 312):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _M64
        self._mt, self._i = mt, 312

    def _twist(self):
        mt = self._mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self._i = 0

    def __call__(self):
        if self._i >= 312:
            self._twist()
        y = self._mt[self._i]
        self._i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _M64


def canonical(g):
    """generate_canonical<double, 53> over a 64-bit engine: one draw, (double)u / 2^64,
    values that round to 1 become the largest double below 1."""
    r = float(g()) / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


class NormalDistribution:
    """std::normal_distribution<double>(mean, stddev), polar method: x, y uniform in
    (-1, 1) until 0 < x^2 + y^2 <= 1, m = sqrt(-2 ln(r2) / r2); returns y m, keeps x m."""

    def __init__(self, mean, stddev):
        self.mean, self.stddev, self._saved = mean, stddev, None

    def __call__(self, g):
        if self._saved is not None:
            r, self._saved = self._saved, None
        else:
            while True:
                x = 2.0 * canonical(g) - 1.0
                y = 2.0 * canonical(g) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            self._saved = x * mult
            r = y * mult
        return r * self.stddev + self.mean


# unit cells in reduced coordinates (lattice_generator.h:162-169 for FCC)
UNIT_CELLS = {
    "FCC": [(0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)],
}


def lattice_with_noise(structure, lattice_size, bounds_max, grid_dims, cell_size, sigma):
    """Particles of the `lattice` operator on a periodic domain [0, bounds_max]^3 (identity
    xform, no shift, one rank, deterministic mode) followed by deterministic
    `gaussian_noise_r`.  Returns (positions (n, 3) in cell order, then generation order within
    a cell; ids).

    generate_particle_lattice.h:277-316: one thread walks k, j, i (each from floor(0 / a) - 1
    to ceil(L / a) + 1) and the unit cell's sites l in order; a site at (i + p_l) * a is kept
    when it lies inside the domain bounds (both ends included, geometry.h:184-203) and is
    appended to the cell floor(x / cell_size) (grid.h:126-134); sites on the domain's upper
    faces fall in ghost cells, which are neither numbered nor kept.
    generate_particle_lattice.h:157, 369-383: ids start after the largest existing id (none:
    1) and run through the domain's cells in index order i + D (j + D k), within a cell in
    insertion order.
    gaussian_noise.h:62-79, 152-159 and gaussian_noise.cpp:32: each inner cell seeds a
    std::mt19937_64 with cell_index * 1023 and a fresh normal_distribution(0, sigma) adds one
    deviate to rx, then ry, then rz of each particle in cell order."""
    a = float(lattice_size)
    L = float(bounds_max)
    D = int(grid_dims)
    cs = float(cell_size)
    basis = UNIT_CELLS[structure]
    lo, hi = math.floor(0.0 / a), math.ceil(L / a)
    cells = {}
    for k in range(lo - 1, hi + 2):
        for j in range(lo - 1, hi + 2):
            for i in range(lo - 1, hi + 2):
                for px, py, pz in basis:
                    x, y, z = (i + px) * a, (j + py) * a, (k + pz) * a
                    if not (0.0 <= x <= L and 0.0 <= y <= L and 0.0 <= z <= L):
                        continue
                    ci, cj, ck = math.floor(x / cs), math.floor(y / cs), math.floor(z / cs)
                    if ci >= D or cj >= D or ck >= D:
                        continue
                    cells.setdefault(ci + D * (cj + D * ck), []).append([x, y, z])
    r, ids = [], []
    next_id = 1
    for c in range(D ** 3):
        g = MT19937_64(c * 1023)
        nd = NormalDistribution(0.0, float(sigma))
        for p in cells.get(c, []):
            p[0] += nd(g)
            p[1] += nd(g)
            p[2] += nd(g)
            r.append(p)
            ids.append(next_id)
            next_id += 1
    return np.array(r, dtype=np.float64).reshape(-1, 3), np.array(ids, dtype=np.uint64)


def read_check_values(path):
    """check_values file (check_values.cpp:369-385, yaml_check_particles.h:24-31): per sampled
    particle id, position, acceleration and velocity, as hex floats in internal units.
    Returns {id: (r, a, v)}."""
    out = {}
    for line in open(path):
        s = line.strip()
        if not s.startswith("- ["):
            continue
        f = [t.strip() for t in s[3:-1].split(",")]
        vals = [float.fromhex(x) for x in f[1:]]
        out[int(f[0])] = (np.array(vals[0:3]), np.array(vals[3:6]), np.array(vals[6:9]))
    return out
