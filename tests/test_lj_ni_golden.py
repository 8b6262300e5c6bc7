"""The reference's own golden for the LJ path: contribs/microStamp/samples/benchmark_lj_snap/
check_values_lj_Ni.dat (copied to tests/golden/), written by check_values (check_values.cpp)
after the sample deck input_lj_Ni.msp: 16,384 Ni atoms on an FCC lattice (a = 3.48 A) in a
periodic 55.68 A box of 4^3 cells, deterministic gaussian noise of 0.05 A on the positions,
LJ eps = 0.3729 eV, sigma = 2.2808 A, rcut = 4.1 A, rcut_inc = 2 A, dt = 2 fs, 100 velocity
Verlet steps, forces divided by the mass (58.693 Da).  The golden holds position,
acceleration and velocity of 100 sampled atoms; check_values passes when each is within
1e-5 (A, A/ps^2, A/ps) of it.

The initial state comes from oracle/lattice_noise.py (the lattice order, the ids and the
per-cell seeded noise restated; its deviates are pinned against the compiled C++ standard
library).  Unit-mass particles with eps / m give the reference's accelerations
(divide_force_by_type_scalar)."""
import json
import os

import numpy as np
import pytest

from oracle.lattice_noise import (MT19937_64, NormalDistribution, ev_to_internal, lattice_with_noise,
                                  read_check_values)

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "check_values_lj_Ni.dat")
L, CELL, D = 55.68, 13.92, 4
SIGMA, RCUT, SKIN, DT, STEPS, MASS = 2.2808, 4.1, 2.0, 0.002, 100, 58.693
THRESH = 1e-5  # check_values' pos/vel/acc thresholds in input_lj_Ni.msp


@pytest.fixture(scope="module")
def ni_state():
    r, ids = lattice_with_noise("FCC", 3.48, L, D, CELL, 0.05)
    return r, ids


def eps_per_mass(e_ev=0.3729):
    return ev_to_internal(e_ev) / MASS


def errors(out, gold):
    idx = {int(i): q for q, i in enumerate(out["id"])}
    er, ea, ev = [], [], []
    for i, (r, a, v) in gold.items():
        q = idx[i]
        dr = r - out["r"][q]
        dr -= L * np.round(dr / L)  # check_values compares periodic images
        er.append(np.linalg.norm(dr))
        ea.append(np.linalg.norm(a - out["f"][q]))
        ev.append(np.linalg.norm(v - out["v"][q]))
    return max(er), max(ea), max(ev)


def test_mt19937_64_known_answer():
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042  # the C++ standard's required 10000th value
    fx = json.load(open(os.path.join(HERE, "golden", "noise_fixture.json")))
    assert fx["mt19937_64_10000th_default_seed"] == 9981545732273789042


def test_noise_deviates_equal_libstdcxx():
    fx = json.load(open(os.path.join(HERE, "golden", "noise_fixture.json")))
    for c, vals in fx["cells"].items():
        g = MT19937_64(int(c) * 1023)
        nd = NormalDistribution(0.0, fx["sigma"])
        got = [nd(g) for _ in vals]
        assert got == [float.fromhex(v) for v in vals], f"cell {c}"


def test_lattice_counts_and_ids(ni_state):
    r, ids = ni_state
    assert r.shape == (16384, 3)
    np.testing.assert_array_equal(ids, np.arange(1, 16385, dtype=np.uint64))
    # 256 atoms per cell, numbered cell by cell; noise of 0.05 A around the lattice sites
    cell = lambda p: int(np.floor(p[0] / CELL) + D * (np.floor(p[1] / CELL) + D * np.floor(p[2] / CELL)))
    site = np.round(r / 1.74) * 1.74
    assert np.abs(r - site).max() < 0.3 and 0.045 < np.std(r - site) < 0.055
    assert cell(site[0] + 1e-9) == 0 and cell(site[16383] + 1e-9) == 63


def _oracle_run(r, ids, steps=STEPS, e_ev=0.3729):
    from oracle.oracle import Oracle
    o = Oracle((0, 0, 0), (L, L, L), (D, D, D), CELL, ghost_dist=RCUT + SKIN)
    o.set_particles(r, None, ids)
    e = eps_per_mass(e_ev)
    o.init(e, SIGMA, RCUT, SKIN)
    res = o.run(steps, DT, e, SIGMA, RCUT, SKIN, rebuild_every=0)
    return o.export(), res


def test_oracle_reproduces_reference_golden(ni_state):
    """All 100 sampled atoms within check_values' thresholds (measured: 1e-10 or better)."""
    r, ids = ni_state
    gold = read_check_values(GOLDEN)
    assert len(gold) == 100
    out, res = _oracle_run(r, ids)
    er, ea, ev = errors(out, gold)
    print(f"oracle vs golden: |dr| {er:.3g} A, |da| {ea:.3g} A/ps^2, |dv| {ev:.3g} A/ps ({res})")
    assert er < THRESH and ea < THRESH and ev < THRESH
    assert er < 1e-9 and ea < 1e-8 and ev < 1e-9


def test_golden_pins_step_count_and_units(ni_state):
    """What the golden rules out: one step fewer or more, and the CODATA 2018 charge in the
    eV conversion (8e-6 relative in eps) each miss it by far more than the thresholds."""
    r, ids = ni_state
    gold = read_check_values(GOLDEN)
    for steps in (STEPS - 1, STEPS + 1):
        er, ea, ev = errors(_oracle_run(r, ids, steps=steps)[0], gold)
        assert er > 1e-3 and ea > 10.0, (steps, er, ea)
    e2018 = 0.3729 * 1.602176634e-19 / 1.6021892e-19  # the same eV in the CODATA 2018 charge
    er, ea, ev = errors(_oracle_run(r, ids, e_ev=e2018)[0], gold)
    assert ea > 1e-2 and ev > 1e-4, (er, ea, ev)


@pytest.mark.gpu
def test_gpu_fp64_reproduces_reference_golden(ni_state):
    """The library's FP64 run (generic per-slot lists: a 13.92 A cell holds 256 atoms, too
    many for an LDS brick) meets check_values' thresholds and equals the oracle bitwise."""
    from exanbody_amd import PRECISION_FP64, Box
    r, ids = ni_state
    gold = read_check_values(GOLDEN)
    b = Box((0, 0, 0), (L, L, L), (D, D, D), CELL)
    b.nbh_dist(RCUT, SKIN)
    b.set_particles(r, None, ids)
    e = eps_per_mass()
    b.init(e, SIGMA, RCUT, SKIN, PRECISION_FP64)
    res = b.run(STEPS, DT, e, SIGMA, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=0)
    out = b.export()
    b.close()
    er, ea, ev = errors(out, gold)
    print(f"GPU FP64 vs golden: |dr| {er:.3g} A, |da| {ea:.3g} A/ps^2, |dv| {ev:.3g} A/ps ({res})")
    assert er < THRESH and ea < THRESH and ev < THRESH
    ref, _ = _oracle_run(r, ids)
    o = {int(i): q for q, i in enumerate(ref["id"])}
    perm = np.array([o[int(i)] for i in out["id"]])
    for k in ("r", "v", "f"):
        np.testing.assert_array_equal(out[k].view(np.uint64), ref[k][perm].view(np.uint64), err_msg=f"{k} bitwise vs oracle")
