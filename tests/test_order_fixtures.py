"""Within-cell particle order after move_particles and in ghost images, against hand-derived
fixtures (tests/golden/order_fixture.json).

move_particles packs a cell's stayers by swap-with-last and appends movers in (source cell
index, source position) order (move_particles_across_cells.h:105-153,186-214); ghost images
keep source order (update_ghosts_comm_scheme.cpp:449-470).  Neighbour streams index
particles by their position in a cell, so these orders are part of the bit-exact contract.
The oracle is checked on the CPU, the device library through its C ABI; two oracle builds
with a planted ordering mistake each must fail the same checks.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "order_fixture.json")))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


def _arrays(case):
    parts = case["particles"]
    r = np.array([p["r"] for p in parts], np.float64)
    v = np.array([p.get("v", [0.0, 0.0, 0.0]) for p in parts], np.float64)
    ids = np.array([p["id"] for p in parts], np.uint64)
    return r, v, ids


def _cells_of(exp):
    """cell index -> (ids, x) in slot order, from an export with ghosts."""
    out = {}
    for c in np.unique(exp["cell"]):
        m = exp["cell"] == c
        out[int(c)] = (exp["id"][m].tolist(), exp["r"][m, 0].tolist())
    return out


def _check(case, exp):
    cells = _cells_of(exp)
    for c, want in case["expected_cells"].items():
        ids, x = cells.get(int(c), ([], []))
        assert ids == want["id"], f"cell {c}: ids {ids}, expected {want['id']}"
        assert x == want["x"], f"cell {c}: x {x}, expected {want['x']}"


def _oracle(case, libpath=None):
    d = case["domain"]
    return Oracle(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"], periodic=d["periodic"],
                  ghost_dist=case["rcut"] + case["skin"], libpath=libpath)


def oracle_move(libpath=None):
    case = FIX["move"]
    r, v, ids = _arrays(case)
    o = _oracle(case, libpath)
    o.set_particles(r, v, ids)
    o.zero_force()
    o.push_f_v_r(1.0, 1.0)
    o.move_particles()
    return o.export(ghosts=False)


def oracle_ghosts(libpath=None):
    case = FIX["ghost"]
    r, _, ids = _arrays(case)
    o = _oracle(case, libpath)
    o.set_particles(r, None, ids)
    o.ghost_update_all()
    return o.export(ghosts=True)


def test_fixture_is_self_consistent():
    """Every particle is accounted for once in the move fixture's expected cells."""
    case = FIX["move"]
    got = sorted(i for c in case["expected_cells"].values() for i in c["id"])
    assert got == sorted(p["id"] for p in case["particles"])


def test_move_order_oracle():
    _check(FIX["move"], oracle_move())


def test_ghost_order_oracle():
    _check(FIX["ghost"], oracle_ghosts())


@pytest.mark.parametrize("mutant,which", [("stable_pack", "move"), ("reverse_ghosts", "ghost")])
def test_planted_order_mistakes_fail(mutant, which):
    """An oracle with stable compaction instead of swap-with-last, or with ghost images in
    reversed source order, fails the fixture; the other fixture still passes."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "mutants"], check=True)
    path = os.path.join(ORACLE_DIR, "_build", f"libxnb_oracle_{mutant}.so")
    run = {"move": oracle_move, "ghost": oracle_ghosts}
    with pytest.raises(AssertionError):
        _check(FIX[which], run[which](path))
    other = "ghost" if which == "move" else "move"
    _check(FIX[other], run[other](path))


@pytest.mark.gpu
def test_move_order_gpu():
    from exanbody_amd import Box
    case = FIX["move"]
    d = case["domain"]
    r, v, ids = _arrays(case)
    b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"], periodic=d["periodic"])
    b.nbh_dist(case["rcut"], case["skin"])
    b.set_particles(r, v, ids)
    b.zero_particle_force()
    b.push_f_v_r(1.0, 1.0)
    b.move_particles()
    _check(case, b.export(ghosts=False))
    b.close()


@pytest.mark.gpu
def test_ghost_order_gpu():
    from exanbody_amd import Box
    case = FIX["ghost"]
    d = case["domain"]
    r, _, ids = _arrays(case)
    b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"], periodic=d["periodic"])
    b.nbh_dist(case["rcut"], case["skin"])
    b.set_particles(r, None, ids)
    b.ghost_update_all()
    _check(case, b.export(ghosts=True))
    b.close()
