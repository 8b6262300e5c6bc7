"""The force-functor contract (include/xnb_compute_pair.h): the pair traversal runs any functor
with the reference's buffer-less form.  Two functors are instantiated: Lennard-Jones (the
production force, batched form) and the gravitational functor of contribs/pi
(gravitational_force.cu:44-82, buffer-less form, per-type masses).

CPU: the oracle's gravitational_force is pinned by closed forms (two bodies: |f| = G m_a m_b /
r^2 along the separation; a body at the centre of a cube of equal masses feels no force;
Newton's third law sums every force to zero; the cutoff drops a pair one ulp past rcut).
GPU: the device traversal equals the oracle bitwise on brick lists, per-slot lists and with
ghost cells; the LJ instantiation keeps its bitwise parity (the other parity tests)."""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64
from oracle.oracle import Oracle, OracleError
from tests.helpers import RCUT, SKIN, fcc_inputs, make_pair

G = 0.75
MASS = np.array([1.0, 2.5, 0.4])


def _oracle_of(r, L, types, ncell=3, rcut=2.5):
    o = Oracle((0.0,) * 3, (L,) * 3, (ncell,) * 3, L / ncell, ghost_dist=rcut + 0.3)
    o.set_particles(np.asarray(r, np.float64), None, np.arange(len(r), dtype=np.uint64), np.asarray(types, np.uint8))
    o.init(1.0, 1.0, rcut, 0.3)
    o.zero_force()
    return o


def _forces_by_id(o):
    e = o.export()
    f = np.zeros_like(e["f"])
    f[e["id"].astype(np.int64)] = e["f"]
    return f


def test_two_bodies_closed_form():
    L = 9.0
    r = [[3.1, 4.0, 4.2], [4.3, 4.5, 3.6]]
    o = _oracle_of(r, L, [0, 1])
    o.gravitational_force(G, MASS, 2.5)
    f = _forces_by_id(o)
    d = np.array(r[1]) - np.array(r[0])
    dist = np.linalg.norm(d)
    expect = G * MASS[0] * MASS[1] / dist ** 2 * d / dist
    np.testing.assert_allclose(f[0], expect, rtol=1e-14, atol=0)
    np.testing.assert_allclose(f[1], -expect, rtol=1e-14, atol=0)


def test_cube_of_equal_masses_balances():
    L = 9.0
    c = np.array([4.5, 4.5, 4.5])
    corners = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)], np.float64) * 0.7
    r = np.vstack([c, c + corners])
    o = _oracle_of(r, L, [2] + [1] * 8)
    o.gravitational_force(G, MASS, 2.5)
    f = _forces_by_id(o)
    mag = G * MASS[2] * MASS[1] / (3 * 0.49)
    assert np.abs(f[0]).max() <= 1e-14 * mag
    assert np.linalg.norm(f[1]) > 0.5 * mag


def test_newton_third_law_and_cutoff():
    r, _, ids, L = fcc_inputs(4, jitter=0.05, seed=5)
    types = (np.arange(len(r)) % 3).astype(np.uint8)
    o, _, _ = make_pair(r, None, ids, L, types=types, gpu=False)
    o.init(1.0, 1.0, RCUT, SKIN)
    o.zero_force()
    o.gravitational_force(G, MASS, RCUT)
    f = o.export()["f"]
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).sum()
    # a pair at exactly rcut is seen, one a few ulps beyond it is not
    a = np.array([2.0, 3.0, 4.0])
    for bx, seen in ((4.5, True), (np.nextafter(4.5, 5.0), False)):
        o2 = _oracle_of([a, [bx, 3.0, 4.0]], 9.0, [0, 0])
        o2.gravitational_force(G, MASS, 2.5)
        assert np.any(o2.export()["f"] != 0.0) == seen, bx


def test_type_without_mass_is_an_error():
    o = _oracle_of([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0]], 9.0, [0, 3])
    with pytest.raises(OracleError):
        o.gravitational_force(G, MASS, 2.5)


def _gpu_pair(seed, ncell=6):
    r, v, ids, L = fcc_inputs(ncell, jitter=0.1, temperature=1.0, seed=seed)
    types = np.random.default_rng(seed).integers(0, 3, len(r)).astype(np.uint8)
    o, b, dom = make_pair(r, v, ids, L, types=types)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    return o, b


@pytest.mark.gpu
@pytest.mark.parametrize("bricks,ghost", [(None, False), ("0", False), (None, True)])
def test_gravity_bitwise(monkeypatch, bricks, ghost):
    if bricks is None:
        monkeypatch.delenv("XNB_BRICKS", raising=False)
    else:
        monkeypatch.setenv("XNB_BRICKS", bricks)
    o, b = _gpu_pair(41)
    if ghost:
        o.chunk_neighbors(RCUT + SKIN)
        b.chunk_neighbors(RCUT + SKIN, all_cells=True)
    # accumulates into the stored LJ force, in stream order, as the buffer-less functor does
    o.gravitational_force(G, MASS, RCUT, ghost=ghost)
    b.gravitational_force(G, MASS, RCUT, ghost=ghost)
    oe, ge = o.export(ghosts=ghost), b.export(ghosts=ghost)
    np.testing.assert_array_equal(oe["id"], ge["id"])
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))
    np.testing.assert_array_equal(oe["ep"].view(np.uint64), ge["ep"].view(np.uint64))
    # alone: zero first, then gravity only
    o.zero_force(ghost)
    b.zero_particle_force(ghost)
    o.gravitational_force(G, MASS, RCUT, ghost=ghost)
    b.gravitational_force(G, MASS, RCUT, ghost=ghost)
    oe, ge = o.export(ghosts=ghost), b.export(ghosts=ghost)
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))
    assert np.abs(ge["f"]).max() > 0


@pytest.mark.gpu
def test_gravity_rejects_type_without_mass():
    from exanbody_amd import XnbError
    o, b = _gpu_pair(42)
    with pytest.raises(XnbError):
        b.gravitational_force(G, MASS[:2], RCUT)
