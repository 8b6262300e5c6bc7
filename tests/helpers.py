"""Shared test helpers: an oracle and a device Box built from the same seeded inputs."""
import numpy as np

from exanbody_amd import Box, cubic_domain, generate_fcc
from oracle.oracle import Oracle

RHO = 0.8442
RCUT = 2.5
SKIN = 0.3


def fcc_inputs(ncell=8, rho=RHO, jitter=0.01, temperature=0.72, seed=1, n_atoms=0):
    r, v, ids, L = generate_fcc(ncell, rho, n_atoms=n_atoms, jitter=jitter, temperature=temperature, seed=seed)
    return r, v, ids, L


def fcc_lattice(cells, a, jitter=0.0, temperature=0.0, seed=1):
    """FCC lattice of cells[0] x cells[1] x cells[2] unit cells of edge a (numpy, any box
    shape), Gaussian position jitter and zero-mean Gaussian velocities of variance
    temperature.  Returns (r[n,3], v[n,3], id[n], box lengths)."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(k) for k in cells], indexing="ij"), -1).reshape(-1, 3)
    r = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)
    if jitter:
        r = r + rng.normal(0.0, jitter, r.shape)
    lengths = np.array(cells, np.float64) * a
    r = np.mod(r, lengths)
    v = rng.normal(0.0, np.sqrt(temperature), r.shape) if temperature else np.zeros_like(r)
    v -= v.mean(0)
    return r, v, np.arange(len(r), dtype=np.uint64), lengths


def make_pair_domain(r, v, ids, dom, types=None, rcut=RCUT, skin=SKIN, gpu=True):
    """Oracle and (optionally) device Box of domain dom over the same particles; nbh_dist
    applied to both."""
    periodic = dom["periodic"]
    o = Oracle(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"], periodic=periodic,
               ghost_dist=rcut + skin)
    o.set_particles(r, v, ids, types)
    b = None
    if gpu:
        b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"], periodic=periodic)
        b.nbh_dist(rcut, skin)
        b.set_particles(r, v, ids, types)
    return o, b


def make_pair(r, v, ids, L, types=None, min_cell=RCUT + SKIN, rcut=RCUT, skin=SKIN, gpu=True, periodic=(1, 1, 1)):
    """Oracle and (optionally) device Box over the same particles in a cubic box of edge L."""
    dom = cubic_domain(L, min_cell, periodic=periodic)
    o, b = make_pair_domain(r, v, ids, dom, types, rcut, skin, gpu)
    return o, b, dom


def assert_same_layout(o_exp, g_exp, what=""):
    """Same particles in the same cells and the same within-cell order (bitwise positions)."""
    assert o_exp["cell"].shape == g_exp["cell"].shape, f"{what}: particle counts differ"
    np.testing.assert_array_equal(o_exp["cell"], g_exp["cell"], err_msg=f"{what}: cell of each slot")
    np.testing.assert_array_equal(o_exp["id"], g_exp["id"], err_msg=f"{what}: ids in cell order")
    np.testing.assert_array_equal(o_exp["r"].view(np.uint64), g_exp["r"].view(np.uint64),
                                  err_msg=f"{what}: positions (bitwise)")


def rel_err(a, b):
    """max and RMS of |a-b| per atom relative to the RMS norm of b."""
    d = np.linalg.norm(a - b, axis=1)
    ref = np.sqrt(np.mean(np.sum(b * b, axis=1)))
    return d.max() / ref, np.sqrt(np.mean(d * d)) / ref
