"""Multi-material LJ (compute_pair_multimat) on the GPU vs the oracle: two species with
Lorentz-Berthelot mixing and per-pair cutoffs rc = 2.5 sigma_ij.  FP64 forces, energies and a
run with rebuilds are bitwise the oracle's (brick lists and generic lists); mixed precision
within the stated tolerance."""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED
from tests.helpers import assert_same_layout, fcc_inputs, make_pair, rel_err

pytestmark = pytest.mark.gpu

SIG = np.array([1.0, 1.2, 1.4])           # pair ids (0,0), (0,1), (1,1)
EPS = np.array([1.0, np.sqrt(0.8), 0.8])  # Lorentz-Berthelot
RC = 2.5 * SIG
SKIN = 0.3


def _pair(seed=31, ncell=6, frac_b=0.2):
    r, v, ids, L = fcc_inputs(ncell, rho=0.5, jitter=0.05, temperature=0.8, seed=seed)
    types = (np.random.default_rng(seed).random(r.shape[0]) < frac_b).astype(np.uint8)
    o, b, dom = make_pair(r, v, ids, L, types=types, min_cell=RC.max() + SKIN, rcut=RC.max(), skin=SKIN)
    o.init(1.0, 1.0, RC.max(), SKIN)
    b.init(1.0, 1.0, RC.max(), SKIN, PRECISION_FP64)
    o.zero_force()
    b.zero_particle_force()
    o.lj_force_multimat(2, EPS, SIG, RC)
    return o, b


@pytest.mark.parametrize("bricks", [None, "0"])
def test_multimat_forces_bitwise(monkeypatch, bricks):
    if bricks is not None:
        monkeypatch.setenv("XNB_BRICKS", bricks)
    else:
        monkeypatch.delenv("XNB_BRICKS", raising=False)
    o, b = _pair()
    b.lennard_jones_force_multimat(2, EPS, SIG, RC, precision=PRECISION_FP64)
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "multimat")
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))
    np.testing.assert_array_equal(oe["ep"].view(np.uint64), ge["ep"].view(np.uint64))
    assert np.abs(oe["f"]).max() > 0
    b.zero_particle_force()
    b.lennard_jones_force_multimat(2, EPS, SIG, RC, precision=PRECISION_MIXED)
    mx, rms = rel_err(b.export()["f"], oe["f"])
    assert rms <= 1e-5, (mx, rms)


def test_multimat_run_bitwise():
    """xnb_run_multimat (fused steps, rebuild every 10) equals the oracle's operator sequence."""
    o, b = _pair(seed=32)
    b.lennard_jones_force_multimat(2, EPS, SIG, RC, precision=PRECISION_FP64)
    dt, steps, every = 0.002, 25, 10
    nbh = RC.max() + SKIN
    for k in range(1, steps + 1):
        o.push_f_v_r(dt, 1.0)
        o.push_f_v(dt, 0.5)
        if k % every == 0:
            o.move_particles()
            o.backup_r()
            o.ghost_update_all()
            o.chunk_neighbors(nbh)
        else:
            o.ghost_update_r()
        o.zero_force()
        o.lj_force_multimat(2, EPS, SIG, RC)
        o.push_f_v(dt, 0.5)
    res = b.run_multimat(steps, dt, 2, EPS, SIG, RC, SKIN, precision=PRECISION_FP64, rebuild_every=every)
    assert res["rebuilds"] == steps // every
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after multimat run")
    for k in ("v", "f", "ep"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)


def test_multimat_run_pair_radii_then_full_lists():
    """xnb_run_multimat keeps lists of each type pair's radius; the operators that follow
    (stream export, the multimat force) see full lists again: the CS=1 streams of every inner
    cell equal the oracle's built on the same positions, and the forces stay bitwise."""
    o, b = _pair(seed=33)
    b.lennard_jones_force_multimat(2, EPS, SIG, RC, precision=PRECISION_FP64)
    dt, steps, every = 0.002, 22, 10
    nbh = RC.max() + SKIN
    for k in range(1, steps + 1):
        o.push_f_v_r(dt, 1.0)
        o.push_f_v(dt, 0.5)
        if k % every == 0:
            o.move_particles()
            o.rebuild_amr()
            o.backup_r()
            o.ghost_update_all()
            o.chunk_neighbors(nbh)
        else:
            o.ghost_update_r()
        o.zero_force()
        o.lj_force_multimat(2, EPS, SIG, RC)
        o.push_f_v(dt, 0.5)
    b.run_multimat(steps, dt, 2, EPS, SIG, RC, SKIN, precision=PRECISION_FP64, rebuild_every=every)
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after multimat run")
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))
    o.chunk_neighbors(nbh)  # full lists on the current positions, as the device rebuilds them
    gs, gmax = b.streams()
    os_ = o.streams()
    dims, gl, nc = b.grid_info()
    inner = 0
    for c in range(nc):
        k, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
            np.testing.assert_array_equal(gs[c], os_[c], err_msg=f"stream of cell {c}")
            inner += 1
    assert inner > 0
    o.zero_force()
    b.zero_particle_force()
    o.lj_force_multimat(2, EPS, SIG, RC)
    b.lennard_jones_force_multimat(2, EPS, SIG, RC, precision=PRECISION_FP64)
    np.testing.assert_array_equal(o.export()["f"].view(np.uint64), b.export()["f"].view(np.uint64))
