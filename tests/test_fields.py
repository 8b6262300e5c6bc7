"""Field sets with extra per-particle fields (a grid variant carrying more than r, v, f, id,
type): the fields travel with their particle through move_particles, rebuild_amr, ghost
images, batched runs and snapshots, and leave the dynamics untouched.  Each field is a fixed
function of the particle id, so every slot, inner and ghost, can be checked exactly."""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED, Batch, Box, cubic_domain, generate_fcc
from tests.helpers import RCUT, SKIN

pytestmark = pytest.mark.gpu

FIELDS = (lambda i: i * 1.0, lambda i: 0.5 * i + 3.0, lambda i: np.sin(i))


def _box(ncell, rho, temp, seed, extra=True, types=False):
    r, v, ids, L = generate_fcc(ncell, rho, jitter=0.02, temperature=temp, seed=seed)
    dom = cubic_domain(L, RCUT + SKIN)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(RCUT, SKIN)
    x = [f(ids.astype(np.float64)) for f in FIELDS] if extra else None
    t = (ids % 2).astype(np.uint8) if types else None
    b.set_particles(r, v, ids, t, extra=x)
    return b


def _check_fields(e, what):
    ids = e["id"].astype(np.float64)
    assert len(e["extra"]) == len(FIELDS)
    for k, f in enumerate(FIELDS):
        np.testing.assert_array_equal(e["extra"][k], f(ids), err_msg=f"{what}: field {k}")


@pytest.mark.parametrize("rho,prec", [(0.8442, PRECISION_FP64), (1.6, PRECISION_MIXED)])
def test_fields_follow_particles_and_leave_dynamics_unchanged(rho, prec):
    steps, every = 30, 5
    b = _box(8, rho, 1.5, 21)
    ref = _box(8, rho, 1.5, 21, extra=False)
    for x in (b, ref):
        x.init(1.0, 1.0, RCUT, SKIN, prec)
        x.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=prec, rebuild_every=every)
    e = b.export(ghosts=True)
    assert b.num_particles()[1] > 0
    _check_fields(e, "inner and ghost slots after the run")
    er = ref.export(ghosts=True)
    assert "extra" not in er
    for k in ("cell", "id"):
        np.testing.assert_array_equal(e[k], er[k])
    for k in ("r", "v", "f", "ep"):
        np.testing.assert_array_equal(e[k].view(np.uint64), er[k].view(np.uint64), err_msg=k)


def test_fields_in_a_batch_and_snapshot():
    steps = 23
    alone = _box(6, 0.8, 1.2, 22)
    alone.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    alone.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
    b = _box(6, 0.8, 1.2, 22)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    batch = Batch()
    other = _box(5, 0.75, 1.0, 23, extra=False)
    other.init(1.0, 1.0, RCUT, SKIN, PRECISION_MIXED)
    s = batch.attach(b, steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
    batch.attach(other, 17, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_MIXED, rebuild_every=10)
    while batch.step(4):
        pass
    batch.detach(s)
    ea, eb = alone.export(ghosts=True), b.export(ghosts=True)
    _check_fields(eb, "batched box")
    for k in ("r", "v"):
        np.testing.assert_array_equal(ea[k].view(np.uint64), eb[k].view(np.uint64), err_msg=k)
    for k in range(len(FIELDS)):
        np.testing.assert_array_equal(ea["extra"][k], eb["extra"][k])
    # snapshot into a fresh context of the same domain keeps the fields, and the run goes on
    snap = b.snapshot()
    dims, _, _ = b.grid_info()
    d = b.export()
    c = _box(6, 0.8, 1.2, 22, extra=False)  # same domain; its particles are replaced
    c.restore(snap)
    ec = c.export()
    np.testing.assert_array_equal(ec["id"], d["id"])
    _check_fields(ec, "restored box")
    c.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    c.run(10, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
    _check_fields(c.export(ghosts=True), "restored box after 10 more steps")
    batch.close()


def test_grid_without_type_and_id():
    r, v, _, L = generate_fcc(4, 0.8, jitter=0.02, temperature=1.0, seed=24)
    dom = cubic_domain(L, RCUT + SKIN)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(RCUT, SKIN)
    b.set_particles(r, v)  # no id, no type field: ids 0..n-1, every type 0
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    e = b.export()
    np.testing.assert_array_equal(np.sort(e["id"]), np.arange(len(r), dtype=np.uint64))
    assert np.all(e["type"] == 0)
