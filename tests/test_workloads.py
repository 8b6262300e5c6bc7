"""The configured workloads at their own sizes, against the oracle or against each box run
alone:

* 8,388,608-atom FCC fluid (FCC 128^3): FP64 through 20 steps with rebuilds at steps 10 and
  20, bitwise the oracle's (layout, velocities, per-atom forces and energies, the CS=1 streams
  of sampled inner cells); mixed forces on the final positions within 1e-5 RMS of the
  oracle's FP64 forces.
* The two-species slab with droplets (multimat, Lorentz-Berthelot, rc = 2.5 sigma_ij) at a
  reduced size with the same construction: FP64 run with rebuilds bitwise the oracle's.
* The 512-box batched ensemble's job list: all 512 boxes batched, 16 of them (the smallest
  and the largest among them) bitwise equal to the same box run alone.
* Serving with a 200,000-atom job and an HBM budget that binds: every job's result bitwise
  equal to the job run alone, the budget never exceeded.
"""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED, Batch, Box
from tests.helpers import RCUT, SKIN, assert_same_layout, fcc_inputs, make_pair, make_pair_domain, rel_err

pytestmark = pytest.mark.gpu

DT = 0.005


def _inner_cells(dims, gl):
    i = np.arange(dims[0])
    j = np.arange(dims[1])
    k = np.arange(dims[2])
    kk, jj, ii = np.meshgrid(k, j, i, indexing="ij")
    inner = ((ii >= gl) & (ii < dims[0] - gl) & (jj >= gl) & (jj < dims[1] - gl) & (kk >= gl) & (kk < dims[2] - gl))
    return np.nonzero(inner.ravel())[0]


def test_fcc128_fp64_after_rebuilds_bitwise():
    r, v, ids, L = fcc_inputs(128, temperature=0.72, jitter=0.01, seed=2)
    assert r.shape[0] == 8388608
    o, b, dom = make_pair(r, v, ids, L)
    assert dom["grid_dims"] == (76, 76, 76)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    ro = o.run(20, DT, 1.0, 1.0, RCUT, SKIN, rebuild_every=10)
    rg = b.run(20, DT, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=10)
    assert ro == rg and rg["rebuilds"] == 2
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after 20 steps")
    for k in ("v", "f", "ep"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    # CS=1 streams of 400 inner cells drawn at random, built at step 20 from the moved atoms
    dims, gl, nc = b.grid_info()
    cells = np.random.default_rng(5).choice(_inner_cells(dims, gl), 400, replace=False)
    gs, gmax = b.streams(cells)
    assert gmax == o.max_neighbors()
    for c in cells:
        np.testing.assert_array_equal(gs[c], o.stream(int(c)), err_msg=f"stream of cell {c}")
    # mixed forces on the final positions against the oracle's FP64 forces
    b.zero_particle_force()
    b.lennard_jones_force(1.0, 1.0, RCUT, precision=PRECISION_MIXED)
    mx, rms = rel_err(b.export()["f"], oe["f"])
    assert rms <= 1e-5, (mx, rms)


def test_two_species_slab_fp64_run_bitwise():
    from exanbody_amd.workloads import lorentz_berthelot, slab_droplets
    eps, sig, rc = lorentz_berthelot()
    r, v, ids, types, dom = slab_droplets(n_atoms=60000, ncell=32, n_droplets=8, seed=3)
    assert dom["grid_dims"] == (16, 16, 16)
    nbh = rc.max() + SKIN
    o, b = make_pair_domain(r, v, ids, dom, types=types, rcut=rc.max(), skin=SKIN)
    o.init(1.0, 1.0, rc.max(), SKIN)
    b.init(1.0, 1.0, rc.max(), SKIN, PRECISION_FP64)
    o.zero_force()
    o.lj_force_multimat(2, eps, sig, rc)
    b.zero_particle_force()
    b.lennard_jones_force_multimat(2, eps, sig, rc, precision=PRECISION_FP64)
    np.testing.assert_array_equal(o.export()["f"].view(np.uint64), b.export()["f"].view(np.uint64), "initial forces")
    dt, steps, every = 0.004, 30, 10
    for k in range(1, steps + 1):  # numerical_scheme with compute_pair_multimat forces
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
        o.lj_force_multimat(2, eps, sig, rc)
        o.push_f_v(dt, 0.5)
    res = b.run_multimat(steps, dt, 2, eps, sig, rc, SKIN, precision=PRECISION_FP64, rebuild_every=every)
    assert res["rebuilds"] == steps // every
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "slab after the run")
    for k in ("v", "f", "ep"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    # vacuum: a third of the cells or more hold no atom
    assert np.count_nonzero(b.cell_counts() == 0) > 0.3 * np.prod(dom["grid_dims"])
    b.zero_particle_force()
    b.lennard_jones_force_multimat(2, eps, sig, rc, precision=PRECISION_MIXED)
    mx, rms = rel_err(b.export()["f"], oe["f"])
    assert rms <= 1e-5, (mx, rms)


def _job_box(j):
    d = j["domain"]
    b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"])
    b.nbh_dist(j["rcut"], j["skin"])
    b.set_particles(j["r"], j["v"], j["id"])
    b.init(j["epsilon"], j["sigma"], j["rcut"], j["skin"], j["precision"])
    return b


def _same(a, b, what):
    for k in ("cell", "id"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    for k in ("r", "v", "f", "ep"):
        np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64), err_msg=f"{what}: {k} (bitwise)")


def test_batched_512_box_job_list_equals_alone():
    from exanbody_amd.workloads import batch_job_list
    steps = 20
    jobs = batch_job_list(512, steps)
    atoms = np.array([j["r"].shape[0] for j in jobs])
    order = np.argsort(atoms, kind="stable")
    check = sorted(set(order[np.linspace(0, len(order) - 1, 16).round().astype(int)].tolist()))
    assert len(check) == 16 and int(np.argmin(atoms)) in check and int(np.argmax(atoms)) in check
    batch = Batch()
    boxes = [_job_box(j) for j in jobs]
    slots = [batch.attach(b, steps, DT, 1.0, 1.0, RCUT, SKIN, precision=j["precision"], rebuild_every=j["rebuild_every"])
             for b, j in zip(boxes, jobs)]
    left = len(slots)
    while left:
        left = batch.step(steps)
    res = {}
    for q, s in enumerate(slots):
        assert batch.status(s) == (steps, 1)
        res[q] = batch.detach(s)
    for q in check:
        a = _job_box(jobs[q])
        ra = a.run(steps, DT, 1.0, 1.0, RCUT, SKIN, precision=jobs[q]["precision"], rebuild_every=jobs[q]["rebuild_every"])
        assert res[q] == ra, f"box {q}: rebuild / violation counts"
        _same(boxes[q].export(), a.export(), f"box {q} ({atoms[q]} atoms)")
        a.close()
    for b in boxes:
        b.close()
    batch.close()


def test_serving_large_job_binding_budget():
    from exanbody_amd import _lib
    from exanbody_amd.serve import job_mix, make_job, serve
    jobs = job_mix(7, rate=100.0, atoms=(1000, 20000), steps=(20, 40), seed=6)
    big = make_job(200000, 0.80, 1.0, 30, PRECISION_MIXED, seed=77, arrival=0.005)
    jobs = sorted(jobs + [big], key=lambda j: j["arrival"])
    L = _lib.load()
    need = [L.xnb_job_bytes(j["r"].shape[0]) for j in jobs]
    budget = max(need) + sorted(need)[len(need) // 2]  # the 200k job fits only beside the smaller half
    assert sum(need) > budget
    results, stats = serve(jobs, hbm_budget_bytes=budget, steps_per_iter=5, outputs=True)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    assert stats["peak_bytes"] <= budget
    assert stats["peak_active"] < len(jobs)
    for q, (j, res) in enumerate(zip(jobs, results)):
        b = _job_box(j)
        rr = b.run(j["steps"], j["dt"], j["epsilon"], j["sigma"], j["rcut"], j["skin"], precision=j["precision"],
                   rebuild_every=j["rebuild_every"])
        out = b.export()
        b.close()
        assert res["status"] == 0 and (res["rebuilds"], res["violations"]) == (rr["rebuilds"], rr["violations"]), q
        np.testing.assert_array_equal(res["id"], out["id"], err_msg=f"job {q} ids")
        np.testing.assert_array_equal(res["r"].view(np.uint64), out["r"].view(np.uint64), err_msg=f"job {q} r")
        np.testing.assert_array_equal(res["v"].view(np.uint64), out["v"].view(np.uint64), err_msg=f"job {q} v")
