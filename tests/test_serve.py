"""Serving loop (xnb_serve_run): jobs arriving while others run are admitted into the shared
batch; each job's final state equals the same job advanced alone by xnb_run, bitwise, in the
mix and when served one at a time; latency statistics are consistent."""
import numpy as np
import pytest

from exanbody_amd import Box
from exanbody_amd.serve import job_bytes, job_mix, serve

pytestmark = pytest.mark.gpu


def _alone(j):
    d = j["domain"]
    b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"])
    b.nbh_dist(j["rcut"], j["skin"])
    b.set_particles(j["r"], j["v"], j["id"])
    b.init(j["epsilon"], j["sigma"], j["rcut"], j["skin"], j["precision"])
    res = b.run(j["steps"], j["dt"], j["epsilon"], j["sigma"], j["rcut"], j["skin"], precision=j["precision"],
                rebuild_every=j["rebuild_every"])
    out = b.export()
    b.close()
    return out, res


@pytest.mark.parametrize("max_active,steps_per_iter", [(0, 1), (1, 1), (0, 7)])
def test_serve_results_equal_alone(max_active, steps_per_iter):
    """Several steps per loop iteration (jobs finish inside an iteration) change nothing either."""
    jobs = job_mix(8, rate=200.0, atoms=(300, 3000), steps=(15, 45), seed=3)
    ref = [_alone(j) for j in jobs]
    results, stats = serve(jobs, max_active=max_active, steps_per_iter=steps_per_iter, outputs=True)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    if max_active == 1:
        assert stats["peak_active"] == 1
    lat = []
    for q, (res, (out, runres)) in enumerate(zip(results, ref)):
        assert res["status"] == 0
        assert (res["rebuilds"], res["violations"]) == (runres["rebuilds"], runres["violations"]), q
        np.testing.assert_array_equal(res["id"], out["id"], err_msg=f"job {q} ids")
        np.testing.assert_array_equal(res["r"].view(np.uint64), out["r"].view(np.uint64), err_msg=f"job {q} r")
        np.testing.assert_array_equal(res["v"].view(np.uint64), out["v"].view(np.uint64), err_msg=f"job {q} v")
        assert res["finish"] >= res["start"] >= jobs[q]["arrival"] - 1e-9
        lat.append(res["latency"])
    assert stats["p50_latency_s"] == pytest.approx(np.percentile(lat, 50))
    assert stats["p99_latency_s"] == pytest.approx(np.percentile(lat, 99))
    steps = sum(j["r"].shape[0] * j["steps"] for j in jobs)
    assert stats["atom_steps"] == pytest.approx(steps)


def test_serve_budget_limits_admission():
    jobs = job_mix(6, rate=0.0, atoms=(2000, 2000), steps=(10, 10), seed=4)
    need = [job_bytes(j) for j in jobs]
    budget = max(a + b for a, b in zip(need, need[1:])) + 1
    assert min(a + b + c for a, b, c in zip(need, need[1:], need[2:])) > budget
    results, stats = serve(jobs, hbm_budget_bytes=budget)
    assert stats["completed"] == 6 and stats["peak_active"] == 2
    assert stats["peak_bytes"] <= budget


def test_job_bytes_bounds_real_allocations():
    """The admission estimate xnb_job_bytes_for is an upper bound of the device memory each
    job's context really allocates in the loop (counted at every allocation and free), over
    sizes from 300 to 200k atoms, densities from 0.6 to 1.6 and list radii 2.8 and 4.4, FP64
    and mixed; the summed peak of the contexts stays within the budget the loop admitted
    against."""
    from exanbody_amd import _lib
    from exanbody_amd.serve import make_job
    L = _lib.load()
    jobs = []
    q = 0
    for n in (300, 4000, 30000, 200000):
        for rho, rcut, skin in ((0.6, 2.5, 0.3), (0.9, 2.5, 0.3), (1.2, 2.5, 0.3), (1.6, 2.5, 0.3), (0.8, 4.0, 0.4)):
            jobs.append(make_job(n, rho, 1.0, 12, _lib.PRECISION_FP64 if q % 2 == 0 else _lib.PRECISION_MIXED, seed=700 + q,
                                 rcut=rcut, skin=skin, dt=0.002))
            q += 1
    budget = sum(job_bytes(j) for j in jobs[-5:])
    results, stats = serve(jobs, hbm_budget_bytes=budget)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    for j, res in zip(jobs, results):
        n = j["r"].shape[0]
        est = job_bytes(j)
        print(f"n={n} rho={n / np.prod(np.array(j['domain']['bounds_max']) - j['domain']['bounds_min']):.2f} rl={j['rcut'] + j['skin']} "
              f"peak={res['peak_device_bytes']} estimate={est} ratio={res['peak_device_bytes'] / est:.3f}", flush=True)
        assert 0 < res["peak_device_bytes"] <= est
    assert stats["peak_bytes"] <= budget
    assert 0 < stats["peak_device_bytes"] <= stats["peak_bytes"]


def test_serve_failed_setup_is_reported():
    """A job whose set-up fails (grid that does not tile its domain) ends with that error while
    the jobs around it complete with the same results as alone."""
    jobs = job_mix(4, rate=500.0, atoms=(500, 1500), steps=(10, 20), seed=9)
    bad = dict(jobs[1])
    bad["domain"] = dict(bad["domain"], cell_size=bad["domain"]["cell_size"] * 1.01)
    jobs[1] = bad
    ref = {q: _alone(j) for q, j in enumerate(jobs) if q != 1}
    results, stats = serve(jobs, outputs=True)
    assert stats["completed"] == 3 and stats["failed"] == 1
    from exanbody_amd._lib import XNB_E_DOMAIN
    assert results[1]["status"] == XNB_E_DOMAIN
    for q, (out, _) in ref.items():
        assert results[q]["status"] == 0
        np.testing.assert_array_equal(results[q]["r"].view(np.uint64), out["r"].view(np.uint64), err_msg=f"job {q} r")
