"""Serving loop (xnb_serve_run): jobs arriving while others run are admitted into the shared
batch; each job's final state equals the same job advanced alone by xnb_run, bitwise, in the
mix and when served one at a time; latency statistics are consistent."""
import numpy as np
import pytest

from exanbody_amd import Box
from exanbody_amd.serve import job_mix, serve

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
    from exanbody_amd import _lib
    one = _lib.load().xnb_job_bytes(2000)
    results, stats = serve(jobs, hbm_budget_bytes=2 * one + 1)
    assert stats["completed"] == 6 and stats["peak_active"] == 2
    assert stats["peak_bytes"] <= 2 * one + 1
