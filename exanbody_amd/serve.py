"""Continuous-batching request loop (xnb_serve_run, include/xnb_amd.h) and the synthetic MD
job mix it is measured on."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import Domain, RunParams, LJParams, XnbError
from .box import cubic_domain, generate_fcc


class Job(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("n", C.c_size_t), ("r", C.c_void_p), ("v", C.c_void_p), ("id", C.c_void_p),
                ("domain", Domain), ("run", RunParams), ("r_out", C.c_void_p), ("v_out", C.c_void_p), ("id_out", C.c_void_p)]


class JobResult(C.Structure):
    _fields_ = [("start_s", C.c_double), ("finish_s", C.c_double), ("latency_s", C.c_double), ("status", C.c_int),
                ("n_rebuilds", C.c_int64), ("n_violations", C.c_int64), ("peak_device_bytes", C.c_size_t)]


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int), ("steps_per_iter", C.c_int),
                ("autotune", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("atom_steps", C.c_double), ("atom_steps_per_s", C.c_double),
                ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double), ("max_latency_s", C.c_double),
                ("mean_active", C.c_double), ("completed", C.c_int), ("failed", C.c_int), ("peak_active", C.c_int),
                ("peak_bytes", C.c_size_t), ("peak_device_bytes", C.c_size_t)]


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_job(n_atoms, rho, temperature, steps, precision, seed, arrival=0.0, dt=0.005, rcut=2.5, skin=0.3, rebuild_every=10,
             jitter=0.01):
    """A periodic LJ box of n_atoms on a (partially filled) FCC lattice, as a job dict."""
    ncell = max(2, int(np.ceil((n_atoms / 4.0) ** (1.0 / 3.0))))
    r, v, ids, L = generate_fcc(ncell, rho, n_atoms=n_atoms, jitter=jitter, temperature=temperature, seed=seed)
    return dict(arrival=float(arrival), r=r, v=v, id=ids, domain=cubic_domain(L, rcut + skin), steps=int(steps),
                precision=int(precision), dt=dt, rcut=rcut, skin=skin, rebuild_every=int(rebuild_every), epsilon=1.0,
                sigma=1.0)


def job_mix(n_jobs, rate, atoms=(1000, 200000), rho=(0.70, 0.90), temperature=(0.5, 1.5), steps=(100, 2000), seed=1):
    """Poisson arrivals at `rate` jobs/s; atoms log-uniform, density / temperature / steps uniform,
    alternating FP64 and mixed precision."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs)) if rate > 0 else np.zeros(n_jobs)
    jobs = []
    for q in range(n_jobs):
        n = int(round(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1])))))
        jobs.append(make_job(n, rng.uniform(*rho), rng.uniform(*temperature), int(rng.integers(steps[0], steps[1] + 1)),
                             _lib.PRECISION_FP64 if q % 2 == 0 else _lib.PRECISION_MIXED, seed=1000 * seed + q,
                             arrival=float(t[q])))
    return jobs


def _job_struct(j, out=None):
    """The xnb_job of a job dict, and the arrays it points into (keep them alive)."""
    r = np.ascontiguousarray(j["r"], np.float64)
    v = np.ascontiguousarray(j["v"], np.float64)
    ids = np.ascontiguousarray(j["id"], np.uint64)
    d = Domain()
    dom = j["domain"]
    for a in range(3):
        d.bounds_min[a] = dom["bounds_min"][a]
        d.bounds_max[a] = dom["bounds_max"][a]
        d.grid_dims[a] = int(dom["grid_dims"][a])
        d.periodic[a] = int(dom["periodic"][a])
    d.cell_size = dom["cell_size"]
    rp = RunParams(j["dt"], LJParams(j["epsilon"], j["sigma"]), j["rcut"], j["skin"], j["precision"], j["rebuild_every"],
                   j["steps"], 0, 0)
    job = Job(j["arrival"], r.shape[0], _ptr(r), _ptr(v), _ptr(ids), d, rp, _ptr(out["r"]) if out else None,
              _ptr(out["v"]) if out else None, _ptr(out["id"]) if out else None)
    return job, [r, v, ids]


def job_bytes(j):
    """Device bytes the loop budgets for a job dict (xnb_job_bytes_for)."""
    job, keep = _job_struct(j)
    return int(_lib.load().xnb_job_bytes_for(C.byref(job)))


def serve(jobs, device=0, hbm_budget_bytes=0, max_active=0, steps_per_iter=1, outputs=False, autotune=False):
    """Run the loop over job dicts (arrival-ordered).  Returns (results, stats); with outputs,
    each result carries the final r, v, id in cell order."""
    L = _lib.load()
    n = len(jobs)
    arr = (Job * max(n, 1))()
    keep = []
    outs = []
    for q, j in enumerate(jobs):
        o = None
        if outputs:
            o = {k: np.zeros_like(np.ascontiguousarray(j[k], np.uint64 if k == "id" else np.float64)) for k in ("r", "v", "id")}
        outs.append(o)
        arr[q], held = _job_struct(j, o)
        keep += held
    res = (JobResult * max(n, 1))()
    st = ServeStats()
    cfg = ServeConfig(int(device), int(hbm_budget_bytes), int(max_active), int(steps_per_iter), int(bool(autotune)))
    rc = L.xnb_serve_run(arr, n, C.byref(cfg), res, C.byref(st))
    if rc != 0:
        raise XnbError(rc, "xnb_serve_run failed")
    results = []
    for q in range(n):
        x = res[q]
        d = dict(start=x.start_s, finish=x.finish_s, latency=x.latency_s, status=x.status, rebuilds=x.n_rebuilds,
                 violations=x.n_violations, peak_device_bytes=x.peak_device_bytes)
        if outputs:
            d.update(outs[q])
        results.append(d)
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    return results, stats
