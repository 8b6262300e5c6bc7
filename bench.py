"""Benchmark: LJ atom-steps/s (rc=2.5 sigma), mixed vs FP64 on one MI355X.

Headline workload: FCC 128^3 unit cells (8,388,608 atoms), rho*=0.8442, T*=0.72 seeded
Maxwell-Boltzmann velocities, +-0.01 sigma jitter, rc=2.5, skin=0.3, dt=0.005, neighbour
rebuild every 10 steps, velocity-Verlet.  Both precisions run the same device-resident
step loop (push, trigger, ghost refresh or full rebuild, force + push).  value = mixed
atom-steps/s; the FP64 plain form is timed beside it.

Per precision: W warm-up steps, then a rebuild (xnb_init), so the K timed steps start on
fresh lists and rebuild at steps 10, 20, ... of the timed run; the timed run has no HIP
event timers.  The same K steps are then replayed from a snapshot with HIP-event timers on
every kernel class (per-kernel averages, the roofline's launch duration), and the energy
drift is fitted over a window of steps after them.

Beside it, in the same JSON line: many boxes in one batch against one box at a time
("batched"), the continuous-batching serving loop with p50/p99 job latency ("serving"), and
the two-species slab on the multimat path ("two_species_slab"), each with its force error
and energy drift.

Prints ONE JSON line on stdout (the last line); progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LJ atom-steps/s (rc=2.5σ) mixed vs FP64 on 1 GPU; batched p50/p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (published; needs packed FP32, which these kernels do not use)
B_ALG = 144.0             # algorithmic bytes per atom-step: r, v, f read and written once (FP64)

WORKLOADS = {
    "c2": dict(ncell=128, seed=2, label="FCC 128^3 LJ fluid, 8,388,608 atoms"),
    "c1": dict(ncell=32, seed=1, label="FCC 32^3 LJ fluid, 131,072 atoms"),
}
RHO, T0, JITTER, RCUT, SKIN, DT, EVERY = 0.8442, 0.72, 0.01, 2.5, 0.3, 0.005, 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_error(f_test, f_ref):
    """max / RMS of |f_test - f_ref| relative to the RMS reference force, and the largest
    per-atom relative error."""
    d = np.linalg.norm(f_test - f_ref, axis=1)
    fn = np.linalg.norm(f_ref, axis=1)
    rms_ref = float(np.sqrt(np.mean(fn * fn))) if fn.size else 0.0
    if rms_ref == 0.0:
        return dict(rms_rel=0.0, max_rel=0.0, max_rel_per_atom=0.0, rms_force=0.0, atoms=int(fn.size))
    return dict(rms_rel=float(np.sqrt(np.mean(d * d)) / rms_ref), max_rel=float(d.max() / rms_ref),
                max_rel_per_atom=float(np.max(d / np.maximum(fn, 1e-300))), rms_force=rms_ref, atoms=int(fn.size))


def forces_both(box, eps=1.0, sigma=1.0, rcut=RCUT, keep=None):
    """FP64 and mixed forces on the box's current positions; the box is left holding the
    forces of precision `keep` (FP64 when None)."""
    from exanbody_amd import PRECISION_FP64, PRECISION_MIXED
    out = {}
    for p in (PRECISION_MIXED, PRECISION_FP64) if keep != PRECISION_MIXED else (PRECISION_FP64, PRECISION_MIXED):
        box.zero_particle_force()
        box.lennard_jones_force(eps, sigma, rcut, precision=p)
        out[p] = box.export()["f"]
    return out[PRECISION_FP64], out[PRECISION_MIXED]


def drift_fit(steps, energies, n_atoms):
    """Least-squares slope of total energy against step, per atom and per 1000 steps."""
    s = np.asarray(steps, np.float64)
    e = np.asarray(energies, np.float64)
    slope = float(np.polyfit(s, e, 1)[0]) if len(s) >= 2 else 0.0
    return dict(per_atom_per_1000_steps=slope * 1000.0 / n_atoms,
                rel_per_1000_steps=slope * 1000.0 / abs(float(e.mean())) if e.size and e.mean() != 0 else 0.0,
                window_steps=[int(s[0]), int(s[-1])] if len(s) else None, points=int(len(s)),
                e_first_per_atom=float(e[0] / n_atoms) if len(e) else None,
                e_last_per_atom=float(e[-1] / n_atoms) if len(e) else None)


def drift_window(box, start_step, chunks, every, run, n_atoms):
    """Energy after each of `chunks` runs of `every` steps (each run ends on its rebuild step
    and stores the per-atom energies), fitted."""
    steps, es = [], []
    ke, pe = box.energies()
    steps.append(start_step)
    es.append(ke + pe)
    for q in range(chunks):
        run(every)
        ke, pe = box.energies()
        steps.append(start_step + (q + 1) * every)
        es.append(ke + pe)
    return drift_fit(steps, es, n_atoms)


def timers(box):
    from exanbody_amd._lib import TIMER_FORCE, TIMER_GHOST, TIMER_NBH, TIMER_PUSH, TIMER_REBIN
    kern = {}
    for name, tid in (("force", TIMER_FORCE), ("nbh_build", TIMER_NBH), ("push", TIMER_PUSH), ("ghost", TIMER_GHOST),
                      ("rebin", TIMER_REBIN)):
        ms, cnt = box.timer_read(tid)
        kern[name] = dict(total_ms=round(ms, 4), launches=int(cnt), avg_ms=round(ms / cnt, 5) if cnt else None)
    return kern


def missed_pairs(box, snap, counts, every, run, init, max_checks=6):
    """In-cutoff pairs absent from the lists at steps where an atom had moved past max_displ
    with no rebuild: the run is replayed from the snapshot to that step, the in-cutoff pairs of
    the stale lists are counted, the lists are rebuilt on the same positions and counted again
    (lists are exact at a rebuild, so the difference is what the stale lists missed).  The
    last `max_checks` such steps are checked."""
    viol = [k for k in range(1, len(counts) + 1) if counts[k - 1] > 0 and (every <= 0 or k % every != 0)]
    out = dict(violation_steps=viol, checked=[], missed_in_cutoff_pairs=[])
    for k in viol[-max_checks:]:
        box.restore(snap)
        run(k)
        _, stale = box.pair_counts()
        init()
        _, fresh = box.pair_counts()
        out["checked"].append(k)
        out["missed_in_cutoff_pairs"].append(int(fresh) - int(stale))
    out["total_missed"] = int(sum(out["missed_in_cutoff_pairs"]))
    return out


def run_precision(inputs, dom, precision, steps, warmup, every, want_error, drift_chunks):
    from exanbody_amd import Box, PRECISION_FP64, PRECISION_MIXED
    r, v, ids = inputs
    n = r.shape[0]
    box = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    box.nbh_dist(RCUT, SKIN)
    box.set_particles(r, v, ids)
    prec = PRECISION_MIXED if precision == "mixed" else PRECISION_FP64

    def init():
        box.init(1.0, 1.0, RCUT, SKIN, prec)

    def run(k):
        return box.run(k, DT, 1.0, 1.0, RCUT, SKIN, precision=prec, rebuild_every=every)

    init()
    out = {}
    if want_error:
        f64, fmx = forces_both(box, keep=prec)
        out["force_error_t0"] = rel_error(fmx, f64)
    if warmup > 0:
        run(warmup)
    init()  # the timed run starts on freshly built lists
    snap = box.snapshot()
    ke0, pe0 = box.energies()
    box.synchronize()
    t0 = time.perf_counter()
    res = run(steps)
    box.synchronize()
    wall = time.perf_counter() - t0
    counts = box.run_step_counts(steps)
    ke1, pe1 = box.energies()
    p_list, p_cut = box.pair_counts()
    if want_error:
        f64, fmx = forces_both(box, keep=prec)
        out["force_error_end"] = rel_error(fmx, f64)
    out.update(dict(
        atom_steps_per_s=n * steps / wall, ms_per_step=1e3 * wall / steps, wall_s=wall, rebuilds=res["rebuilds"],
        guard_violations=res["violations"],
        energy_timed_run=dict(e_start_per_atom=(ke0 + pe0) / n, e_end_per_atom=(ke1 + pe1) / n, steps=steps),
        pairs=dict(list_entries_per_atom=p_list / n, cutoff_pairs_per_atom=p_cut / n)))
    # replay of the same steps with HIP-event timers on every kernel class
    box.restore(snap)
    box.synchronize()
    box.timer_reset()
    box.timer_enable(True)
    t0 = time.perf_counter()
    run(steps)
    box.synchronize()
    out["timed_replay_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
    box.timer_enable(False)
    out["kernels"] = timers(box)
    if drift_chunks > 0 and every > 0:
        out["energy_drift"] = drift_window(box, steps, drift_chunks, every, run, n)
    if res["violations"] > 0:
        out["missed_pairs"] = missed_pairs(box, snap, counts, every, run, init)
    box.close()
    return out


def cpu_baseline(inputs, dom, cpu_steps, repeats):
    """The oracle's step loop on the host cores, in a child process with OpenMP threads bound
    to physical cores (median of `repeats` runs after a discarded warm-up run)."""
    import psutil
    r, v, ids = inputs
    phys = psutil.cpu_count(logical=False) or os.cpu_count()
    allowed = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", phys) or phys)
    threads = max(1, min(phys, allowed, cap))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "inputs.npz")
        np.savez(path, r=r, v=v, id=ids, bounds_min=np.array(dom["bounds_min"]), bounds_max=np.array(dom["bounds_max"]),
                 grid_dims=np.array(dom["grid_dims"]), cell_size=np.array(dom["cell_size"]))
        p = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), path, "--steps", str(cpu_steps),
                            "--repeats", str(repeats), "--every", str(EVERY)], env=env, capture_output=True, text=True,
                           check=True)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    return dict(value=res["atom_steps_per_s"], unit="atom-steps/s", cores=res["threads"], kind="port",
                sample=f"same workload ({r.shape[0]} atoms), {cpu_steps} steps incl. {cpu_steps // EVERY} rebuilds, "
                       f"median of {repeats} runs ({', '.join(f'{w:.1f}' for w in res['walls_s'])} s) after a discarded "
                       f"2-step warm-up; FP64 oracle (C/OpenMP restatement of the reference operators)",
                cpu_model=model, physical_cores=phys, affinity_cpus=allowed,
                omp=dict(OMP_NUM_THREADS=threads, OMP_PROC_BIND="close", OMP_PLACES="cores"))


def _box_of(j, precision):
    from exanbody_amd import Box
    d = j["domain"]
    b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"])
    b.nbh_dist(j.get("rcut", RCUT), j.get("skin", SKIN))
    b.set_particles(j["r"], j["v"], j["id"])
    b.init(1.0, 1.0, j.get("rcut", RCUT), j.get("skin", SKIN), precision)
    return b


def job_drift(j, precision, steps, every=EVERY):
    """Energy drift of one job's trajectory, fitted over its steps (run alone in runs of
    `every` steps: the same trajectory as one run, bit for bit, and as the batched one)."""
    b = _box_of(j, precision)
    n = j["r"].shape[0]
    chunks = max(1, steps // every)
    d = drift_window(b, 0, chunks, every,
                     lambda k: b.run(k, DT, 1.0, 1.0, RCUT, SKIN, precision=precision, rebuild_every=every), n)
    b.close()
    return d


def summarize_drifts(drifts):
    a = np.array([d["per_atom_per_1000_steps"] for d in drifts])
    return dict(boxes=len(drifts), median_abs_per_atom_per_1000_steps=float(np.median(np.abs(a))) if a.size else None,
                max_abs_per_atom_per_1000_steps=float(np.max(np.abs(a))) if a.size else None)


def batched_bench(n_boxes, steps, alone_sample, drift_sample, seed=5):
    """Many independent boxes (atoms log-uniform in [500, 50000], density U[0.70, 0.90],
    T* U[0.5, 1.5], mixed precision, rebuild every 10) advanced by one batch, one launch per
    kernel per step, against one box at a time (xnb_run) on a sample of the same boxes."""
    from exanbody_amd import Batch, PRECISION_MIXED
    from exanbody_amd.workloads import batch_job_list
    jobs = batch_job_list(n_boxes, steps, seed=seed, rebuild_every=EVERY)

    t = time.perf_counter()
    boxes = [_box_of(j, PRECISION_MIXED) for j in jobs]
    setup = time.perf_counter() - t
    batch = Batch()
    slots = [batch.attach(b, steps, DT, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_MIXED, rebuild_every=EVERY) for b in boxes]
    # workgroup shapes: the batch autotuner times the batched build and force under each
    # candidate brick LDS cap before the first step (not in the timed region)
    t = time.perf_counter()
    lds_frac = batch.autotune()
    tune_s = time.perf_counter() - t
    boxes[0].synchronize()
    t0 = time.perf_counter()
    left = len(slots)
    while left:
        left = batch.step(steps)
    recoveries = sum(batch.recoveries(s) for s in slots)
    for s in slots:
        batch.detach(s)  # synchronises
    wall = time.perf_counter() - t0
    graph_steps, captures = batch.graph_stats()
    atoms = np.array([j["r"].shape[0] for j in jobs])
    # mixed vs FP64 force on every box's final positions
    f64s, fmxs = [], []
    for b in boxes:
        f64, fmx = forces_both(b)
        f64s.append(f64)
        fmxs.append(fmx)
        b.close()
    batch.close()
    err = rel_error(np.concatenate(fmxs), np.concatenate(f64s))
    # one box at a time on a sample of the same boxes
    pick = np.random.default_rng(seed + 1).choice(n_boxes, size=min(alone_sample, n_boxes), replace=False)
    t_alone, as_alone = 0.0, 0.0
    for q in pick:
        b = _box_of(jobs[q], PRECISION_MIXED)
        b.synchronize()
        t1 = time.perf_counter()
        b.run(steps, DT, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_MIXED, rebuild_every=EVERY)
        b.synchronize()
        t_alone += time.perf_counter() - t1
        as_alone += atoms[q] * steps
        b.close()
    drifts = [job_drift(jobs[q], PRECISION_MIXED, steps) for q in pick[:drift_sample]]
    batched = float(atoms.sum() * steps / wall)
    alone = float(as_alone / t_alone)
    return dict(workload=f"{n_boxes} periodic LJ boxes, atoms log-uniform [500, 50000] (total {int(atoms.sum())}), "
                         f"rho* U[0.70, 0.90], T* U[0.5, 1.5], {steps} steps each, mixed, rebuild every {EVERY}",
                atom_steps_per_s=batched, ms_per_step=1e3 * wall / steps, wall_s=wall, setup_s=setup,
                recoveries=int(recoveries), graph_steps=int(graph_steps), graph_captures=int(captures),
                autotuned_brick_lds_fraction=lds_frac, autotune_s=tune_s,
                one_at_a_time_atom_steps_per_s=alone, one_at_a_time_sample=int(len(pick)),
                speedup_vs_one_at_a_time=batched / alone,
                force_error_end=dict(err, note="mixed vs FP64 force on every box's final positions (all boxes pooled)"),
                energy_drift=dict(summarize_drifts(drifts),
                                  note="fitted over each sampled box's steps; the batched trajectory of a box is bitwise "
                                       "its trajectory run alone"))


def serving_bench(n_jobs, rate, drift_sample, seed=9):
    """Poisson arrivals of LJ jobs (atoms log-uniform [1000, 200000], rho* U[0.70, 0.90],
    T* U[0.5, 1.5], steps U[100, 2000], half FP64 / half mixed) through the continuous-batching
    loop, and the same jobs served one at a time."""
    from exanbody_amd import Box, PRECISION_MIXED
    from exanbody_amd.serve import job_mix, serve
    jobs = job_mix(n_jobs, rate, seed=seed)
    out = dict(workload=f"{n_jobs} jobs, Poisson arrivals at {rate}/s, atoms log-uniform [1000, 200000], "
                        "rho* U[0.70, 0.90], T* U[0.5, 1.5], steps U[100, 2000], half FP64 / half mixed")
    # ten steps per loop iteration: arrivals are admitted every ~1-10 ms, far below their 50 ms
    # mean spacing, and the per-iteration host work is amortised
    results = None
    for name, max_active in (("continuous_batching", 0), ("one_at_a_time", 1)):
        res, st = serve(jobs, max_active=max_active, steps_per_iter=10, outputs=(max_active == 0))
        if max_active == 0:
            results = res
        out[name] = {k: (float(v) if isinstance(v, float) else int(v)) for k, v in st.items()}
    out["steps_per_iter"] = 10
    out["throughput_ratio"] = out["continuous_batching"]["atom_steps_per_s"] / out["one_at_a_time"]["atom_steps_per_s"]
    # mixed vs FP64 force on the final positions of every job
    f64s, fmxs = [], []
    for j, res in zip(jobs, results):
        d = j["domain"]
        b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"])
        b.nbh_dist(RCUT, SKIN)
        b.set_particles(res["r"], res["v"], res["id"])
        b.init(1.0, 1.0, RCUT, SKIN, PRECISION_MIXED)
        f64, fmx = forces_both(b)
        f64s.append(f64)
        fmxs.append(fmx)
        b.close()
    out["force_error_end"] = dict(rel_error(np.concatenate(fmxs), np.concatenate(f64s)),
                                  note="mixed vs FP64 force on every job's final positions (all jobs pooled)")
    mixed_jobs = [j for j in jobs if j["precision"] == PRECISION_MIXED][:drift_sample]
    fp64_jobs = [j for j in jobs if j["precision"] != PRECISION_MIXED][:drift_sample]
    out["energy_drift"] = dict(
        mixed=summarize_drifts([job_drift(j, j["precision"], j["steps"]) for j in mixed_jobs]),
        fp64=summarize_drifts([job_drift(j, j["precision"], j["steps"]) for j in fp64_jobs]),
        note="fitted over each sampled job's steps, run alone (a job's result is bitwise the same alone and in the mix)")
    return out


def slab_bench(steps, dt=0.004, every=10):
    """Two-species slab + droplets (4,194,304 atoms, A:B 80:20, sigma_BB 1.4, LB mixing,
    rc = 2.5 sigma_ij), multimat pair path, FP64 and mixed, with the mixed force error."""
    from exanbody_amd import Box, PRECISION_FP64, PRECISION_MIXED
    from exanbody_amd._lib import TIMER_FORCE, TIMER_NBH
    from exanbody_amd.workloads import lorentz_berthelot, slab_droplets
    eps, sig, rc = lorentz_berthelot()
    r, v, ids, types, dom = slab_droplets()
    n = r.shape[0]
    out = dict(workload=f"two-species slab + 64 droplets, {n} atoms, A:B 80:20, sigma_BB 1.4, LB mixing, rc 2.5 sigma_ij, "
                        f"T* 0.8, dt {dt}, {steps} steps, rebuild every {every}, multimat")
    fref = None
    for name, prec in (("fp64", PRECISION_FP64), ("mixed", PRECISION_MIXED)):
        b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
        b.nbh_dist(float(rc.max()), SKIN)
        b.set_particles(r, v, ids, types)
        b.init(1.0, 1.0, float(rc.max()), SKIN, prec)
        b.zero_particle_force()
        b.lennard_jones_force_multimat(2, eps, sig, rc, precision=prec)
        f0 = b.export()["f"]
        if fref is None:
            fref = f0
        else:
            out["mixed_force_error_t0"] = rel_error(f0, fref)

        def run(k):
            return b.run_multimat(k, dt, 2, eps, sig, rc, SKIN, precision=prec, rebuild_every=every)

        b.synchronize()
        t0 = time.perf_counter()
        res = run(steps)
        b.synchronize()
        wall = time.perf_counter() - t0
        b.timer_reset()
        b.timer_enable(True)
        run(every)
        b.synchronize()
        b.timer_enable(False)
        fms, fn = b.timer_read(TIMER_FORCE)
        nms, nn = b.timer_read(TIMER_NBH)
        out[name] = dict(atom_steps_per_s=n * steps / wall, ms_per_step=1e3 * wall / steps, rebuilds=res["rebuilds"],
                         guard_violations=res["violations"], force_avg_ms=fms / max(fn, 1), nbh_build_avg_ms=nms / max(nn, 1),
                         energy_drift=drift_window(b, steps + every, 10, every, run, n))
        b.close()
    out["mixed_over_fp64"] = out["mixed"]["atom_steps_per_s"] / out["fp64"]["atom_steps_per_s"]
    return out


def roofline_of(head, head_prec, n):
    """Roofline of the headline force kernel (the fused LJ force + push of every step but a
    run's last): algorithmic bytes B_alg = 144 B and flops F_alg = 8 per list entry + 24 per
    in-cutoff pair, per atom-step; one launch processes every atom once."""
    fk = head["kernels"]["force"]
    t = fk["avg_ms"] * 1e-3 if fk["avg_ms"] else None
    pl, pc = head["pairs"]["list_entries_per_atom"], head["pairs"]["cutoff_pairs_per_atom"]
    f_alg = 8.0 * pl + 24.0 * pc
    p_flop = (FP32_PEAK_TFLOPS if head_prec == "mixed" else FP64_PEAK_TFLOPS) * 1e12
    ceil_bytes = HBM_PEAK_GBS * 1e9 / B_ALG
    ceil_flops = p_flop / f_alg
    byte_rate = B_ALG * n / t / 1e9 if t else None          # GB/s
    flop_rate = f_alg * n / t / 1e12 if t else None         # TFLOP/s
    hbm_bound = ceil_bytes <= ceil_flops
    return dict(
        bound="hbm" if hbm_bound else "valu",
        achieved=byte_rate if hbm_bound else flop_rate, peak=HBM_PEAK_GBS if hbm_bound else p_flop / 1e12,
        unit="GB/s" if hbm_bound else "TFLOP/s",
        frac=((byte_rate / HBM_PEAK_GBS) if hbm_bound else (flop_rate * 1e12 / p_flop)) if t else None,
        traffic=None,
        byte_frac=byte_rate / HBM_PEAK_GBS if t else None, flop_frac=flop_rate * 1e12 / p_flop if t else None,
        achieved_gbs=byte_rate, achieved_tflops=flop_rate, flop_peak_tflops=p_flop / 1e12,
        ceiling_atom_steps_per_s=min(ceil_bytes, ceil_flops), algorithmic_bytes_per_atom=B_ALG,
        algorithmic_flops_per_atom=f_alg, atoms_per_launch=n, avg_launch_ms=fk["avg_ms"], launches=fk["launches"],
        kernel=f"k_lj_brick<{'true' if head_prec == 'mixed' else 'false'}, false, 2, 8> (LJ force + push_f_v + next step's "
               "push_f_v_r/push_f_v and displacement count; f and ep not stored)",
        note="no MFMA (pair forces are not a dense contraction); frac = achieved / the binding ceiling min(HBM 8 TB/s / "
             "144 B, vector peak / F_alg) with the FP32 peak for the mixed form and the FP64 peak for the plain form; "
             "avg_launch_ms from HIP events on the context stream over the timed steps replayed; traffic (measured HBM "
             "bytes per launch) comes from rocprofv3 --pmc runs, not from this run: see profiles/")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--trigger", choices=["interval", "displacement"], default="interval")
    ap.add_argument("--drift-chunks", type=int, default=20, help="runs of 10 steps in the drift window")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--batch-boxes", type=int, default=512, help="batched workload size (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-alone-sample", type=int, default=32)
    ap.add_argument("--drift-sample", type=int, default=8, help="boxes / jobs whose energy drift is fitted")
    ap.add_argument("--serve-jobs", type=int, default=64, help="serving workload size (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--slab-steps", type=int, default=1000, help="two-species slab workload steps (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from exanbody_amd import Box, cubic_domain, generate_fcc
    w = WORKLOADS[args.workload]
    t = time.perf_counter()
    r, v, ids, L = generate_fcc(w["ncell"], RHO, jitter=JITTER, temperature=T0, seed=w["seed"])
    dom = cubic_domain(L, RCUT + SKIN)
    n = r.shape[0]
    log(f"generated {n} atoms, L={L:.6f}, cells {dom['grid_dims']} edge {dom['cell_size']:.6f} "
        f"in {time.perf_counter() - t:.1f}s")
    every = EVERY if args.trigger == "interval" else 0

    probe = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    copy_gbs = probe.copy_probe(1 << 30, 10)
    probe.close()
    log(f"copy probe: {copy_gbs:.0f} GB/s")

    results = {}
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    for p in precs:
        results[p] = run_precision((r, v, ids), dom, p, args.steps, args.warmup, every,
                                   want_error=(p == "mixed" and not args.no_error), drift_chunks=args.drift_chunks)
        log(p, json.dumps({k: results[p][k] for k in ("atom_steps_per_s", "ms_per_step", "rebuilds", "guard_violations")}))

    head = results["mixed"] if "mixed" in results else results["fp64"]
    head_prec = "mixed" if "mixed" in results else "fp64"
    roofline = roofline_of(head, head_prec, n)
    batched = serving = slab = cpu = None
    if args.batch_boxes > 0:
        log("batched ...")
        batched = batched_bench(args.batch_boxes, args.batch_steps, args.batch_alone_sample, args.drift_sample)
        log("batched", json.dumps(batched))
    if args.serve_jobs > 0:
        log("serving ...")
        serving = serving_bench(args.serve_jobs, args.serve_rate, args.drift_sample)
        log("serving", json.dumps(serving))
    if args.slab_steps > 0:
        log("slab ...")
        slab = slab_bench(args.slab_steps)
        log("slab", json.dumps(slab))
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline((r, v, ids), dom, args.cpu_steps, args.cpu_repeats)
        log("cpu", json.dumps(cpu))
    line = dict(
        metric=METRIC, value=head["atom_steps_per_s"], unit="atom-steps/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=head["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32 pair / f64 accumulate" if head_prec == "mixed" else "f64", data="synthetic",
        config=dict(workload=w["label"] + f", rho*={RHO}, T*={T0}, rc={RCUT}, skin={SKIN}, dt={DT}, "
                                          f"rebuild {'every %d steps' % EVERY if every else 'on displacement'}",
                    atoms=n, precision=head_prec, trigger=args.trigger,
                    mixed_over_fp64=(results["mixed"]["atom_steps_per_s"] / results["fp64"]["atom_steps_per_s"])
                    if len(results) == 2 else None),
        results=results, roofline=roofline, copy_probe_gbs=copy_gbs, batched=batched, serving=serving,
        two_species_slab=slab, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
