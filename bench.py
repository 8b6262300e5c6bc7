"""Benchmark: LJ atom-steps/s (rc=2.5 sigma), mixed vs FP64 on one MI355X.

Headline workload: FCC 128^3 unit cells (8,388,608 atoms), rho*=0.8442, T*=0.72 seeded
Maxwell-Boltzmann velocities, +-0.01 sigma jitter, rc=2.5, skin=0.3, dt=0.005, neighbour
rebuild every 10 steps, velocity-Verlet.  Both precisions run the same device-resident
step loop (push, trigger, ghost refresh or full rebuild, force + push).  value = mixed
atom-steps/s; the FP64 plain form is timed beside it.

Beside it, in the same JSON line: many boxes in one batch against one box at a time
("batched"), the continuous-batching serving loop with p50/p99 job latency ("serving"), and
the two-species slab on the multimat path ("two_species_slab").

Prints ONE JSON line on stdout (the last line); progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LJ atom-steps/s (rc=2.5σ) mixed vs FP64 on 1 GPU; batched p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the headline force kernel k_lj_brick<MIXED, false, 2, 8> on this
# workload: 2 x FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc, calibrated on the copy probe
# (profiles/lj_c2_force_pmc.md)
FORCE_TRAFFIC_BYTES = {"mixed": 2.882e9, "fp64": 3.215e9}
# rocprofv3 VALUBusy (%) of the same kernel on this workload (profiles/lj_c2_force_valu_lds_pmc.md)
FORCE_VALU_BUSY_PCT = {"mixed": 70.7, "fp64": 70.0}
VALU_PEAK_TFLOPS = 78.6  # FP64 vector; scalar FP32 issues at the same rate (packed FP32 disabled)

WORKLOADS = {
    "c2": dict(ncell=128, seed=2, label="FCC 128^3 LJ fluid, 8,388,608 atoms"),
    "c1": dict(ncell=32, seed=1, label="FCC 32^3 LJ fluid, 131,072 atoms"),
}
RHO, T0, JITTER, RCUT, SKIN, DT, EVERY = 0.8442, 0.72, 0.01, 2.5, 0.3, 0.005, 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def force_error(box, precision_ref, precision_test, eps=1.0, sigma=1.0):
    """Force error of precision_test vs precision_ref on the box's current positions."""
    box.zero_particle_force()
    box.lennard_jones_force(eps, sigma, RCUT, precision=precision_ref)
    ref = box.export()
    box.zero_particle_force()
    box.lennard_jones_force(eps, sigma, RCUT, precision=precision_test)
    tst = box.export()
    d = np.linalg.norm(tst["f"] - ref["f"], axis=1)
    fn = np.linalg.norm(ref["f"], axis=1)
    rms_ref = np.sqrt(np.mean(fn * fn))
    # restore reference forces (the step loop continues from them)
    box.zero_particle_force()
    box.lennard_jones_force(eps, sigma, RCUT, precision=precision_ref)
    return dict(rms_rel=float(np.sqrt(np.mean(d * d)) / rms_ref), max_rel=float(d.max() / rms_ref),
                max_rel_per_atom=float(np.max(d / np.maximum(fn, 1e-300))), rms_force=float(rms_ref))


def run_precision(inputs, dom, precision, steps, warmup, every, want_error):
    from exanbody_amd import Box, PRECISION_FP64, PRECISION_MIXED
    from exanbody_amd._lib import TIMER_FORCE, TIMER_GHOST, TIMER_NBH, TIMER_PUSH, TIMER_REBIN
    r, v, ids = inputs
    n = r.shape[0]
    box = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    box.nbh_dist(RCUT, SKIN)
    box.set_particles(r, v, ids)
    prec = PRECISION_MIXED if precision == "mixed" else PRECISION_FP64
    box.init(1.0, 1.0, RCUT, SKIN, prec)
    out = {}
    if want_error:
        out["force_error_t0"] = force_error(box, PRECISION_FP64, PRECISION_MIXED)
        if precision == "mixed":  # init state must be the mixed one
            box.zero_particle_force()
            box.lennard_jones_force(1.0, 1.0, RCUT, precision=prec)
    ke0, pe0 = box.energies()
    if warmup > 0:
        box.run(warmup, DT, 1.0, 1.0, RCUT, SKIN, precision=prec, rebuild_every=every)
    box.synchronize()
    box.timer_reset()
    box.timer_enable(True)
    t0 = time.perf_counter()
    res = box.run(steps, DT, 1.0, 1.0, RCUT, SKIN, precision=prec, rebuild_every=every)
    box.synchronize()
    t1 = time.perf_counter()
    box.timer_enable(False)
    wall = t1 - t0
    kern = {}
    for name, tid in (("force", TIMER_FORCE), ("nbh_build", TIMER_NBH), ("push", TIMER_PUSH), ("ghost", TIMER_GHOST),
                      ("rebin", TIMER_REBIN)):
        ms, cnt = box.timer_read(tid)
        kern[name] = dict(total_ms=round(ms, 4), launches=int(cnt), avg_ms=round(ms / cnt, 5) if cnt else None)
    ke1, pe1 = box.energies()
    e0, e1 = ke0 + pe0, ke1 + pe1
    p_list, p_cut = box.pair_counts()
    out.update(dict(
        atom_steps_per_s=n * steps / wall, ms_per_step=1e3 * wall / steps, wall_s=wall, rebuilds=res["rebuilds"],
        guard_violations=res["violations"], kernels=kern,
        energy=dict(e_start_per_atom=e0 / n, e_end_per_atom=e1 / n, steps_covered=steps + warmup,
                    drift_per_atom_per_1000_steps=(e1 - e0) / n * 1000.0 / (steps + warmup),
                    rel_drift_per_1000_steps=(e1 - e0) / abs(e0) * 1000.0 / (steps + warmup)),
        pairs=dict(list_entries_per_atom=p_list / n, cutoff_pairs_per_atom=p_cut / n)))
    if want_error:
        out["force_error_end"] = force_error(box, PRECISION_FP64, PRECISION_MIXED)
    box.close()
    return out


def cpu_baseline(inputs, dom, cpu_steps):
    from oracle.oracle import Oracle, num_threads
    r, v, ids = inputs
    o = Oracle(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"], ghost_dist=RCUT + SKIN)
    o.set_particles(r, v, ids)
    o.init(1.0, 1.0, RCUT, SKIN)
    t0 = time.perf_counter()
    o.run(cpu_steps, DT, 1.0, 1.0, RCUT, SKIN, rebuild_every=EVERY)
    wall = time.perf_counter() - t0
    return dict(value=r.shape[0] * cpu_steps / wall, unit="atom-steps/s", cores=num_threads(), kind="port",
                sample=f"same workload ({r.shape[0]} atoms), {cpu_steps} steps incl. {cpu_steps // EVERY} rebuilds, "
                       f"FP64 oracle (C/OpenMP restatement of the reference operators), {wall:.1f} s wall")


def batched_bench(n_boxes, steps, alone_sample, seed=5):
    """Many independent boxes (atoms log-uniform in [500, 50000], density U[0.70, 0.90],
    T* U[0.5, 1.5], mixed precision, rebuild every 10) advanced by one batch, one launch per
    kernel per step, against one box at a time (xnb_run) on a sample of the same boxes."""
    from exanbody_amd import Batch, Box, PRECISION_MIXED
    from exanbody_amd.serve import make_job
    rng = np.random.default_rng(seed)
    jobs = []
    for q in range(n_boxes):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
        jobs.append(make_job(n, rng.uniform(0.70, 0.90), rng.uniform(0.5, 1.5), steps, PRECISION_MIXED, seed=7000 + q))

    def make_box(j):
        d = j["domain"]
        b = Box(d["bounds_min"], d["bounds_max"], d["grid_dims"], d["cell_size"])
        b.nbh_dist(RCUT, SKIN)
        b.set_particles(j["r"], j["v"], j["id"])
        b.init(1.0, 1.0, RCUT, SKIN, PRECISION_MIXED)
        return b

    t = time.perf_counter()
    boxes = [make_box(j) for j in jobs]
    setup = time.perf_counter() - t
    batch = Batch()
    slots = [batch.attach(b, steps, DT, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_MIXED, rebuild_every=EVERY) for b in boxes]
    boxes[0].synchronize()
    t0 = time.perf_counter()
    left = len(slots)
    while left:
        left = batch.step(steps)
    for s in slots:
        batch.detach(s)  # synchronises
    wall = time.perf_counter() - t0
    atoms = np.array([j["r"].shape[0] for j in jobs])
    for b in boxes:
        b.close()
    batch.close()
    # one box at a time on a sample of the same boxes
    pick = np.random.default_rng(seed + 1).choice(n_boxes, size=min(alone_sample, n_boxes), replace=False)
    t_alone, as_alone = 0.0, 0.0
    for q in pick:
        b = make_box(jobs[q])
        b.synchronize()
        t1 = time.perf_counter()
        b.run(steps, DT, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_MIXED, rebuild_every=EVERY)
        b.synchronize()
        t_alone += time.perf_counter() - t1
        as_alone += atoms[q] * steps
        b.close()
    batched = float(atoms.sum() * steps / wall)
    alone = float(as_alone / t_alone)
    return dict(workload=f"{n_boxes} periodic LJ boxes, atoms log-uniform [500, 50000] (total {int(atoms.sum())}), "
                         f"rho* U[0.70, 0.90], T* U[0.5, 1.5], {steps} steps each, mixed, rebuild every {EVERY}",
                atom_steps_per_s=batched, ms_per_step=1e3 * wall / steps, wall_s=wall, setup_s=setup,
                one_at_a_time_atom_steps_per_s=alone, one_at_a_time_sample=int(len(pick)),
                speedup_vs_one_at_a_time=batched / alone)


def serving_bench(n_jobs, rate, seed=9):
    """Poisson arrivals of LJ jobs (atoms log-uniform [1000, 200000], rho* U[0.70, 0.90],
    T* U[0.5, 1.5], steps U[100, 2000], half FP64 / half mixed) through the continuous-batching
    loop, and the same jobs served one at a time."""
    from exanbody_amd.serve import job_mix, serve
    jobs = job_mix(n_jobs, rate, seed=seed)
    out = dict(workload=f"{n_jobs} jobs, Poisson arrivals at {rate}/s, atoms log-uniform [1000, 200000], "
                        "rho* U[0.70, 0.90], T* U[0.5, 1.5], steps U[100, 2000], half FP64 / half mixed")
    # ten steps per loop iteration: arrivals are admitted every ~1-10 ms, far below their 50 ms
    # mean spacing, and the per-iteration host work is amortised (p50 0.16 -> 0.13 s measured)
    for name, max_active in (("continuous_batching", 0), ("one_at_a_time", 1)):
        _, st = serve(jobs, max_active=max_active, steps_per_iter=10)
        out[name] = {k: (float(v) if isinstance(v, float) else int(v)) for k, v in st.items()}
    out["steps_per_iter"] = 10
    out["throughput_ratio"] = out["continuous_batching"]["atom_steps_per_s"] / out["one_at_a_time"]["atom_steps_per_s"]
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
            d = np.linalg.norm(f0 - fref, axis=1)
            rms = np.sqrt(np.mean(np.sum(fref * fref, axis=1)))
            out["mixed_force_error_t0"] = dict(rms_rel=float(np.sqrt(np.mean(d * d)) / rms), max_rel=float(d.max() / rms))
        ke0, pe0 = b.energies()
        b.synchronize()
        b.timer_reset()
        b.timer_enable(True)
        t0 = time.perf_counter()
        res = b.run_multimat(steps, dt, 2, eps, sig, rc, SKIN, precision=prec, rebuild_every=every)
        b.synchronize()
        wall = time.perf_counter() - t0
        b.timer_enable(False)
        ke1, pe1 = b.energies()
        fms, fn = b.timer_read(TIMER_FORCE)
        nms, nn = b.timer_read(TIMER_NBH)
        out[name] = dict(atom_steps_per_s=n * steps / wall, ms_per_step=1e3 * wall / steps, rebuilds=res["rebuilds"],
                         guard_violations=res["violations"], force_avg_ms=fms / max(fn, 1), nbh_build_avg_ms=nms / max(nn, 1),
                         energy_drift_per_atom_per_1000_steps=((ke1 + pe1) - (ke0 + pe0)) / n * 1000.0 / steps)
        b.close()
    out["mixed_over_fp64"] = out["mixed"]["atom_steps_per_s"] / out["fp64"]["atom_steps_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--trigger", choices=["interval", "displacement"], default="interval")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--batch-boxes", type=int, default=512, help="batched workload size (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-alone-sample", type=int, default=32)
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
                                   want_error=(p == "mixed" and not args.no_error))
        log(p, json.dumps({k: results[p][k] for k in ("atom_steps_per_s", "ms_per_step", "rebuilds",
                                                        "guard_violations")}))

    head = results["mixed"] if "mixed" in results else results["fp64"]
    head_prec = "mixed" if "mixed" in results else "fp64"
    fk = head["kernels"]["force"]
    pl, pc = head["pairs"]["list_entries_per_atom"], head["pairs"]["cutoff_pairs_per_atom"]
    flops_per_atom = 8.0 * pl + 24.0 * pc
    achieved_tflops = flops_per_atom * n / (fk["avg_ms"] * 1e-3) / 1e12 if fk["avg_ms"] else None
    traffic = FORCE_TRAFFIC_BYTES[head_prec] if args.workload == "c2" else None
    roofline = dict(bound="mfma", achieved=achieved_tflops, peak=VALU_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=(achieved_tflops / VALU_PEAK_TFLOPS) if achieved_tflops else None, traffic=traffic,
                    traffic_gbs=traffic / (fk["avg_ms"] * 1e-3) / 1e9 if (traffic and fk["avg_ms"]) else None,
                    kernel=f"k_lj_brick<{'true' if head_prec == 'mixed' else 'false'}, false, 2, 8> (LJ force + push_f_v + "
                           "next step's push_f_v_r/push_f_v and displacement count, f and ep not stored)",
                    algorithmic_flops_per_atom=flops_per_atom, avg_launch_ms=fk["avg_ms"],
                    valu_busy_pct=FORCE_VALU_BUSY_PCT[head_prec] if args.workload == "c2" else None,
                    note="VALU pair kernel (no MFMA): flops = 8/list entry + 24/in-cutoff pair; peak = FP64 vector "
                         "rate, which scalar FP32 shares with packed FP32 disabled; traffic = measured HBM bytes per "
                         "launch, traffic_gbs = traffic / average launch time (HBM peak 8000 GB/s); valu_busy_pct = "
                         "rocprofv3 VALUBusy of the kernel")
    b_alg = 144.0
    step_ceiling = min(HBM_PEAK_GBS * 1e9 / b_alg, VALU_PEAK_TFLOPS * 1e12 * (2.0 if head_prec == "mixed" else 1.0) /
                       flops_per_atom)
    batched = serving = None
    if args.batch_boxes > 0:
        log("batched ...")
        batched = batched_bench(args.batch_boxes, args.batch_steps, args.batch_alone_sample)
        log("batched", json.dumps(batched))
    if args.serve_jobs > 0:
        log("serving ...")
        serving = serving_bench(args.serve_jobs, args.serve_rate)
        log("serving", json.dumps(serving))
    slab = None
    if args.slab_steps > 0:
        log("slab ...")
        slab = slab_bench(args.slab_steps)
        log("slab", json.dumps(slab))
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline((r, v, ids), dom, args.cpu_steps)
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
        results=results, roofline=roofline,
        step_roofline=dict(b_alg_bytes_per_atom_step=b_alg, flops_per_atom_step=flops_per_atom,
                           ceiling_atom_steps_per_s=step_ceiling, frac=head["atom_steps_per_s"] / step_ceiling),
        copy_probe_gbs=copy_gbs, batched=batched, serving=serving, two_species_slab=slab, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
