"""Batched boxes (xnb_batch): every box advanced in a batch, alone or in the mix, attached and
detached at different steps, ends bitwise equal to the same box advanced by xnb_run alone --
FP64 and mixed precision alike (same per-atom operations in the same order), rebuilt at a
fixed interval or on the displacement trigger."""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED, Batch, Box, cubic_domain, generate_fcc

pytestmark = pytest.mark.gpu

RCUT, SKIN, DT = 2.5, 0.3, 0.005

# (ncell, density, temperature, seed, steps, precision, rebuild_every)
JOBS = [
    (4, 0.84, 0.72, 11, 23, PRECISION_FP64, 10),
    (5, 0.75, 1.20, 12, 31, PRECISION_MIXED, 10),
    (6, 0.90, 0.50, 13, 17, PRECISION_FP64, 5),
    (7, 0.80, 1.00, 14, 40, PRECISION_MIXED, 10),
    (5, 0.70, 1.50, 15, 12, PRECISION_MIXED, 4),
    (8, 0.85, 0.80, 16, 26, PRECISION_FP64, 10),
    # rebuilt on the displacement trigger (hot, so atoms pass max_displ every few steps)
    (5, 0.80, 2.50, 17, 40, PRECISION_FP64, 0),
    (6, 0.75, 3.00, 18, 35, PRECISION_MIXED, 0),
    # about 62 atoms per cell: rebuild_amr reorders every cell at each rebuild
    (10, 1.60, 0.72, 19, 15, PRECISION_FP64, 5),
]


def _box(job):
    ncell, rho, temp, seed = job[:4]
    r, v, ids, L = generate_fcc(ncell, rho, jitter=0.01, temperature=temp, seed=seed)
    dom = cubic_domain(L, RCUT + SKIN)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(RCUT, SKIN)
    b.set_particles(r, v, ids)
    b.init(1.0, 1.0, RCUT, SKIN, job[5])
    return b


def _alone(job):
    b = _box(job)
    res = b.run(job[4], DT, 1.0, 1.0, RCUT, SKIN, precision=job[5], rebuild_every=job[6])
    out = b.export(), res
    b.close()
    return out


def _same(a, b, what):
    for k in ("cell", "id"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    for k in ("r", "v", "f", "ep"):
        np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64), err_msg=f"{what}: {k} (bitwise)")


def test_batch_equals_alone_all_at_once():
    ref = [_alone(j) for j in JOBS]
    batch = Batch()
    boxes, slots = [], []
    for j in JOBS:
        b = _box(j)
        boxes.append(b)
        slots.append(batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6]))
    left = len(JOBS)
    guard = 0
    while left:
        left = batch.step(7)
        guard += 1
        assert guard < 100
    for q, (j, b, s) in enumerate(zip(JOBS, boxes, slots)):
        assert batch.status(s) == (j[4], 1)
        res = batch.detach(s)
        assert res == ref[q][1], f"job {q}: rebuild / violation counts"
        if j[6] == 0:
            assert res["rebuilds"] >= 2 and res["violations"] == 0, f"job {q}: displacement-triggered rebuilds {res}"
        _same(b.export(), ref[q][0], f"job {q}")
        b.close()
    batch.close()


def test_batch_continuous_admission():
    """Jobs arrive while others run (attach between steps) and leave when done; each result
    is the same as alone."""
    ref = [_alone(j) for j in JOBS]
    batch = Batch()
    pending = list(range(len(JOBS)))
    running = {}
    results = {}
    t = 0
    while pending or running:
        # admit one job every 5 steps, at most 3 at a time
        if pending and len(running) < 3 and t % 5 == 0:
            q = pending.pop(0)
            b = _box(JOBS[q])
            j = JOBS[q]
            running[batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6])] = (q, b)
        batch.step(1)
        t += 1
        for s in list(running):
            done, state = batch.status(s)
            assert state >= 0
            if state == 1:
                q, b = running.pop(s)
                assert done == JOBS[q][4]
                results[q] = (batch.detach(s), b.export())
                b.close()
        assert t < 1000
    for q in range(len(JOBS)):
        assert results[q][0] == ref[q][1], f"job {q}: counts"
        _same(results[q][1], ref[q][0], f"job {q}")
    batch.close()


@pytest.mark.parametrize("graph", ["1", "0"])
def test_batch_quiet_steps_from_graph(monkeypatch, graph):
    """Jobs rebuilt at fixed intervals: steps in which no box starts or rebuilds are replayed
    from a captured HIP graph (many replays, with eager rebuild steps and synchronisations in
    between, and re-captures when a box's last step changes the work table); every box still
    ends bitwise equal to the same box run alone, as with graphs switched off."""
    monkeypatch.setenv("XNB_BATCH_GRAPH", graph)
    jobs = [j for j in JOBS if j[6] > 0]
    ref = [_alone(j) for j in jobs]
    batch = Batch()
    boxes, slots = [], []
    for j in jobs:
        b = _box(j)
        boxes.append(b)
        slots.append(batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6]))
    left = len(jobs)
    while left:
        left = batch.step(3)
    replays, captures = batch.graph_stats()
    if graph == "1":
        assert replays >= 5 and captures >= 1, (replays, captures)
    else:
        assert replays == 0 and captures == 0
    for q, (j, b, s) in enumerate(zip(jobs, boxes, slots)):
        assert batch.status(s) == (j[4], 1)
        assert batch.detach(s) == ref[q][1], f"job {q}: rebuild / violation counts"
        _same(b.export(), ref[q][0], f"job {q}")
        b.close()
    batch.close()


def test_batch_rebuild_steps_from_graph(monkeypatch):
    """Boxes that all rebuild every 5 steps: a rebuild step (move, rebuild_amr, backup, ghosts,
    neighbour build, force, clock) is one launch sequence, captured at its second occurrence
    with the same step parity and replayed at the later ones, beside the quiet-step graph.  Of
    the 31 steps, step 1 starts the boxes and step 31 closes them (eager); the quiet steps give
    at most 22 replays (steps 3-4, then four after each rebuild); the rebuild steps alternate
    parity, so 15 and 20 are captured and replayed and 25 and 30 replayed: four more.  Every
    box ends bitwise equal to the same box run alone."""
    monkeypatch.setenv("XNB_BATCH_GRAPH", "1")
    jobs = [
        (6, 0.84, 0.72, 31, 31, PRECISION_MIXED, 5),
        (5, 0.80, 1.00, 32, 31, PRECISION_FP64, 5),
        (7, 0.75, 0.80, 33, 31, PRECISION_MIXED, 5),
        (4, 0.85, 0.60, 34, 31, PRECISION_FP64, 5),
    ]
    ref = [_alone(j) for j in jobs]
    batch = Batch()
    boxes, slots = [], []
    for j in jobs:
        b = _box(j)
        boxes.append(b)
        slots.append(batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6]))
    while batch.step(4):
        pass
    replays, captures = batch.graph_stats()
    rec = [batch.recoveries(s) for s in slots]
    if sum(rec) == 0:  # a recovery changes the box's tables, and with them what is replayed
        assert replays >= 25 and captures >= 2, (replays, captures)
    for q, (j, b, s) in enumerate(zip(jobs, boxes, slots)):
        assert batch.status(s) == (j[4], 1)
        assert batch.detach(s) == ref[q][1], f"job {q}: rebuild / violation counts"
        _same(b.export(), ref[q][0], f"job {q}")
        b.close()
    batch.close()


def test_batch_autotune_keeps_results():
    """The batch autotuner lays every box out under each candidate LDS cap (rebuilding its lists
    and recomputing its forces in place) and keeps the fastest; the boxes then run to the same
    bitwise end states as alone, and tuning after a step is refused."""
    jobs = [JOBS[0], JOBS[1], JOBS[3], JOBS[5]]
    ref = [_alone(j) for j in jobs]
    batch = Batch()
    boxes, slots = [], []
    for j in jobs:
        b = _box(j)
        boxes.append(b)
        slots.append(batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6]))
    frac = batch.autotune()
    assert any(abs(frac - c) < 1e-12 for c in (1.0 / 3.0, 0.5, 0.75, 1.0)), frac
    batch.step(1)
    with pytest.raises(Exception):
        batch.autotune()
    while batch.step(5):
        pass
    for q, (j, b, s) in enumerate(zip(jobs, boxes, slots)):
        assert batch.status(s) == (j[4], 1)
        assert batch.detach(s) == ref[q][1], f"job {q}: rebuild / violation counts"
        _same(b.export(), ref[q][0], f"job {q} (autotuned LDS cap {frac})")
        b.close()
    batch.close()


def test_batch_recovery_path(monkeypatch):
    """With no head room (XNB_BATCH_MARGINS=1,1) rebuilds that grow a brick, a region or the
    ghosts overflow the sizes a box was attached with, and the batch redoes them with the
    one-box operators (recover_box, the ghost-overflow skip of the batched build): recoveries
    happen, and every box still ends bitwise equal to the same box run alone."""
    monkeypatch.setenv("XNB_BATCH_MARGINS", "1.0,1.0")
    jobs = [
        (5, 0.80, 2.50, 21, 30, PRECISION_FP64, 0),
        (6, 0.75, 3.00, 22, 30, PRECISION_MIXED, 5),
        (4, 0.70, 2.00, 23, 25, PRECISION_FP64, 4),
    ]
    ref = [_alone(j) for j in jobs]
    batch = Batch()
    boxes, slots = [], []
    for j in jobs:
        b = _box(j)
        boxes.append(b)
        slots.append(batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6]))
    while batch.step(4):
        pass
    rec = [batch.recoveries(s) for s in slots]
    assert sum(rec) > 0, f"no recovery happened: {rec}"
    for q, (j, b, s) in enumerate(zip(jobs, boxes, slots)):
        assert batch.status(s) == (j[4], 1)
        assert batch.detach(s) == ref[q][1], f"job {q}: rebuild / violation counts"
        _same(b.export(), ref[q][0], f"job {q} ({rec[q]} recoveries)")
        b.close()
    batch.close()


def test_detach_unfinished_box_refuses_run():
    """A box detached before its last step holds part of its next step; xnb_run refuses it."""
    from exanbody_amd import XnbError
    j = JOBS[0]
    batch = Batch()
    b = _box(j)
    s = batch.attach(b, j[4], DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6])
    batch.step(3)
    assert batch.status(s) == (3, 0)
    batch.detach(s)
    with pytest.raises(XnbError):
        b.run(5, DT, 1.0, 1.0, RCUT, SKIN, precision=j[5], rebuild_every=j[6])
    b.close()
    batch.close()


def test_margins_knob_is_checked(monkeypatch):
    monkeypatch.setenv("XNB_BATCH_MARGINS", "0.5,1.0")
    from exanbody_amd import XnbError
    with pytest.raises(XnbError):
        Batch()
