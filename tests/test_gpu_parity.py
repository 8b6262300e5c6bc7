"""GPU (HIP library) vs CPU oracle on the same seeded inputs.

Bar: bit-exact for cell layout, ids, ghost images, neighbour streams and the FP64 trajectory;
mixed precision within the stated force tolerance.
"""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED
from tests.helpers import RCUT, SKIN, assert_same_layout, fcc_inputs, make_pair, rel_err

pytestmark = pytest.mark.gpu


def _init_both(ncell=8, **kw):
    r, v, ids, L = fcc_inputs(ncell, **kw)
    o, b, dom = make_pair(r, v, ids, L)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    return o, b


def test_init_layout_ghosts_streams_forces_bitwise():
    o, b = _init_both(8)
    assert_same_layout(o.export(), b.export(), "inner")
    og, gg = o.export(ghosts=True), b.export(ghosts=True)
    assert_same_layout(og, gg, "inner+ghost")
    # neighbour streams of inner cells: bit-exact CS=1 format
    gs, gmax = b.streams()
    os_ = o.streams()
    dims, gl, nc = b.grid_info()
    counts = b.cell_counts()
    assert gmax == 78  # FCC at rho=0.8442: exactly 78 neighbours within rc+skin
    checked = 0
    for c in range(nc):
        k, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        inner = all(gl <= x < d - gl for x, d in zip((i, j, k), dims))
        if not inner:
            continue
        np.testing.assert_array_equal(gs[c], os_[c], err_msg=f"stream of cell {c}")
        checked += counts[c]
    assert checked == 2048
    oe, ge = o.export(), b.export()
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64), err_msg="FP64 forces bitwise")
    np.testing.assert_array_equal(oe["ep"].view(np.uint64), ge["ep"].view(np.uint64), err_msg="FP64 energies bitwise")


@pytest.mark.parametrize("rebuild_every", [0, 10])
def test_fp64_trajectory_bitwise(rebuild_every):
    o, b = _init_both(8, temperature=1.5, jitter=0.05, seed=7)
    steps = 40
    ro = o.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, rebuild_every=rebuild_every)
    rg = b.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=rebuild_every)
    assert ro == rg
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after run")
    for k in ("v", "f"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    ko, po = o.energies()
    kg, pg = b.energies()
    assert abs(ko - kg) <= 1e-12 * abs(ko) and abs(po - pg) <= 1e-12 * abs(po)


def test_fcc32_full_size_100_steps_bitwise():
    """The reference's CPU-runnable workload at full size: FCC 32^3 (131,072 atoms, 19^3 cells),
    rho 0.8442, T 0.72, jitter 0.01, 100 steps of dt 0.005 rebuilt every 10.  The FP64 run
    equals the oracle bit for bit (layout, velocities, per-atom forces and energies, rebuild and
    guard counts, the CS=1 streams of all 6,859 inner cells at init and after the last rebuild;
    the box totals, summed in another order, within 1e-12); mixed-precision forces at the final
    positions are within 1e-5 RMS of FP64."""
    r, v, ids, L = fcc_inputs(32, temperature=0.72, jitter=0.01, seed=1)
    o, b, dom = make_pair(r, v, ids, L)
    assert dom["grid_dims"] == (19, 19, 19)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)

    def same_streams(what):
        gs, gmax = b.streams()
        os_ = o.streams()
        dims, gl, nc = b.grid_info()
        cells = 0
        for c in range(nc):
            k, rem = divmod(c, dims[0] * dims[1])
            j, i = divmod(rem, dims[0])
            if all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
                np.testing.assert_array_equal(gs[c], os_[c], err_msg=f"{what}: stream of cell {c}")
                cells += 1
        assert cells == 19 ** 3 and gmax == o.max_neighbors()

    same_streams("init")
    ro = o.run(100, 0.005, 1.0, 1.0, RCUT, SKIN, rebuild_every=10)
    rg = b.run(100, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=10)
    assert ro == rg and rg["rebuilds"] == 10
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after 100 steps")
    same_streams("after the run")
    for k in ("v", "f", "ep"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    (ko, po), (kg, pg) = o.energies(), b.energies()
    assert abs(ko - kg) <= 1e-12 * abs(ko) and abs(po - pg) <= 1e-12 * abs(po)
    b.zero_particle_force()
    b.lennard_jones_force(1.0, 1.0, RCUT, precision=PRECISION_MIXED)
    mx, rms = rel_err(b.export()["f"], ge["f"])
    assert rms <= 1e-5, (mx, rms)


def test_fcc128_full_size_properties():
    """The 8,388,608-atom FCC 128^3 workload at full size, through properties that hold at any
    size: with jitter within +-0.01 no FCC shell crosses 2.5 or 2.8, so every atom has exactly
    78 list neighbours and 54 within the cutoff; the pair forces of full lists cancel, so the
    total force is zero up to the rounding of the per-atom sums; ids are conserved; mixed
    forces are within 1e-5 RMS of the FP64 ones."""
    import ctypes as C

    from exanbody_amd import Box, cubic_domain
    r, v, ids, L = fcc_inputs(128, temperature=0.72, jitter=0.01, seed=2)
    n = len(r)
    dom = cubic_domain(L, RCUT + SKIN)
    assert dom["grid_dims"] == (76, 76, 76) and n == 8388608
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(RCUT, SKIN)
    b.set_particles(r, v, ids)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    assert b.pair_counts() == (78 * n, 54 * n)
    tot, mx = C.c_size_t(), C.c_uint32()
    b._check(b.L.xnb_chunk_neighbors_stream(b._h, None, 0, None, C.byref(tot), C.byref(mx)), "stream size")
    assert mx.value == 78
    e = b.export()
    np.testing.assert_array_equal(np.sort(e["id"]), np.arange(n, dtype=np.uint64))
    f = e["f"]
    assert np.all(np.abs(f.sum(axis=0)) <= 1e-9 * np.abs(f).sum(axis=0))
    b.zero_particle_force()
    b.lennard_jones_force(1.0, 1.0, RCUT, precision=PRECISION_MIXED)
    mxe, rms = rel_err(b.export()["f"], f)
    assert rms <= 1e-5, (mxe, rms)


def test_mixed_force_error():
    o, b = _init_both(8, temperature=0.72, seed=3)
    b.run(50, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=10)
    ref = b.export()
    b.zero_particle_force()
    b.lennard_jones_force(1.0, 1.0, RCUT, precision=PRECISION_MIXED)
    mix = b.export()
    mx, rms = rel_err(mix["f"], ref["f"])
    assert rms <= 1e-5, (mx, rms)
    pe_rel = np.abs(mix["ep"] - ref["ep"]).max() / np.abs(ref["ep"]).max()
    assert pe_rel <= 1e-5


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_lj_force_on_ghost_cells(precision):
    """lennard_jones_force with its ghost slot set (lennard_jones.cu:174-181) over lists of every
    cell (chunk_neighbors over all cells; ghost cells see the neighbours inside the grid):
    forces and energies of inner and ghost particles equal the oracle's, bitwise for FP64 and
    within the mixed tolerance otherwise."""
    o, b = _init_both(6, temperature=0.9, seed=23)
    o.chunk_neighbors(RCUT + SKIN)
    b.chunk_neighbors(RCUT + SKIN, all_cells=True)
    o.zero_force(ghost=True)
    b.zero_particle_force(ghost=True)
    o.lj_force(1.0, 1.0, RCUT, ghost=True)
    b.lennard_jones_force(1.0, 1.0, RCUT, precision=precision, ghost=True)
    oe, ge = o.export(ghosts=True), b.export(ghosts=True)
    assert_same_layout(oe, ge, "inner+ghost")
    dims, gl, _ = b.grid_info()
    c = ge["cell"].astype(np.int64)
    ijk = np.stack([c % dims[0], (c // dims[0]) % dims[1], c // (dims[0] * dims[1])], 1)
    ghost = np.any((ijk < gl) | (ijk >= np.asarray(dims) - gl), axis=1)
    assert ghost.any() and np.abs(ge["f"][ghost]).max() > 0
    if precision == PRECISION_FP64:
        for k in ("f", "ep"):
            np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    else:
        mx, rms = rel_err(ge["f"], oe["f"])
        assert rms <= 1e-5, (mx, rms)


def _with_bricks(monkeypatch, value):
    if value is None:
        monkeypatch.delenv("XNB_BRICKS", raising=False)
    else:
        monkeypatch.setenv("XNB_BRICKS", value)


@pytest.mark.parametrize("periodic,min_cell,steps", [((1, 1, 0), RCUT + SKIN, 0), ((1, 1, 1), 0.5 * (RCUT + SKIN) + 0.05, 20)])
def test_ghosts_open_axis_and_two_layers(periodic, min_cell, steps):
    """Ghost cells along a non-periodic axis stay empty; with cells thinner than the ghost
    distance two ghost layers are filled.  Layout with ghost images and FP64 forces equal the
    oracle's, also after a run with rebuilds (two-layer case)."""
    r, v, ids, L = fcc_inputs(7, temperature=0.72, seed=31)
    o, b, dom = make_pair(r, v, ids, L, min_cell=min_cell, periodic=periodic)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    assert_same_layout(o.export(ghosts=True), b.export(ghosts=True), "inner+ghost")
    np.testing.assert_array_equal(o.export()["f"].view(np.uint64), b.export()["f"].view(np.uint64))
    if steps:
        ro = o.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, rebuild_every=5)
        rg = b.run(steps, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
        assert ro == rg
        assert_same_layout(o.export(ghosts=True), b.export(ghosts=True), "inner+ghost after run")
        np.testing.assert_array_equal(o.export()["f"].view(np.uint64), b.export()["f"].view(np.uint64))


@pytest.mark.parametrize("bricks,fuse", [("0", "1"), ("1,1,1", "1"), ("2,2,1", "1"), ("4,2,2", "1"), ("3,2,2", "1"),
                                         ("4,4,2", "1"), ("4,4,2", "0")])
def test_list_paths_bitwise(monkeypatch, bricks, fuse):
    """Generic (per-slot) and brick-tiled lists of several shapes give the oracle's streams
    and the oracle's FP64 forces bit for bit; the run matches with and without the next
    step's first half fused into the force kernel."""
    _with_bricks(monkeypatch, bricks)
    monkeypatch.setenv("XNB_FUSE_PUSH", fuse)
    r, v, ids, L = fcc_inputs(7, temperature=1.2, jitter=0.1, seed=21)
    o, b, dom = make_pair(r, v, ids, L)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    dims, gl, nc = b.grid_info()
    gs, _ = b.streams()
    for c in range(nc):
        k, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
            np.testing.assert_array_equal(gs[c], o.stream(c), err_msg=f"cell {c}")
    oe, ge = o.export(), b.export()
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))
    ro = o.run(30, 0.005, 1.0, 1.0, RCUT, SKIN, rebuild_every=0)
    rg = b.run(30, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=0)
    assert ro == rg
    oe, ge = o.export(), b.export()
    assert_same_layout(oe, ge, "after run")
    np.testing.assert_array_equal(oe["v"].view(np.uint64), ge["v"].view(np.uint64))


@pytest.mark.parametrize("opts", [dict(), dict(build_particle_offset=0), dict(dual_particle_offset=1),
                                  dict(random_access=1), dict(half_symmetric=1), dict(skip_ghosts=1)])
def test_stream_options_all_cells(opts):
    """Every ChunkNeighborsConfig option, ghost cells included (all-cells generic build)."""
    r, v, ids, L = fcc_inputs(6, jitter=0.05, seed=4)
    o, b, dom = make_pair(r, v, ids, L)
    for x in (o, b):
        x.move_particles()
        x.backup_r()
        x.ghost_update_all()
    o.chunk_neighbors(RCUT + SKIN, **opts)
    b.chunk_neighbors(RCUT + SKIN, all_cells=True, **opts)
    gs, gmax = b.streams()
    os_ = o.streams()
    assert len(gs) == len(os_)
    for c in range(len(gs)):
        np.testing.assert_array_equal(gs[c], os_[c], err_msg=f"cell {c} {opts}")
    assert gmax == o.max_neighbors()


@pytest.mark.parametrize("opts", [dict(half_symmetric=1), dict(skip_ghosts=1), dict(dual_particle_offset=1)])
def test_stream_options_brick_path(opts):
    r, v, ids, L = fcc_inputs(6, jitter=0.05, seed=5)
    o, b, dom = make_pair(r, v, ids, L)
    for x in (o, b):
        x.move_particles()
        x.backup_r()
        x.ghost_update_all()
    o.chunk_neighbors(RCUT + SKIN, **opts)
    b.chunk_neighbors(RCUT + SKIN, **opts)
    dims, gl, nc = b.grid_info()
    gs, _ = b.streams()
    for c in range(nc):
        k, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
            np.testing.assert_array_equal(gs[c], o.stream(c), err_msg=f"cell {c} {opts}")


def test_stream_fixture_gpu():
    import json
    import os
    from exanbody_amd import Box
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "stream_fixture.json")))
    d = fx["domain"]
    for name, case in fx["cases"].items():
        b = Box(d["origin"], d["extent"], d["grid_dims"], d["cell_size"])
        b.nbh_dist(2.5, 0.3)
        b.set_particles(np.array(fx["positions"]))
        b.move_particles()
        b.ghost_update_all()
        b.chunk_neighbors(fx["nbh_dist"], **case["config"])
        gs, _ = b.streams()
        for c, want in case["streams"].items():
            assert gs[int(c)].tolist() == want, (name, c)
        b.close()


def test_fp64_sqrt_recip_sequences_are_ieee():
    """The FP64 force kernels compute r = sqrt(d2) and 1.0 / r with their own normal-range
    sequences; they must equal the IEEE correctly rounded results (numpy's sqrt and division)
    bit for bit over the ranges they are used on, including near-midpoint cases."""
    from exanbody_amd import Box, cubic_domain
    dom = cubic_domain(10.0, 2.8)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    rng = np.random.default_rng(7)
    parts = [
        rng.uniform(1e-6, 6.25, 2_000_000),                              # pair d2 range (rc = 2.5)
        np.exp2(rng.uniform(-767.0, 1000.0, 500_000)),                   # whole sqrt range
        np.array([1.0, 2.0, 4.0, 6.25, 0.25, 2.0 ** -767, 2.0 ** 1000]),
    ]
    # near-midpoint squares: y^2 rounded, and its neighbours
    y = rng.uniform(0.5, 2.5, 200_000)
    sq = y * y
    parts += [sq, np.nextafter(sq, 0), np.nextafter(sq, 10)]
    x = np.concatenate(parts)
    s, _ = b.fp64_core(x)
    np.testing.assert_array_equal(s.view(np.uint64), np.sqrt(x).view(np.uint64))
    # reciprocal over its range and on the roots the kernel feeds it
    y = np.concatenate([rng.uniform(1e-3, 2.5, 2_000_000), np.exp2(rng.uniform(-383.0, 383.0, 500_000)),
                        np.sqrt(rng.uniform(1e-6, 6.25, 1_000_000)), 1.0 + np.arange(1, 100_000) * 2.0 ** -52,
                        2.0 - np.arange(1, 100_000) * 2.0 ** -52])
    _, rc = b.fp64_core(y)
    np.testing.assert_array_equal(rc.view(np.uint64), (1.0 / y).view(np.uint64))
    b.close()


@pytest.mark.parametrize("precision,rho", [(PRECISION_FP64, 0.8442), (PRECISION_MIXED, 0.8442), (PRECISION_FP64, 1.6)])
def test_snapshot_restart_is_bitwise(precision, rho):
    """run(12); snapshot; restore into a fresh box; run(12) == run(24) without interruption
    (rebuild every 6: the restart falls on a rebuild step boundary, so the schedule is kept).
    rho 1.6: cells of about 62 atoms, restored in their sub-cell order (rebuild_amr)."""
    from exanbody_amd import Box
    if rho > 1.0:
        r, v, ids, L = fcc_inputs(10, rho=rho, temperature=0.72, jitter=0.02, seed=41)
    else:
        r, v, ids, L = fcc_inputs(7, temperature=1.0, jitter=0.05, seed=41)
    _, a, dom = make_pair(r, v, ids, L)
    _, b, _ = make_pair(r, v, ids, L)
    for x in (a, b):
        x.init(1.0, 1.0, RCUT, SKIN, precision)
    a.run(24, 0.005, 1.0, 1.0, RCUT, SKIN, precision=precision, rebuild_every=6)
    b.run(12, 0.005, 1.0, 1.0, RCUT, SKIN, precision=precision, rebuild_every=6)
    snap = b.snapshot()
    b.close()
    c = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    c.restore(snap)
    c.run(12, 0.005, 1.0, 1.0, RCUT, SKIN, precision=precision, rebuild_every=6)
    ea, ec = a.export(), c.export()
    assert_same_layout(ea, ec, "restart")
    for k in ("v", "f", "ep"):
        np.testing.assert_array_equal(ea[k].view(np.uint64), ec[k].view(np.uint64), err_msg=k)


def test_autotune_keeps_results():
    """The autotuner only changes the brick shape: forces, energies and a later run are bitwise
    those of an untuned box; a second box of the same class reuses the cached shape."""
    r, v, ids, L = fcc_inputs(8, temperature=1.0, jitter=0.05, seed=51)
    _, a, _ = make_pair(r, v, ids, L)
    _, b, _ = make_pair(r, v, ids, L)
    for x in (a, b):
        x.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    f0 = a.export()
    dims = a.autotune(1.0, 1.0, RCUT, PRECISION_FP64)
    assert dims in {(4, 4, 2), (4, 2, 2), (2, 4, 2), (4, 4, 1), (2, 2, 2), (4, 2, 1)}
    f1 = a.export()
    for k in ("r", "f", "ep"):
        np.testing.assert_array_equal(f0[k].view(np.uint64), f1[k].view(np.uint64), err_msg=k)
    assert b.autotune(1.0, 1.0, RCUT, PRECISION_FP64) == dims
    a.run(15, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
    b.run(15, 0.005, 1.0, 1.0, RCUT, SKIN, precision=PRECISION_FP64, rebuild_every=5)
    ea, eb = a.export(), b.export()
    assert_same_layout(ea, eb, "tuned")
    np.testing.assert_array_equal(ea["v"].view(np.uint64), eb["v"].view(np.uint64))
