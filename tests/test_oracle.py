"""CPU checks that pin the oracle (oracle/xnb_oracle.c) independently of its own formulas:
closed forms of the LJ pair, FCC lattice shell sums and symmetry, Newton's third law,
F = -grad E by finite differences, brute-force neighbour sets, a hand-derived stream
fixture, ghost images by brute force, backup_r encoding bounds and move invariants.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, encode_double_u32, restore_u32_double, sub_grid_size

HERE = os.path.dirname(os.path.abspath(__file__))
RHO = 0.8442


def fcc(n, a, jitter=0.0, seed=0):
    base = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    r = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    if jitter:
        r = r + np.random.default_rng(seed).uniform(-jitter, jitter, r.shape)
    L = n * a
    return np.mod(r, L), L


def box(L, r, min_cell=2.8, ghost=2.8, v=None, types=None):
    dims = int(L // min_cell)
    o = Oracle([0, 0, 0], [L, L, L], [dims] * 3, L / dims, ghost_dist=ghost)
    o.set_particles(r, v, None, types)
    return o


def decode_pairs(o):
    """All (id_a, id_b) pairs of inner cells decoded from the CS=1 streams."""
    ex = o.export(ghosts=True)
    dims = o.dims
    cells = ex["cell"]
    starts = {}
    for k, c in enumerate(cells):
        starts.setdefault(int(c), k)
    gl = o.ghost_layers
    pairs = set()
    counts = o.cell_counts()
    for ca in range(o.num_cells):
        k, rem = divmod(ca, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if not all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
            continue
        st = o.stream(ca)
        n = counts[ca]
        if n == 0:
            assert st.size == 0
            continue
        assert st[0] <= 2 and st[1] == 0
        nt = int(st[0])
        off = st[:(n * nt + 1) * 2].view(np.uint32)
        base = (n * nt + 1) * 2
        for pa in range(n):
            pos = base + int(off[pa]) - nt
            ncg = int(st[pos]); pos += 1
            last = None
            for _ in range(ncg):
                enc = int(st[pos]); nch = int(st[pos + 1]); pos += 2
                assert enc >= 1057
                ri, rj, rk = (enc & 31) - 16, ((enc >> 5) & 31) - 16, ((enc >> 10) & 31) - 16
                cb = ca + (rk * dims[1] + rj) * dims[0] + ri
                for _ in range(nch):
                    pb = int(st[pos]); pos += 1
                    key = (enc, pb)
                    assert last is None or key > last, "stream not sorted by (cell_enc, p_b)"
                    last = key
                    assert not (cb == ca and pb == pa)
                    pairs.add((int(ex["id"][starts[ca] + pa]), int(ex["id"][starts[cb] + pb])))
            assert pos == base + int(off[pa + 1]) - nt
    return pairs


# ---------------------------------------------------------------- LJ closed forms

@pytest.mark.parametrize("r", [0.9, 2 ** (1 / 6), 1.0, 1.5, 2.4999])
def test_two_atom_closed_form(r):
    L = 12.0
    o = box(L, np.array([[3.0, 3.0, 3.0], [3.0 + r, 3.0, 3.0]]))
    o.init(1.0, 1.0, 2.5, 0.3)
    e = o.export()
    f = e["f"][np.argsort(e["id"])]
    ep = e["ep"][np.argsort(e["id"])]
    # textbook LJ: E = 4(r^-12 - r^-6), F on atom 0 along +x = -dE/dr * (-1) ... sign fixed by physics:
    # repulsive (r < 2^1/6) pushes atom 0 towards -x, attractive pulls it towards +x.
    E = 4 * (r ** -12 - r ** -6)
    Fmag = 24 * (2 * r ** -13 - r ** -7)  # = -dE/dr, >0 means repulsive
    assert f[0, 0] == pytest.approx(-Fmag, rel=1e-12, abs=1e-14)
    assert f[1, 0] == pytest.approx(Fmag, rel=1e-12, abs=1e-14)
    assert abs(f[0, 1]) + abs(f[0, 2]) == 0.0
    assert ep[0] == pytest.approx(E, rel=1e-12, abs=1e-14) and ep[1] == ep[0]
    if r == 1.0:
        assert abs(ep[0]) < 1e-15            # E(sigma) = 0
    if r == 2 ** (1 / 6):
        assert ep[0] == pytest.approx(-1.0, rel=1e-14)  # well depth
        assert abs(f[0, 0]) < 1e-13          # force vanishes at the minimum


def test_cutoff_boundary():
    L = 12.0
    rc = 2.5
    inside = np.nextafter(rc, 0)
    outside = np.nextafter(rc, 10)
    for d, expect in ((inside, True), (outside, False)):
        # x0 = 0.5: 0.5 + d is exact at this magnitude, so the pair distance is d itself
        o = box(L, np.array([[0.5, 3.0, 3.0], [0.5 + d, 3.0, 3.0]]))
        o.init(1.0, 1.0, rc, 0.3)
        has = np.abs(o.export()["f"]).max() > 0
        assert (0.5 + d) - 0.5 == d
        assert has == (d * d <= rc * rc)
        assert has == expect


def test_fcc_shells_counts_zero_force_and_lattice_sum():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(8, a)
    o = box(L, r)
    o.init(1.0, 1.0, 2.5, 0.3)
    assert o.max_neighbors() == 78
    pairs = decode_pairs(o)
    assert len(pairs) == 78 * len(r)
    e = o.export()
    assert np.abs(e["f"]).max() < 1e-12
    shells = [(a / math.sqrt(2), 12), (a, 6), (a * math.sqrt(1.5), 24), (a * math.sqrt(2), 12)]
    lattice_sum = sum(n * 4 * (s ** -12 - s ** -6) for s, n in shells)
    np.testing.assert_allclose(e["ep"], lattice_sum, rtol=1e-13)


def test_newton_third_law_and_gradient():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(5, a, jitter=0.05, seed=3)
    o = box(L, r)
    o.init(1.0, 1.0, 2.5, 0.3)
    e = o.export()
    F = e["f"]
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).max()
    # F = -dPE/dx by central differences on one atom (no pair near the cutoff for this h)
    k = int(e["id"][7])
    h = 1e-6
    order = np.argsort(e["id"])
    fk = F[order][k]
    for ax in range(3):
        pe = []
        for s in (+1, -1):
            rr = r.copy()
            rr[k, ax] += s * h
            q = box(L, rr)
            q.init(1.0, 1.0, 2.5, 0.3)
            pe.append(q.energies()[1])
        g = (pe[0] - pe[1]) / (2 * h)
        assert -g == pytest.approx(fk[ax], rel=1e-5, abs=1e-6)


# ------------------------------------------------------------- neighbour lists

def brute_pairs(r, L, cut):
    d = r[:, None, :] - r[None, :, :]
    d -= L * np.round(d / L)
    d2 = (d * d).sum(-1)
    i, j = np.nonzero((d2 <= cut * cut) & (d2 > 0))
    return set(zip(i.tolist(), j.tolist()))


def test_neighbour_set_brute_force():
    a = (4 / 0.7) ** (1 / 3)
    r, L = fcc(5, a, jitter=0.3, seed=5)
    o = box(L, r)
    o.init(1.0, 1.0, 2.5, 0.3)
    assert decode_pairs(o) == brute_pairs(r, L, 2.8)


def test_stream_fixture():
    fx = json.load(open(os.path.join(HERE, "golden", "stream_fixture.json")))
    d = fx["domain"]
    for name, case in fx["cases"].items():
        o = Oracle(d["origin"], d["extent"], d["grid_dims"], d["cell_size"], ghost_dist=fx["nbh_dist"])
        o.set_particles(np.array(fx["positions"]))
        o.move_particles()
        o.ghost_update_all()
        cfg = dict(case["config"])
        o.chunk_neighbors(fx["nbh_dist"], **cfg)
        dims = o.dims
        for c in range(o.num_cells):
            k, rem = divmod(c, dims[0] * dims[1])
            j, i = divmod(rem, dims[0])
            if not all(1 <= x < n - 1 for x, n in zip((i, j, k), dims)):
                continue
            want = case["streams"].get(str(c), [])
            assert o.stream(c).tolist() == want, (name, c)


# ----------------------------------------------------------------------- ghosts

def test_ghost_images_brute_force():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(4, a, jitter=0.05, seed=9)
    o = box(L, r)
    o.move_particles()
    o.ghost_update_all()
    g = o.export(ghosts=True)
    inner = o.num_particles()
    ghosts = set()
    counts = o.cell_counts()
    dims = o.dims
    gl = o.ghost_layers
    for k in range(len(g["id"])):
        c = int(g["cell"][k])
        kk, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if not all(gl <= x < n - gl for x, n in zip((i, j, kk), dims)):
            ghosts.add((int(g["id"][k]), tuple(np.round(g["r"][k], 12))))
    assert len(ghosts) == o.num_ghost_particles() == len(g["id"]) - inner
    want = set()
    lo, hi = -2.8, L + 2.8
    for p in range(len(r)):
        for s in np.ndindex(3, 3, 3):
            sh = np.array(s) - 1
            if not sh.any():
                continue
            q = r[p] + sh * L
            if np.all(q >= lo) and np.all(q <= hi):
                want.add((p, tuple(np.round(q, 12))))
    assert ghosts == want
    assert counts.sum() == inner + o.num_ghost_particles()


# --------------------------------------------------------------- backup / move

def test_backup_encoding_bounds():
    rng = np.random.default_rng(1)
    cs = 2.828793585486328
    for o in (0.0, 17.3, 211.5):
        for x in o + rng.uniform(0, cs, 200):
            u = encode_double_u32(x, o, cs)
            y = restore_u32_double(u, o, cs)
            assert abs(y - x) <= cs * 2.0 ** -32 * (1 + 1e-9) + 4 * np.spacing(max(abs(x), 1.0))


def test_trigger_threshold():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(4, a, jitter=0.01, seed=2)
    v = np.zeros_like(r)
    v[5] = [0.14 / 0.01, 0, 0]  # moves 0.14 in one push of dt=0.01
    o = box(L, r, v=v)
    o.init(0.0, 1.0, 2.5, 0.3)  # epsilon 0: no forces, free flight
    o.push_f_v_r(0.01, 1.0)
    assert o.particle_displ_over(0.15) == 0
    o.push_f_v_r(0.01, 1.0)
    assert o.particle_displ_over(0.15) == 1


def test_move_invariants_and_free_flight():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(4, a, jitter=0.01, seed=4)
    rng = np.random.default_rng(0)
    v = rng.normal(0, 20.0, r.shape)
    o = box(L, r, v=v)
    o.init(0.0, 1.0, 2.5, 0.3)
    o.push_f_v_r(0.1, 1.0)  # free flight: f = 0, r <- r + v dt
    o.move_particles()
    e = o.export()
    assert sorted(e["id"].tolist()) == list(range(len(r)))
    order = np.argsort(e["id"])
    expect = np.mod(r + v * 0.1, L)
    got = e["r"][order]
    dd = got - expect
    dd -= L * np.round(dd / L)
    assert np.abs(dd).max() < 1e-12
    assert np.all(got >= 0) and np.all(got < L)
    dims = o.dims
    cs = L / (dims[0] - 2)
    cells = e["cell"][order]
    for p in range(len(r)):
        k, rem = divmod(int(cells[p]), dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        for ax, idx in enumerate((i, j, k)):
            lo = (idx - 1) * cs
            assert lo - 1e-12 <= got[p, ax] <= lo + cs + 1e-12


def test_lone_atom_and_empty_runs():
    """A single atom whose images lie beyond the cutoff flies ballistically through a run with
    rebuilds (velocity unchanged, position r0 + n dt v wrapped into the box); an empty box
    runs and reports zero energies."""
    L, steps, dt = 9.0, 40, 0.005
    r0 = np.array([[4.5, 0.01, 8.99]])
    v0 = np.array([[0.3, -0.7, 0.4]])
    o = Oracle([0, 0, 0], [L] * 3, [3] * 3, L / 3, ghost_dist=2.8)
    o.set_particles(r0, v0)
    o.init(1.0, 1.0, 2.5, 0.3)
    res = o.run(steps, dt, 1.0, 1.0, 2.5, 0.3, rebuild_every=5)
    assert res["rebuilds"] == steps // 5
    e = o.export()
    np.testing.assert_array_equal(e["v"], v0)
    assert not e["f"].any() and not e["ep"].any()
    dd = e["r"] - np.mod(r0 + steps * dt * v0, L)
    assert np.abs(dd).max() < 1e-12
    assert np.all(e["r"] >= 0) and np.all(e["r"] < L)
    o = Oracle([0, 0, 0], [L] * 3, [3] * 3, L / 3, ghost_dist=2.8)
    o.set_particles(np.zeros((0, 3)))
    o.init(1.0, 1.0, 2.5, 0.3)
    o.run(steps, dt, 1.0, 1.0, 2.5, 0.3, rebuild_every=5)
    assert o.num_particles() == 0 and o.energies() == (0.0, 0.0)


def _sorted_state(o):
    e = o.export()
    k = np.argsort(e["id"])
    return e["r"][k], e["v"][k]


def test_verlet_time_reversibility():
    """Velocity Verlet is time reversible: 50 steps, velocities negated, 50 more steps return
    every atom to its start (round-off only).  A position update r += v dt + c f dt^2 returns
    with an error (c - 1/2)(f0 + f1) dt^2 unless c = 1/2, and a missing half kick also fails."""
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(6, a, jitter=0.05, seed=1)
    v = np.random.default_rng(2).normal(0, 1.0, r.shape)
    v -= v.mean(0)
    o = box(L, r, v=v)
    o.init(1.0, 1.0, 2.5, 0.3)
    o.run(50, 0.005, 1.0, 1.0, 2.5, 0.3)
    r1, v1 = _sorted_state(o)
    moved = r1 - r
    moved -= L * np.round(moved / L)
    assert np.abs(moved).max() > 0.3
    o = box(L, r1, v=-v1)
    o.init(1.0, 1.0, 2.5, 0.3)
    o.run(50, 0.005, 1.0, 1.0, 2.5, 0.3)
    r2, v2 = _sorted_state(o)
    dr = r2 - r
    dr -= L * np.round(dr / L)
    assert np.abs(dr).max() < 1e-10
    assert np.abs(v2 + v).max() < 1e-9


def test_verlet_energy_error_is_second_order():
    """Total energy KE + PE of a cold crystal (no pair crosses the cutoff, whose unshifted
    potential would make E jump) over a fixed time of 0.4: the largest deviation from E(0)
    falls by 4 when dt halves (measured ratios 4.006 and 4.000), as for a second-order
    symplectic integrator; a first-order update gives 2."""
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(6, a, jitter=0.003, seed=3)
    v = np.random.default_rng(4).normal(0, np.sqrt(0.005), r.shape)
    v -= v.mean(0)
    dev = []
    for dt in (0.004, 0.002, 0.001):
        o = box(L, r, v=v)
        o.init(1.0, 1.0, 2.5, 0.3)
        e0 = sum(o.energies())
        worst = 0.0
        for _ in range(40):
            o.run(int(round(0.01 / dt)), dt, 1.0, 1.0, 2.5, 0.3, rebuild_every=10)
            worst = max(worst, abs(sum(o.energies()) - e0))
        dev.append(worst)
        assert worst < 1e-6 * abs(e0) * (dt / 0.004) ** 2 * 1.5
    assert 3.6 < dev[0] / dev[1] < 4.4 and 3.6 < dev[1] / dev[2] < 4.4, dev


def test_sub_grid_size_closed_form():
    """sub_grid_size(n, 6.5) = floor(cbrt(n / 6.5)), 1 below 2 and at most 16, equals the
    integer form: the largest s with 13 s^3 <= 2 n (for n >= 52), for every n up to 60,000."""
    for n in range(0, 60000):
        if n == 0:
            want = 0
        elif 13 * 8 > 2 * n:
            want = 1
        else:
            k = 2
            while 13 * (k + 1) ** 3 <= 2 * n:
                k += 1
            want = min(k, 16)
        assert sub_grid_size(n) == want, n


def test_rebuild_amr_stable_sub_cell_order():
    """rebuild_amr on a dense fluid (about 62 particles per cell, sub-grid side 2): in each
    cell of 52 or more particles, the new order is a permutation of the old one, sorted by
    sub-cell index (k s + j) s + i of the clamped cell-relative coordinates, and particles of
    one sub-cell keep their previous relative order; sparser cells are unchanged."""
    a = (4 / 1.6) ** (1 / 3)
    r, L = fcc(10, a, jitter=0.05, seed=5)
    o = box(L, r)
    before = o.export()
    o.rebuild_amr()
    after = o.export()
    dims = o.dims
    gl = o.ghost_layers
    cs = L / (dims[0] - 2 * gl)
    counts = o.cell_counts()
    dense = 0
    for c in np.unique(before["cell"]):
        sel_b, sel_a = before["cell"] == c, after["cell"] == c
        ib, ia = before["id"][sel_b], after["id"][sel_a]
        assert sorted(ib.tolist()) == sorted(ia.tolist())
        n = int(counts[c])
        sg = sub_grid_size(n)
        if sg <= 1:
            np.testing.assert_array_equal(ib, ia)
            continue
        dense += 1
        k, rem = divmod(int(c), dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        lo = np.array([(i - gl) * cs, (j - gl) * cs, (k - gl) * cs])
        q = np.clip(np.trunc((after["r"][sel_a] - lo) / cs * sg), 0, sg - 1).astype(np.int64)
        key = (q[:, 2] * sg + q[:, 1]) * sg + q[:, 0]
        assert np.all(np.diff(key) >= 0)
        rank_before = {int(x): t for t, x in enumerate(ib)}
        for kk in np.unique(key):
            pos = [rank_before[int(x)] for x in ia[key == kk]]
            assert pos == sorted(pos)
        assert len(np.unique(key)) > 1
    assert dense > 0


def test_multimat_single_type_equals_singlemat():
    a = (4 / RHO) ** (1 / 3)
    r, L = fcc(4, a, jitter=0.05, seed=8)
    o = box(L, r)
    o.init(1.0, 1.0, 2.5, 0.3)
    f1 = o.export()["f"].copy()
    o.zero_force()
    o.lj_force_multimat(1, [1.0], [1.0], [2.5])
    f2 = o.export()["f"]
    np.testing.assert_array_equal(f1.view(np.uint64), f2.view(np.uint64))


def test_multimat_lorentz_berthelot_pairs():
    # two atoms of types (0,1): pair id 1 parameters must be used, per-pair cutoff respected
    L = 14.0
    for d, inside in ((2.9, True), (3.1, False)):
        o = box(L, np.array([[3.0, 3.0, 3.0], [3.0 + d, 3.0, 3.0]]), min_cell=3.8, ghost=3.8,
                types=np.array([0, 1], np.uint8))
        o.init(0.0, 1.0, 3.5, 0.3)
        o.zero_force()
        eps = [1.0, math.sqrt(0.8), 0.8]
        sig = [1.0, 1.2, 1.4]
        o.lj_force_multimat(2, eps, sig, [2.5 * s for s in sig])
        e = o.export()
        f = e["f"][np.argsort(e["id"])]
        if inside:
            s = 1.2 / d
            Fmag = 24 * math.sqrt(0.8) * (2 * s ** 12 - s ** 6) / d
            assert f[1, 0] == pytest.approx(Fmag, rel=1e-12)
        else:
            assert np.abs(f).max() == 0.0
