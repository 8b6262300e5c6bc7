"""The neighbour build's FP32 prefilter: candidates whose FP32 squared distance lies inside an
error band around nbh_dist^2 (or near 0) are decided by the FP64 test, so the lists are those
of the FP64 build.

The GPU cases put many candidates exactly on the cutoff sphere (a simple cubic lattice of
spacing nbh_dist / 2, where FP64 rounding alone decides each pair) and near or exactly on top
of each other (FP64 keeps a pair at d2 > 0 and drops exact duplicates), and compare streams
and FP64 forces with the oracle bit for bit.  The CPU case checks the band's error bound
against the FP32 arithmetic it covers.
"""
import numpy as np
import pytest

from exanbody_amd import PRECISION_FP64, PRECISION_MIXED, Box, cubic_domain
from tests.helpers import RCUT, SKIN, fcc_inputs, make_pair


def _inner_streams_equal(o, b):
    dims, gl, nc = b.grid_info()
    gs, _ = b.streams()
    checked = 0
    for c in range(nc):
        k, rem = divmod(c, dims[0] * dims[1])
        j, i = divmod(rem, dims[0])
        if all(gl <= x < d - gl for x, d in zip((i, j, k), dims)):
            np.testing.assert_array_equal(gs[c], o.stream(c), err_msg=f"cell {c}")
            checked += 1
    return checked


@pytest.mark.gpu
def test_candidates_on_the_cutoff_sphere():
    nbh = RCUT + SKIN
    a = nbh / 2.0
    n = 12
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    r = g * a
    L = n * a
    v = np.zeros_like(r)
    ids = np.arange(len(r), dtype=np.uint64)
    o, b, dom = make_pair(r, v, ids, L)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    # the (2, 0, 0) lattice neighbours sit at nbh_dist up to rounding: some are in, some out
    d2 = (2 * a) ** 2
    assert abs(d2 - nbh * nbh) < 1e-12
    assert _inner_streams_equal(o, b) > 0
    oe, ge = o.export(), b.export()
    np.testing.assert_array_equal(oe["f"].view(np.uint64), ge["f"].view(np.uint64))


@pytest.mark.gpu
def test_coincident_and_near_coincident_atoms():
    r, v, ids, L = fcc_inputs(6, jitter=0.05, seed=8)
    n = len(r)
    near = r[0:10] + np.array([1e-9, 0.0, 0.0])       # FP64 d2 = 1e-18 > 0: kept both ways
    dup = r[10:20].copy()                               # d2 == 0: never kept
    close = r[20:30] + 1e-7                             # FP32 d2 near 0: decided in FP64
    r2 = np.concatenate([r, near, dup, close])
    v2 = np.concatenate([v, v[0:10], v[10:20], v[20:30]])
    ids2 = np.arange(len(r2), dtype=np.uint64)
    o, b, dom = make_pair(r2, v2, ids2, L)
    o.init(1.0, 1.0, RCUT, SKIN)
    b.init(1.0, 1.0, RCUT, SKIN, PRECISION_FP64)
    assert _inner_streams_equal(o, b) > 0
    assert len(r2) == n + 30


def _run_fcc(flag, monkeypatch, r, v, ids, L, precision):
    monkeypatch.setenv("XNB_NBH_PREFILTER", flag)
    dom = cubic_domain(L, RCUT + SKIN)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(RCUT, SKIN)
    b.set_particles(r, v, ids)
    b.init(1.0, 1.0, RCUT, SKIN, precision)
    res = b.run(40, 0.005, 1.0, 1.0, RCUT, SKIN, precision=precision, rebuild_every=10)
    e = b.export()
    gs, gmax = b.streams()
    b.close()
    return res, e, gs, gmax


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_prefilter_on_off_same_run_at_scale(monkeypatch, precision):
    """256,000 atoms of a hot, disordered FCC start, 40 steps with a rebuild every 10: the
    prefiltered and the all-FP64 builds give the same lists, so the runs agree bit for bit."""
    r, v, ids, L = fcc_inputs(40, temperature=1.5, jitter=0.08, seed=12)
    ra, ea, gsa, gma = _run_fcc("1", monkeypatch, r, v, ids, L, precision)
    rb, eb, gsb, gmb = _run_fcc("0", monkeypatch, r, v, ids, L, precision)
    assert ra == rb and ra["rebuilds"] == 4
    assert gma == gmb
    assert len(gsa) == len(gsb)
    for c in range(len(gsa)):
        np.testing.assert_array_equal(gsa[c], gsb[c], err_msg=f"cell {c}")
    for k in ("r", "v", "f", "id"):
        np.testing.assert_array_equal(ea[k].view(np.uint64) if ea[k].dtype == np.float64 else ea[k],
                                      eb[k].view(np.uint64) if eb[k].dtype == np.float64 else eb[k], err_msg=k)


@pytest.mark.gpu
def test_prefilter_on_off_same_multimat_slab(monkeypatch):
    """Per-pair list radii on a two-species slab with vacuum and droplets (150,000 atoms),
    30 FP64 steps with a rebuild every 10: the same trajectory with and without the prefilter."""
    from exanbody_amd.workloads import lorentz_berthelot, slab_droplets
    eps, sig, rc = lorentz_berthelot()
    r, v, ids, types, dom = slab_droplets(n_atoms=150000, ncell=48, n_droplets=8, seed=4)
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("XNB_NBH_PREFILTER", flag)
        b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
        b.nbh_dist(float(rc.max()), SKIN)
        b.set_particles(r, v, ids, types)
        b.init(1.0, 1.0, float(rc.max()), SKIN, PRECISION_FP64)
        res = b.run_multimat(30, 0.004, 2, eps, sig, rc, SKIN, precision=PRECISION_FP64, rebuild_every=10)
        outs.append((res, b.export()))
        b.close()
    (ra, ea), (rb, eb) = outs
    assert ra == rb and ra["rebuilds"] == 3
    for k in ("r", "v", "f"):
        np.testing.assert_array_equal(ea[k].view(np.uint64), eb[k].view(np.uint64), err_msg=k)
    np.testing.assert_array_equal(ea["id"], eb["id"])


def _band(max2, cmax):
    """The kernel's band half-width H and near-zero bound (xnb_brick.hip, PREF branch)."""
    u = 2.0 ** -24
    C = cmax * (1.0 + 2.0 ** -20)
    E = 4.01 * u * C
    R = np.sqrt(max2)
    H = 2.0 * (2.0 * np.sqrt(3.0) * R * E + 3.0 * E * E + 4.0 * u * max2)
    return H, 16.0 * E * E


def _fma32(a, b, c):
    """fmaf on float32 arrays: the product is exact in float64, one rounding of the sum to
    float32 (a float64 sum can round once more first; that changes the result by far less
    than the bound tested here)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _dot_band(T2, cmax):
    """The per-pair-radius build's superset band H2 on t + |a_f|^2 (xnb_brick.hip, PREF), and
    its distance-independent part (the rounding of the norms and of the three fmas)."""
    u = 2.0 ** -24
    C = cmax * (1.0 + 2.0 ** -20)
    E2 = 2.01 * u * C
    H2 = 2.0 * (45.0 * u * C * C + 2.0 * np.sqrt(3.0) * np.sqrt(T2) * E2 + 3.0 * E2 * E2)
    return H2, 45.0 * u * C * C + 3.0 * E2 * E2


@pytest.mark.parametrize("cmax", [2.0, 9.8, 40.0])
def test_dot_form_band_bounds_fp32_error(cmax):
    """The per-pair-radius build's superset test evaluates t = |b_f|^2 - 2 a_f.b_f with three
    fmas on the staged FP32 coordinates and their staged FP32 squared norms; t + |a_f|^2 stays
    within half its band of the FP64 d2 near the cutoff, and within the band's
    distance-independent part plus its distance term for near-coincident pairs."""
    rng = np.random.default_rng(5)
    nbh = RCUT + SKIN
    max2 = nbh * nbh
    H2, H0 = _dot_band(max2, cmax)
    m = 400000
    centre = rng.uniform(-1e3, 1e3, 3)
    a = rng.uniform(-cmax, cmax, (m, 3))
    d = rng.normal(size=(m, 3))
    scale = np.where(rng.random(m) < 0.1, rng.uniform(0.0, 0.05, m), nbh * rng.uniform(0.98, 1.02, m))
    d *= (scale / np.linalg.norm(d, axis=1))[:, None]
    b = np.clip(a + d, -cmax, cmax)
    pa, pb = a + centre, b + centre
    ex = pa - pb
    d2 = (ex[:, 0] * ex[:, 0] + ex[:, 1] * ex[:, 1]) + ex[:, 2] * ex[:, 2]
    fa = (pa - centre).astype(np.float32)
    fb = (pb - centre).astype(np.float32)

    def norm2(f):
        return _fma32(f[:, 0], f[:, 0], _fma32(f[:, 1], f[:, 1], f[:, 2] * f[:, 2]))

    wa, wb = norm2(fa), norm2(fb)
    m2 = np.float32(-2.0)
    t = _fma32(m2 * fa[:, 0], fb[:, 0], _fma32(m2 * fa[:, 1], fb[:, 1], _fma32(m2 * fa[:, 2], fb[:, 2], wb)))
    err = np.abs(t.astype(np.float64) + wa.astype(np.float64) - d2)
    near = d2 > 0.5 * max2
    assert err[near].max() <= 0.5 * H2
    close = d2 < 0.01
    assert err[close].max() <= H0 + 2.0 * np.sqrt(3.0) * 0.1 * 2.01 * 2.0 ** -24 * cmax * 2.0
    # the band stays a small fraction of the candidates at the bench's brick size
    if cmax == 9.8:
        assert H2 < 2e-3
        frac = np.mean(np.abs(t[near].astype(np.float64) + wa[near] - max2) <= H2)
        assert frac < 0.05


@pytest.mark.parametrize("cmax", [2.0, 9.8, 40.0])
def test_band_bounds_fp32_error(cmax):
    """For staged coordinates up to cmax, |d2f - d2| stays below half the band (the factor 2
    the kernel keeps in H) over pairs near the cutoff, and the near-zero bound separates FP32
    zeros from FP64 distances that could round to 0."""
    rng = np.random.default_rng(3)
    nbh = RCUT + SKIN
    max2 = nbh * nbh
    H, z = _band(max2, cmax)
    m = 400000
    centre = rng.uniform(-1e3, 1e3, 3)
    a = rng.uniform(-cmax, cmax, (m, 3))
    d = rng.normal(size=(m, 3))
    d *= (nbh * rng.uniform(0.98, 1.02, m) / np.linalg.norm(d, axis=1))[:, None]
    b = np.clip(a + d, -cmax, cmax)
    pa, pb = a + centre, b + centre
    ex = pa - pb
    d2 = (ex[:, 0] * ex[:, 0] + ex[:, 1] * ex[:, 1]) + ex[:, 2] * ex[:, 2]
    fa = (pa - centre).astype(np.float32)
    fb = (pb - centre).astype(np.float32)
    df = fa - fb
    d2f = _fma32(df[:, 2], df[:, 2], _fma32(df[:, 1], df[:, 1], df[:, 0] * df[:, 0]))
    err = np.abs(d2f.astype(np.float64) - d2)
    assert err.max() <= 0.5 * H
    assert H < 1e-4 * max2 * (cmax / 2.0)
    # pairs the band cannot decide are rare at the bench's brick size
    if cmax == 9.8:
        frac = np.mean(np.abs(d2f - max2) <= H)
        assert frac < 0.01
    assert z < 1e-6
