"""Synthetic inputs of the heterogeneous workload: a two-species liquid slab with droplets."""
import numpy as np

from .box import cubic_domain


def lorentz_berthelot(sigma_a=1.0, sigma_b=1.4, eps_a=1.0, eps_b=0.8, rc_factor=2.5):
    """Pair tables in unique-pair-id order (0,0), (0,1), (1,1)."""
    sig = np.array([sigma_a, 0.5 * (sigma_a + sigma_b), sigma_b])
    eps = np.array([eps_a, np.sqrt(eps_a * eps_b), eps_b])
    return eps, sig, rc_factor * sig


def slab_droplets(n_atoms=4_194_304, ncell=128, n_droplets=64, radius=5.0, frac_b=0.2, temperature=0.8, jitter=0.01, seed=1,
                  rho=0.55, min_cell=3.8):
    """Periodic cubic box of ncell FCC unit cells per side at number density rho: the lower
    half along z is a liquid slab, the upper half vacuum holding n_droplets spheres of the
    given radius (seeded, non-overlapping); slab sites are dropped at random so the total is
    n_atoms.  Species B (type 1) is a random frac_b of the atoms.  Velocities Maxwell-Boltzmann
    at `temperature`, zero total momentum, exact temperature with 3N - 3 degrees of freedom.
    Returns (r, v, ids, types, domain)."""
    rng = np.random.default_rng(seed)
    a = (4.0 / rho) ** (1.0 / 3.0)
    L = ncell * a
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    half = ncell // 2
    i, j, k = np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(half), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    slab = ((cells[:, None, :] + basis[None, :, :]) * a).reshape(-1, 3)
    # droplet centres in the vacuum half, at least one radius + a lattice constant from the slab
    zlo, zhi = L / 2 + radius + a, L - radius - a
    centres = []
    while len(centres) < n_droplets:
        c = np.array([rng.uniform(0, L), rng.uniform(0, L), rng.uniform(zlo, zhi)])
        ok = True
        for d in centres:
            e = c - d
            e[:2] -= L * np.round(e[:2] / L)
            if np.dot(e, e) < (2 * radius + a) ** 2:
                ok = False
                break
        if ok:
            centres.append(c)
    m = int(np.ceil(radius / a)) + 1
    g = np.arange(-m, m + 1, dtype=np.float64)
    gi, gj, gk = np.meshgrid(g, g, g, indexing="ij")
    local = ((np.stack([gi.ravel(), gj.ravel(), gk.ravel()], axis=1)[:, None, :] + basis[None, :, :]) * a).reshape(-1, 3)
    drops = []
    for c in centres:
        base = np.floor(c / a) * a
        p = base + local
        keep = np.sum((p - c) ** 2, axis=1) <= radius * radius
        drops.append(p[keep])
    drops = np.concatenate(drops)
    n_slab = n_atoms - drops.shape[0]
    if n_slab <= 0 or n_slab > slab.shape[0]:
        raise ValueError("droplets do not fit the atom budget")
    slab = slab[np.sort(rng.choice(slab.shape[0], n_slab, replace=False))]
    r = np.concatenate([slab, drops])
    r += rng.uniform(-jitter, jitter, r.shape)
    r %= L
    n = r.shape[0]
    types = (rng.random(n) < frac_b).astype(np.uint8)
    v = rng.normal(0.0, 1.0, (n, 3))
    v -= v.mean(axis=0)
    v *= np.sqrt(temperature * (3 * n - 3) / np.sum(v * v))
    ids = np.arange(n, dtype=np.uint64)
    return r, v, ids, types, cubic_domain(L, min_cell)


def batch_job_list(n_boxes=512, steps=500, seed=5, precision=None, rebuild_every=10):
    """The batched ensemble: n_boxes periodic LJ boxes, atoms log-uniform in [500, 50000]
    (FCC-filled in lattice order, jittered +-0.01), density U[0.70, 0.90], T* U[0.5, 1.5],
    mixed precision unless given, as job dicts (serve.make_job)."""
    from . import _lib
    from .serve import make_job
    prec = _lib.PRECISION_MIXED if precision is None else precision
    rng = np.random.default_rng(seed)
    jobs = []
    for q in range(n_boxes):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
        jobs.append(make_job(n, rng.uniform(0.70, 0.90), rng.uniform(0.5, 1.5), steps, prec, seed=7000 + q,
                             rebuild_every=rebuild_every))
    return jobs
