"""Python handle on one device-resident box (an xnb_ctx), calling the C ABI by operator name.

Method names follow the reference YAML operators: move_particles, backup_r,
ghost_update_all, ghost_update_r, particle_displ_over, chunk_neighbors,
zero_particle_force, lennard_jones_force, push_f_v_r, push_f_v.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import XnbError


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def generate_fcc(ncell, rho, n_atoms=0, jitter=0.01, temperature=0.72, seed=1, velocities=True):
    """Seeded FCC lattice (host only).  Returns (r[n,3], v[n,3] or None, id[n], box_length)."""
    L = _lib.load()
    n = 4 * ncell ** 3 if n_atoms <= 0 else n_atoms
    r = np.empty((n, 3), np.float64)
    v = np.empty((n, 3), np.float64) if velocities else None
    ids = np.empty(n, np.uint64)
    box = C.c_double()
    rc = L.xnb_generate_fcc(int(ncell), float(rho), int(n), float(jitter), float(temperature), int(seed), _ptr(r),
                            _ptr(v), _ptr(ids), C.byref(box))
    if rc != 0:
        raise XnbError(rc, "xnb_generate_fcc failed")
    return r, v, ids, box.value


def cubic_domain(box_length, min_cell, periodic=(1, 1, 1), origin=0.0):
    """Domain of a cubic box with the largest cell count whose cell edge is >= min_cell."""
    dims = max(1, int(box_length // min_cell))
    while dims > 1 and box_length / dims < min_cell:
        dims -= 1
    cs = box_length / dims
    return dict(bounds_min=(origin,) * 3, bounds_max=(origin + box_length,) * 3, grid_dims=(dims,) * 3, cell_size=cs,
                periodic=tuple(periodic))


def device_count():
    n = C.c_int()
    _lib.load().xnb_device_count(C.byref(n))
    return n.value


class Box:
    def __init__(self, bounds_min, bounds_max, grid_dims, cell_size, periodic=(1, 1, 1), device=0):
        self.L = _lib.load()
        d = _lib.Domain()
        for a in range(3):
            d.bounds_min[a] = bounds_min[a]
            d.bounds_max[a] = bounds_max[a]
            d.grid_dims[a] = int(grid_dims[a])
            d.periodic[a] = int(periodic[a])
        d.cell_size = cell_size
        self._h = C.c_void_p()
        self._last_steps = 0
        self._check(self.L.xnb_create(C.byref(self._h), C.byref(d), int(device)), "xnb_create", ctx=False)

    def _check(self, rc, what, ctx=True):
        if rc != 0:
            msg = self.L.xnb_last_error(self._h if ctx else None)
            raise XnbError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if getattr(self, "_borrowed", False):  # owned by a Deck
            self._h = None
            return
        if getattr(self, "_h", None) is not None and self._h.value:
            self.L.xnb_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- slots / state
    def nbh_dist(self, rcut_max, rcut_inc):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.xnb_nbh_dist(self._h, rcut_max, rcut_inc, C.byref(a), C.byref(b), C.byref(c)), "nbh_dist")
        return dict(nbh_dist_lab=a.value, max_displ=b.value, ghost_dist=c.value)

    def set_particles(self, r, v=None, ids=None, types=None, extra=None):
        """extra: a list of up to 4 arrays of n values (extra per-particle fields)."""
        r = np.ascontiguousarray(r, np.float64)
        v = None if v is None else np.ascontiguousarray(v, np.float64)
        ids = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        types = None if types is None else np.ascontiguousarray(types, np.uint8)
        xs = [np.ascontiguousarray(x, np.float64) for x in (extra or [])]
        ptrs = (C.c_void_p * max(len(xs), 1))(*[x.ctypes.data for x in xs])
        self._check(self.L.xnb_set_particles_with_fields(self._h, r.shape[0], _ptr(r), _ptr(v), _ptr(ids), _ptr(types),
                                                         len(xs), ptrs), "set_particles")

    def num_particles(self):
        a, b = C.c_size_t(), C.c_size_t()
        self.L.xnb_num_particles(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def grid_info(self):
        dims = np.zeros(3, np.int64)
        gl, nc = C.c_int(), C.c_size_t()
        self.L.xnb_grid_info(self._h, _ptr(dims), C.byref(gl), C.byref(nc))
        return dims, gl.value, nc.value

    def cell_counts(self):
        _, _, nc = self.grid_info()
        out = np.zeros(nc, np.uint32)
        self._check(self.L.xnb_get_cell_counts(self._h, _ptr(out)), "get_cell_counts")
        return out

    def export(self, ghosts=False):
        n_in, n_g = self.num_particles()
        n = n_in + (n_g if ghosts else 0)
        out = dict(cell=np.zeros(n, np.uint32), r=np.zeros((n, 3)), v=np.zeros((n, 3)), f=np.zeros((n, 3)),
                   ep=np.zeros(n), id=np.zeros(n, np.uint64), type=np.zeros(n, np.uint8))
        self._check(self.L.xnb_get_particles(self._h, int(ghosts), _ptr(out["cell"]), _ptr(out["r"]), _ptr(out["v"]),
                                             _ptr(out["f"]), _ptr(out["ep"]), _ptr(out["id"]), _ptr(out["type"])),
                    "get_particles")
        nxc = C.c_int()
        self._check(self.L.xnb_field_count(self._h, C.byref(nxc)), "field_count")
        nx = nxc.value
        if nx:
            xs = [np.zeros(n) for _ in range(nx)]
            ptrs = (C.c_void_p * nx)(*[x.ctypes.data for x in xs])
            self._check(self.L.xnb_get_fields(self._h, int(ghosts), nx, ptrs), "get_fields")
            out["extra"] = xs
        return out

    # ---- operators
    def move_particles(self):
        self._check(self.L.xnb_move_particles(self._h), "move_particles")

    def backup_r(self):
        self._check(self.L.xnb_backup_r(self._h), "backup_r")

    def rebuild_amr(self, sub_grid_density=6.5, enforced_ordering=1):
        self._check(self.L.xnb_rebuild_amr(self._h, float(sub_grid_density), int(enforced_ordering)), "rebuild_amr")

    def ghost_update_all(self):
        self._check(self.L.xnb_ghost_update_all(self._h), "ghost_update_all")

    def ghost_update_r(self):
        self._check(self.L.xnb_ghost_update_r(self._h), "ghost_update_r")

    def particle_displ_over(self, threshold):
        n = C.c_uint64()
        self._check(self.L.xnb_particle_displ_over(self._h, threshold, C.byref(n)), "particle_displ_over")
        return n.value

    def chunk_neighbors(self, nbh_dist, build_particle_offset=1, dual_particle_offset=0, random_access=0,
                        half_symmetric=0, skip_ghosts=0, all_cells=False):
        cfg = _lib.ChunkNeighborsConfig(build_particle_offset, dual_particle_offset, random_access, half_symmetric,
                                        skip_ghosts)
        self._check(self.L.xnb_chunk_neighbors(self._h, nbh_dist, C.byref(cfg), int(all_cells)), "chunk_neighbors")

    def streams(self, cells=None):
        """Per-cell CS=1 uint16 streams and the max neighbour count: a list of arrays, one per
        cell, or a dict {cell: array} for the given cells."""
        tot, mx = C.c_size_t(), C.c_uint32()
        self._check(self.L.xnb_chunk_neighbors_stream(self._h, None, 0, None, C.byref(tot), C.byref(mx)), "stream size")
        _, _, nc = self.grid_info()
        buf = np.zeros(max(tot.value, 1), np.uint16)
        offs = np.zeros(nc + 1, np.uint64)
        self._check(self.L.xnb_chunk_neighbors_stream(self._h, _ptr(buf), buf.size, _ptr(offs), C.byref(tot),
                                                      C.byref(mx)), "stream export")
        if cells is not None:
            return {int(c): buf[offs[c]:offs[c + 1]].copy() for c in cells}, mx.value
        return [buf[offs[c]:offs[c + 1]].copy() for c in range(nc)], mx.value

    def zero_particle_force(self, ghost=False):
        self._check(self.L.xnb_zero_particle_force(self._h, int(ghost)), "zero_particle_force")

    def divide_force_by_type_scalar(self, values, ghost=False):
        """f /= values[type] (divide_force_by_type_scalar, e.g. with the type masses)."""
        v = np.ascontiguousarray(values, np.float64)
        self._check(self.L.xnb_divide_force_by_type_scalar(self._h, int(v.size), v.ctypes.data, int(ghost)),
                    "divide_force_by_type_scalar")

    def lennard_jones_force(self, epsilon, sigma, rcut, precision=_lib.PRECISION_FP64, ghost=False):
        p = _lib.LJParams(epsilon, sigma)
        self._check(self.L.xnb_lennard_jones_force(self._h, C.byref(p), rcut, int(precision), int(ghost)),
                    "lennard_jones_force")

    def lennard_jones_force_multimat(self, ntypes, eps_pair, sigma_pair, rcut_pair, precision=_lib.PRECISION_FP64):
        e = np.ascontiguousarray(eps_pair, np.float64)
        s = np.ascontiguousarray(sigma_pair, np.float64)
        rc = np.ascontiguousarray(rcut_pair, np.float64)
        self._check(self.L.xnb_lennard_jones_force_multimat(self._h, int(ntypes), _ptr(e), _ptr(s), _ptr(rc),
                                                            int(precision)), "lennard_jones_force_multimat")

    def gravitational_force(self, G, type_mass, rcut, ghost=False):
        """f += G m_a m_b dr / r^3 over pairs within rcut, masses per particle type."""
        m = np.ascontiguousarray(type_mass, np.float64)
        self._check(self.L.xnb_gravitational_force(self._h, float(G), int(m.size), _ptr(m), float(rcut), int(bool(ghost))),
                    "gravitational_force")

    def push_f_v_r(self, dt, dt_scale=1.0):
        self._check(self.L.xnb_push_f_v_r(self._h, dt, dt_scale), "push_f_v_r")

    def push_f_v(self, dt, dt_scale=0.5):
        self._check(self.L.xnb_push_f_v(self._h, dt, dt_scale), "push_f_v")

    def energies(self):
        ke, pe = C.c_double(), C.c_double()
        self._check(self.L.xnb_simulation_energies(self._h, C.byref(ke), C.byref(pe)), "simulation_energies")
        return ke.value, pe.value

    # ---- composite
    def init(self, epsilon, sigma, rcut, rcut_inc, precision=_lib.PRECISION_FP64):
        p = _lib.LJParams(epsilon, sigma)
        self._check(self.L.xnb_init(self._h, C.byref(p), rcut, rcut_inc, int(precision)), "init")

    def run(self, steps, dt, epsilon, sigma, rcut, rcut_inc, precision=_lib.PRECISION_FP64, rebuild_every=0):
        p = _lib.RunParams(dt, _lib.LJParams(epsilon, sigma), rcut, rcut_inc, int(precision), int(rebuild_every),
                           int(steps), 0, 0)
        self._check(self.L.xnb_run(self._h, C.byref(p)), "run")
        self._last_steps = int(steps)
        return dict(rebuilds=p.n_rebuilds, violations=p.n_violations)

    def run_step_counts(self, n=None):
        """particle_displ_over count of every step of the last run (index k - 1 for step k)."""
        n = self._last_steps if n is None else n
        out = np.zeros(n, np.uint64)
        self._check(self.L.xnb_run_step_counts(self._h, _ptr(out), n), "run_step_counts")
        return out

    def run_multimat(self, steps, dt, ntypes, eps_pair, sigma_pair, rcut_pair, rcut_inc, precision=_lib.PRECISION_FP64,
                     rebuild_every=0):
        e = np.ascontiguousarray(eps_pair, np.float64)
        s = np.ascontiguousarray(sigma_pair, np.float64)
        rc = np.ascontiguousarray(rcut_pair, np.float64)
        p = _lib.RunParams(dt, _lib.LJParams(1.0, 1.0), float(rc.max()), rcut_inc, int(precision), int(rebuild_every), int(steps),
                           0, 0)
        self._check(self.L.xnb_run_multimat(self._h, C.byref(p), int(ntypes), _ptr(e), _ptr(s), _ptr(rc)), "run_multimat")
        self._last_steps = int(steps)
        return dict(rebuilds=p.n_rebuilds, violations=p.n_violations)

    def autotune(self, epsilon, sigma, rcut, precision=_lib.PRECISION_FP64):
        """Brick shape chosen by timing (cached per box class); returns it (0s off the brick path)."""
        p = _lib.LJParams(epsilon, sigma)
        dims = np.zeros(3, np.int32)
        self._check(self.L.xnb_autotune(self._h, C.byref(p), rcut, int(precision), _ptr(dims)), "autotune")
        return tuple(int(x) for x in dims)

    # ---- snapshots
    def snapshot(self):
        """Bytes of the box's full step-loop state (see xnb_snapshot_save)."""
        n = C.c_size_t()
        self._check(self.L.xnb_snapshot_size(self._h, C.byref(n)), "snapshot_size")
        buf = np.zeros(n.value, np.uint8)
        self._check(self.L.xnb_snapshot_save(self._h, _ptr(buf), buf.size), "snapshot_save")
        return buf

    def restore(self, snap):
        snap = np.ascontiguousarray(snap, np.uint8)
        self._check(self.L.xnb_snapshot_load(self._h, _ptr(snap), snap.size), "snapshot_load")

    # ---- measurement
    def timer_enable(self, on=True):
        self.L.xnb_timer_enable(self._h, int(on))

    def timer_reset(self):
        self.L.xnb_timer_reset(self._h)

    def timer_read(self, timer):
        ms, n = C.c_double(), C.c_int64()
        self._check(self.L.xnb_timer_read(self._h, int(timer), C.byref(ms), C.byref(n)), "timer_read")
        return ms.value, n.value

    def pair_counts(self):
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.L.xnb_pair_counts(self._h, C.byref(a), C.byref(b)), "pair_counts")
        return a.value, b.value

    def copy_probe(self, nbytes, iters=10):
        g = C.c_double()
        self._check(self.L.xnb_copy_probe(self._h, int(nbytes), int(iters), C.byref(g)), "copy_probe")
        return g.value

    def fp64_core(self, x):
        """Device sqrt / reciprocal sequences of the FP64 force kernels applied to x."""
        x = np.ascontiguousarray(x, np.float64)
        sq, rc = np.empty_like(x), np.empty_like(x)
        self._check(self.L.xnb_fp64_core(self._h, _ptr(x), x.size, _ptr(sq), _ptr(rc)), "fp64_core")
        return sq, rc

    def device_bytes(self):
        """(device bytes held now, most held at once) by this context."""
        a, b = C.c_size_t(), C.c_size_t()
        self._check(self.L.xnb_device_bytes(self._h, C.byref(a), C.byref(b)), "device_bytes")
        return a.value, b.value

    def synchronize(self):
        self._check(self.L.xnb_synchronize(self._h), "synchronize")
