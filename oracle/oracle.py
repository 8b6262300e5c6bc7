"""ctypes binding of the CPU oracle (oracle/xnb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(exanbody_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libxnb_oracle.so")
_libs = {}


class OracleError(RuntimeError):
    pass


class NbhConfig(C.Structure):
    _fields_ = [("build_particle_offset", C.c_int), ("dual_particle_offset", C.c_int),
                ("random_access", C.c_int), ("half_symmetric", C.c_int), ("skip_ghosts", C.c_int)]


class RunParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("epsilon", C.c_double), ("sigma", C.c_double),
                ("rcut", C.c_double), ("rcut_inc", C.c_double), ("rebuild_every", C.c_int),
                ("steps", C.c_int64), ("n_rebuilds", C.c_int64), ("n_violations", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(path=None):
    """The oracle library (path: another build of the same source, e.g. a planted-mistake one)."""
    path = path or _LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P, D, I, I64, SZ = C.c_void_p, C.c_double, C.c_int, C.c_int64, C.c_size_t
        L.op_create.argtypes = [C.POINTER(P), P, P, P, D, P, D]
        L.op_destroy.argtypes = [P]
        L.op_ghost_layers.argtypes = [P]
        L.op_grid_dims.argtypes = [P, P]
        L.op_num_cells.argtypes = [P]; L.op_num_cells.restype = SZ
        L.op_num_particles.argtypes = [P]; L.op_num_particles.restype = SZ
        L.op_num_ghost_particles.argtypes = [P]; L.op_num_ghost_particles.restype = SZ
        L.op_cell_count.argtypes = [P, SZ]; L.op_cell_count.restype = C.c_uint32
        L.op_set_particles.argtypes = [P, SZ, P, P, P, P]
        L.op_export.argtypes = [P, I, P, P, P, P, P, P, P]; L.op_export.restype = SZ
        for name in ("op_move_particles", "op_ghost_update_all", "op_ghost_update_r", "op_backup_r"):
            getattr(L, name).argtypes = [P]
        L.op_rebuild_amr.argtypes = [P, D]
        L.op_sub_grid_size.argtypes = [C.c_uint32, D]; L.op_sub_grid_size.restype = I
        L.op_particle_displ_over.argtypes = [P, D]; L.op_particle_displ_over.restype = I64
        L.op_chunk_neighbors.argtypes = [P, D, C.POINTER(NbhConfig)]
        L.op_zero_force.argtypes = [P, I]
        L.op_lj_force.argtypes = [P, D, D, D, I]
        L.op_lj_force_multimat.argtypes = [P, I, P, P, P]
        L.op_gravitational_force.argtypes = [P, D, I, P, D, I]
        L.op_push_f_v_r.argtypes = [P, D, D]
        L.op_push_f_v.argtypes = [P, D, D]
        L.op_stream_size.argtypes = [P, SZ]; L.op_stream_size.restype = SZ
        L.op_stream_data.argtypes = [P, SZ]; L.op_stream_data.restype = C.POINTER(C.c_uint16)
        L.op_max_neighbors.argtypes = [P]; L.op_max_neighbors.restype = C.c_uint32
        L.op_init.argtypes = [P, D, D, D, D]
        L.op_run.argtypes = [P, C.POINTER(RunParams)]
        L.op_energies.argtypes = [P, C.POINTER(D), C.POINTER(D)]
        L.op_encode_double_u32.argtypes = [D, D, D]; L.op_encode_double_u32.restype = C.c_uint32
        L.op_restore_u32_double.argtypes = [C.c_uint32, D, D]; L.op_restore_u32_double.restype = D
        L.op_num_threads.restype = I
        _libs[path] = L
    return _libs[path]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what, allow=()):
    if rc != 0 and rc not in allow:
        raise OracleError(f"{what} failed with code {rc}")
    return rc


class Oracle:
    """One periodic box on the CPU oracle (per-cell SoA grid with ghost layers)."""

    def __init__(self, origin, extent, grid_dims, cell_size, periodic=(1, 1, 1), ghost_dist=2.8, libpath=None):
        L = lib(libpath)
        self._h = C.c_void_p()
        o = np.ascontiguousarray(origin, dtype=np.float64)
        e = np.ascontiguousarray(extent, dtype=np.float64)
        g = np.ascontiguousarray(grid_dims, dtype=np.int64)
        p = np.ascontiguousarray(periodic, dtype=np.int32)
        _check(L.op_create(C.byref(self._h), _ptr(o), _ptr(e), _ptr(g), float(cell_size), _ptr(p),
                           float(ghost_dist)), "op_create")
        self.L = L

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.L.op_destroy(h)
            self._h = None

    @property
    def h(self):
        return self._h

    @property
    def ghost_layers(self):
        return self.L.op_ghost_layers(self._h)

    @property
    def dims(self):
        d = np.zeros(3, dtype=np.int64)
        self.L.op_grid_dims(self._h, _ptr(d))
        return d

    @property
    def num_cells(self):
        return self.L.op_num_cells(self._h)

    def num_particles(self):
        return self.L.op_num_particles(self._h)

    def num_ghost_particles(self):
        return self.L.op_num_ghost_particles(self._h)

    def cell_counts(self):
        return np.array([self.L.op_cell_count(self._h, c) for c in range(self.num_cells)], dtype=np.uint32)

    def set_particles(self, r, v=None, ids=None, types=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        n = r.shape[0]
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ids = np.arange(n, dtype=np.uint64) if ids is None else np.ascontiguousarray(ids, dtype=np.uint64)
        types = None if types is None else np.ascontiguousarray(types, dtype=np.uint8)
        _check(self.L.op_set_particles(self._h, n, _ptr(r), _ptr(v), _ptr(ids), _ptr(types)), "op_set_particles")

    def export(self, ghosts=False):
        n = self.num_particles() + (self.num_ghost_particles() if ghosts else 0)
        out = dict(cell=np.zeros(n, np.uint32), r=np.zeros((n, 3)), v=np.zeros((n, 3)), f=np.zeros((n, 3)),
                   ep=np.zeros(n), id=np.zeros(n, np.uint64), type=np.zeros(n, np.uint8))
        k = self.L.op_export(self._h, int(ghosts), _ptr(out["cell"]), _ptr(out["r"]), _ptr(out["v"]),
                             _ptr(out["f"]), _ptr(out["ep"]), _ptr(out["id"]), _ptr(out["type"]))
        assert k == n
        return out

    def move_particles(self):
        _check(self.L.op_move_particles(self._h), "move_particles")

    def ghost_update_all(self):
        _check(self.L.op_ghost_update_all(self._h), "ghost_update_all")

    def ghost_update_r(self):
        _check(self.L.op_ghost_update_r(self._h), "ghost_update_r")

    def backup_r(self):
        _check(self.L.op_backup_r(self._h), "backup_r")

    def rebuild_amr(self, sub_grid_density=6.5):
        _check(self.L.op_rebuild_amr(self._h, float(sub_grid_density)), "rebuild_amr")

    def particle_displ_over(self, threshold):
        return int(self.L.op_particle_displ_over(self._h, float(threshold)))

    def chunk_neighbors(self, nbh_dist, build_particle_offset=1, dual_particle_offset=0, random_access=0,
                        half_symmetric=0, skip_ghosts=0):
        cfg = NbhConfig(build_particle_offset, dual_particle_offset, random_access, half_symmetric, skip_ghosts)
        return _check(self.L.op_chunk_neighbors(self._h, float(nbh_dist), C.byref(cfg)), "chunk_neighbors")

    def zero_force(self, ghost=False):
        _check(self.L.op_zero_force(self._h, int(ghost)), "zero_force")

    def lj_force(self, epsilon, sigma, rcut, ghost=False):
        _check(self.L.op_lj_force(self._h, float(epsilon), float(sigma), float(rcut), int(ghost)), "lj_force")

    def lj_force_multimat(self, ntypes, eps_pair, sigma_pair, rcut_pair):
        e = np.ascontiguousarray(eps_pair, np.float64)
        s = np.ascontiguousarray(sigma_pair, np.float64)
        rc = np.ascontiguousarray(rcut_pair, np.float64)
        _check(self.L.op_lj_force_multimat(self._h, int(ntypes), _ptr(e), _ptr(s), _ptr(rc)), "lj_force_multimat")

    def gravitational_force(self, G, type_mass, rcut, ghost=False):
        m = np.ascontiguousarray(type_mass, np.float64)
        _check(self.L.op_gravitational_force(self._h, float(G), int(m.size), _ptr(m), float(rcut), int(bool(ghost))),
               "gravitational_force")

    def push_f_v_r(self, dt, dt_scale=1.0):
        _check(self.L.op_push_f_v_r(self._h, float(dt), float(dt_scale)), "push_f_v_r")

    def push_f_v(self, dt, dt_scale=0.5):
        _check(self.L.op_push_f_v(self._h, float(dt), float(dt_scale)), "push_f_v")

    def stream(self, cell):
        n = self.L.op_stream_size(self._h, cell)
        if n == 0:
            return np.zeros(0, np.uint16)
        p = self.L.op_stream_data(self._h, cell)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def streams(self):
        return [self.stream(c) for c in range(self.num_cells)]

    def max_neighbors(self):
        return int(self.L.op_max_neighbors(self._h))

    def init(self, epsilon, sigma, rcut, rcut_inc):
        _check(self.L.op_init(self._h, float(epsilon), float(sigma), float(rcut), float(rcut_inc)), "init")

    def run(self, steps, dt, epsilon, sigma, rcut, rcut_inc, rebuild_every=0):
        p = RunParams(dt, epsilon, sigma, rcut, rcut_inc, rebuild_every, steps, 0, 0)
        _check(self.L.op_run(self._h, C.byref(p)), "run")
        return dict(rebuilds=p.n_rebuilds, violations=p.n_violations)

    def energies(self):
        ke, pe = C.c_double(), C.c_double()
        self.L.op_energies(self._h, C.byref(ke), C.byref(pe))
        return ke.value, pe.value


def encode_double_u32(x, o, r):
    return lib().op_encode_double_u32(float(x), float(o), float(r))


def sub_grid_size(n, density=6.5):
    return int(lib().op_sub_grid_size(int(n), float(density)))


def restore_u32_double(x, o, r):
    return lib().op_restore_u32_double(int(x), float(o), float(r))


def num_threads():
    return lib().op_num_threads()
