"""Python handle on a batch of boxes (xnb_batch): many independent periodic boxes advanced
together, one launch per kernel per step over all of them (include/xnb_amd.h, batches)."""
import ctypes as C

from . import _lib
from ._lib import XnbError


class Batch:
    def __init__(self, device=0):
        self.L = _lib.load()
        self._h = C.c_void_p()
        rc = self.L.xnb_batch_create(C.byref(self._h), int(device))
        if rc != 0:
            raise XnbError(rc, "xnb_batch_create failed")
        self._boxes = {}

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.xnb_batch_error(self._h)
            raise XnbError(rc, f"{what}: {msg.decode() if msg else ''}")

    def attach(self, box, steps, dt, epsilon, sigma, rcut, rcut_inc, precision=_lib.PRECISION_FP64, rebuild_every=10):
        """Attach an initialised Box for `steps` steps; returns its slot.  rebuild_every 0
        rebuilds on the displacement trigger (particle_displ_over), like Box.run."""
        p = _lib.RunParams(dt, _lib.LJParams(epsilon, sigma), rcut, rcut_inc, int(precision), int(rebuild_every), int(steps),
                           0, 0)
        slot = C.c_int()
        self._check(self.L.xnb_batch_attach(self._h, box._h, C.byref(p), C.byref(slot)), "attach")
        self._boxes[slot.value] = box
        return slot.value

    def step(self, nsteps=1):
        """Advance all unfinished boxes by up to nsteps; returns the number still unfinished."""
        n = C.c_int()
        self._check(self.L.xnb_batch_step(self._h, int(nsteps), C.byref(n)), "step")
        return n.value

    def status(self, slot):
        done, state = C.c_int64(), C.c_int()
        self._check(self.L.xnb_batch_box_status(self._h, int(slot), C.byref(done), C.byref(state)), "status")
        return done.value, state.value

    def recoveries(self, slot):
        n = C.c_int64()
        self._check(self.L.xnb_batch_box_recoveries(self._h, int(slot), C.byref(n)), "recoveries")
        return n.value

    def graph_stats(self):
        """(steps replayed from a captured HIP graph, graphs captured)."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.L.xnb_batch_graph_stats(self._h, C.byref(a), C.byref(b)), "graph_stats")
        return a.value, b.value

    def autotune(self):
        """Choose the batch's brick LDS cap (and with it every box's brick shape) by timing the
        batched build and force of the attached boxes, before their first step; returns the
        kept fraction of the CU's LDS."""
        f = C.c_double()
        self._check(self.L.xnb_batch_autotune(self._h, C.byref(f)), "autotune")
        return f.value

    def detach(self, slot):
        """Detach a box; returns dict(rebuilds, violations) like Box.run."""
        p = _lib.RunParams()
        self._check(self.L.xnb_batch_detach(self._h, int(slot), C.byref(p)), "detach")
        self._boxes.pop(slot, None)
        return dict(rebuilds=p.n_rebuilds, violations=p.n_violations)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.L.xnb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
