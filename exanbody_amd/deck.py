"""Python handle on an operator deck (xnb_deck): YAML-subset decks with exaNBody's operator
names, run through the C ABI.  `exanbody_amd/decks/lj_verlet.msp` holds the step loop."""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import XnbError

DECK_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Deck:
    def __init__(self, *texts, files=("lj_verlet.msp",)):
        self.L = _lib.load()
        self._h = C.c_void_p()
        self.L.xnb_deck_create(C.byref(self._h))
        for f in files:
            self.load(open(os.path.join(DECK_DIR, f)).read())
        for t in texts:
            self.load(t)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.L.xnb_deck_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def load(self, text):
        rc = self.L.xnb_deck_load(self._h, text.encode())
        if rc != 0:
            raise XnbError(rc, "deck load: " + self.L.xnb_deck_error(self._h).decode())

    def set_particles(self, r, v=None, ids=None, types=None):
        r = np.ascontiguousarray(r, np.float64)
        v = None if v is None else np.ascontiguousarray(v, np.float64)
        ids = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        types = None if types is None else np.ascontiguousarray(types, np.uint8)
        self.L.xnb_deck_set_particles(self._h, r.shape[0], _ptr(r), _ptr(v), _ptr(ids), _ptr(types))

    def run(self, entry="simulation"):
        rc = self.L.xnb_deck_run(self._h, entry.encode() if entry else None)
        if rc != 0:
            raise XnbError(rc, "deck run: " + self.L.xnb_deck_error(self._h).decode())
        return self.L.xnb_deck_log(self._h).decode()

    def slot(self, name):
        buf = C.create_string_buffer(256)
        rc = self.L.xnb_deck_slot(self._h, name.encode(), buf, 256)
        return buf.value.decode() if rc == 0 else None

    def box(self):
        """The device box created by the deck's `domain` operator (borrowed handle)."""
        from .box import Box
        h = self.L.xnb_deck_context(self._h)
        if not h:
            return None
        b = Box.__new__(Box)
        b.L = self.L
        b._h = C.c_void_p(h)
        b._borrowed = True
        return b


def domain_block(bounds_min, bounds_max, grid_dims, cell_size, periodic=(1, 1, 1)):
    f = lambda v: "[" + ", ".join(repr(float(x)) for x in v) + "]"
    return ("domain:\n"
            f"  cell_size: {float(cell_size)!r}\n"
            f"  grid_dims: [{', '.join(str(int(x)) for x in grid_dims)}]\n"
            f"  bounds: [{f(bounds_min)}, {f(bounds_max)}]\n"
            f"  periodic: [{', '.join('true' if p else 'false' for p in periodic)}]\n")
