"""Verlet list with a skin for the HBM pair path (kupsmi_pair_forces_verlet).

The list lives in a device handle owned by this object; every call re-uses it
and the kernels rebuild a system's list on the GPU once one of its atoms has
moved more than skin / 2 since the last build (no host synchronisation).
Cells, charges, species, exclusion ids and sizes must stay fixed while the
list is reused; call ``invalidate()`` after changing any of them.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .. import _lib


def neighbor_capacity(sizes, tril, cutoff: float, skin: float, margin: float = 1.4,
                      extra: int = 64) -> int:
    """Entries per atom for a list of radius cutoff + skin: margin x the mean
    number of atoms in that sphere at the densest system's density, + extra.
    An atom whose list overflows still gets exact forces (the kernel walks the
    cell stencil for it), only slower."""
    tril = np.asarray(tril, float).reshape(-1, 6)
    vol = np.abs(tril[:, 0] * tril[:, 2] * tril[:, 5])
    rho = float(np.max(np.asarray(sizes, float) / vol)) if len(vol) else 0.0
    return int(math.ceil(margin * rho * 4.0 / 3.0 * math.pi * (cutoff + skin) ** 3)) + extra


class VerletList:
    def __init__(self, skin: float, max_neighbors: int):
        if skin < 0:
            raise ValueError("skin must be non-negative")
        self.skin = float(skin)
        self.max_neighbors = int(max_neighbors)
        h = C.c_void_p()
        _lib.check(_lib.load().kupsmi_verlet_create(C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def freeze(self, frozen: bool = True) -> None:
        """While frozen, a call that would resize the handle's buffers fails
        (for handles used by recorded work)."""
        _lib.check(_lib.load().kupsmi_verlet_freeze(self._h, int(frozen)))

    def profile(self, enable: bool = True) -> tuple[float, int]:
        """(summed HIP-event milliseconds, calls) of the main pair pass since
        the previous call (synchronises on those events), then turns the
        timing of later calls on or off."""
        ms = C.c_double()
        n = C.c_longlong()
        _lib.check(_lib.load().kupsmi_verlet_profile(self._h, int(enable), C.byref(ms),
                                                     C.byref(n)))
        return ms.value, n.value

    def invalidate(self):
        _lib.check(_lib.load().kupsmi_verlet_invalidate(self._h, _lib.stream_ptr()))

    def counters(self) -> tuple[int, int]:
        """(system rebuilds, overflowed atom lists) since the handle's buffers
        were sized (synchronises the stream)."""
        out = (C.c_ulonglong * 2)()
        _lib.check(_lib.load().kupsmi_verlet_counters(self._h, _lib.stream_ptr(), out))
        return int(out[0]), int(out[1])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().kupsmi_verlet_destroy(h)
            except Exception:
                pass
            self._h = None
