"""Parity output of the force kernels: the pair set a force evaluation used.

``PairSetRecorder(capacity).arm()`` makes the next pair-evaluating call of
the library context (``kupsmi_md_lj`` through the fused MD propagators,
``kupsmi_lj_forces``, ``kupsmi_pair_forces``) or of a Verlet handle
(``arm(verlet)``, ``kupsmi_pair_forces_verlet``) write every directed pair
(i, j, shift) its force loops kept, with the terms that kept it (bit 0 LJ,
bit 1 real-space Coulomb).  The armed call runs the recording instantiation
of the same kernels: same source, same pair decisions, plus the writes.  The
reference has no such output; its pair sets are its neighbor lists
(core/neighborlist.py:115-211), which the recorded sets are compared with.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from .capacity import CapacityError


class PairSetRecorder:
    def __init__(self, capacity: int, device=None):
        dev = device or _lib.require_gpu()
        self.capacity = int(capacity)
        self.ij = torch.empty((max(self.capacity, 1), 2), dtype=torch.int32, device=dev)
        self.shift = torch.empty((max(self.capacity, 1), 3), dtype=torch.int32, device=dev)
        self.terms = torch.empty(max(self.capacity, 1), dtype=torch.uint8, device=dev)
        # the library writes an unsigned 64-bit count; a signed tensor holds the same bits
        self.count = torch.zeros(1, dtype=torch.int64, device=dev)

    def arm(self, verlet=None) -> "PairSetRecorder":
        lib = _lib.load()
        args = (self.capacity, _lib.ptr(self.ij), _lib.ptr(self.shift), _lib.ptr(self.terms),
                _lib.ptr(self.count))
        if verlet is not None:
            _lib.check(lib.kupsmi_verlet_record_pairs(verlet.handle, *args))
        else:
            _lib.check(lib.kupsmi_ctx_record_pairs(_lib.ctx(), *args))
        return self

    def result(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(ij (E, 2), shift (E, 3), terms (E,)) on the host, in no particular
        order; raises CapacityError when the pairs did not fit."""
        torch.cuda.synchronize()
        n = int(self.count.item())
        if n > self.capacity:
            raise CapacityError(f"pair record needs {n} entries, capacity is {self.capacity}", n)
        return (self.ij[:n].cpu().numpy(), self.shift[:n].cpu().numpy(),
                self.terms[:n].cpu().numpy())
