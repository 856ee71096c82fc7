"""Neighbor lists with the reference's plugin interface.

``CellListNeighborList`` / ``DenseNearestNeighborList`` mirror
core/neighborlist.py:902-1077: ``nl(lh, rh, systems, cutoffs,
rh_index_remap)`` returns the directed ``Edges`` of every system.  The pair
set is decided on the GPU in FP64 by ``kupsmi_neighborlist`` with the
reference's semantics (strict cutoff, image grid for narrow cells,
minimum-image-only exclusion).  Capacities follow the reference: the edge
buffer holds ``avg_edges * n`` entries, an overflow raises CapacityError or,
with ``fix=True``, regrows to the next power of two and recomputes
(core/capacity.py:175-234, core/propagator.py:137-172).
``neighborlist_changes`` mirrors core/neighborlist.py:1424-1515: the edges a
particle move removes and adds, from ``kupsmi_neighborlist_changes``.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .. import _lib
from .batch import ParticleBatch, SystemBatch, system_offsets
from .capacity import CapacityError, next_power_of_two


@dataclass
class Edges:
    """Directed pairs: indices (E, 2) and integer image shifts (E, 1, 3);
    d_ij = r_j - r_i + shift @ L (core/neighborlist.py:115-211)."""

    indices: torch.Tensor
    shifts: torch.Tensor

    def __len__(self) -> int:
        return self.indices.shape[0]

    def difference_vectors(self, positions: torch.Tensor, system: torch.Tensor,
                           systems: SystemBatch) -> torch.Tensor:
        lat = systems.lattice()[system[self.indices[:, 0]].long()]
        sh = self.shifts[:, 0].to(lat.dtype)
        off = torch.einsum("ej,ejk->ek", sh, lat)
        p = positions.to(lat.dtype)
        return p[self.indices[:, 1].long()] - p[self.indices[:, 0].long()] + off


@dataclass
class DenseNearestNeighborList:
    """All same-system pairs (core/neighborlist.py:902-982)."""

    avg_edges: int = 64
    fix: bool = True

    def __call__(self, lh: ParticleBatch, rh: ParticleBatch | None, systems: SystemBatch,
                 cutoffs: torch.Tensor, rh_index_remap=None) -> Edges:
        if rh is not None or rh_index_remap is not None:
            raise NotImplementedError(
                "rh queries are served by neighborlist_changes (the edges a move removes and "
                "adds) and inside the GCMC kernels; the full list takes rh=None")
        lib = _lib.load()
        n_sys = systems.n_sys
        sys_off, _ = system_offsets(lh.system, n_sys)
        excl = lh.exclusion
        if excl is None:
            excl = torch.arange(lh.n, dtype=torch.int32, device=lh.positions.device)
        cut = torch.as_tensor(cutoffs, dtype=torch.float64, device=lh.positions.device)
        cut = cut.expand(n_sys).contiguous() if cut.dim() == 0 or cut.numel() == 1 else cut.contiguous()
        # converted inputs are bound to names so they outlive the launches
        pos = lh.positions.to(torch.float32).contiguous()
        excl = excl.to(torch.int32).contiguous()
        count = torch.zeros(1, dtype=torch.int64, device=lh.positions.device)
        avg = self.avg_edges
        while True:
            cap = max(avg * lh.n, 1)
            ij = torch.empty((cap, 2), dtype=torch.int32, device=lh.positions.device)
            sh = torch.empty((cap, 3), dtype=torch.int32, device=lh.positions.device)
            _lib.check(lib.kupsmi_neighborlist(
                _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), lh.n,
                _lib.ptr(pos), _lib.ptr(excl),
                _lib.ptr(systems.tril), _lib.ptr(cut), cap, _lib.ptr(ij), _lib.ptr(sh),
                _lib.ptr(count)))
            required = int(count.item())
            if required <= cap:
                self.avg_edges = avg
                return Edges(ij[:required], sh[:required].reshape(-1, 1, 3))
            if not self.fix:
                raise CapacityError(
                    f"neighbor list needs {required} edges, capacity is {cap}", required)
            avg = next_power_of_two(-(-required // max(lh.n, 1)))


@dataclass
class CellListNeighborList(DenseNearestNeighborList):
    """Cell-list neighbor list (core/neighborlist.py:985-1077).  Its pair set is
    the dense set; the export decides every pair in FP64, the MD/GCMC kernels
    use LDS cell lists internally."""


@dataclass
class NeighborListChangesResult:
    added: Edges
    removed: Edges


def _inputs(lh: ParticleBatch, systems: SystemBatch, cutoffs):
    dev = lh.positions.device
    sys_off, _ = system_offsets(lh.system, systems.n_sys)
    excl = lh.exclusion
    if excl is None:
        excl = torch.arange(lh.n, dtype=torch.int32, device=dev)
    cut = torch.as_tensor(cutoffs, dtype=torch.float64, device=dev)
    n_sys = systems.n_sys
    cut = cut.expand(n_sys).contiguous() if cut.dim() == 0 or cut.numel() == 1 else cut.contiguous()
    return (sys_off, lh.positions.to(torch.float32).contiguous(),
            excl.to(torch.int32).contiguous(), cut)


def neighborlist_changes(neighborlist: DenseNearestNeighborList, lh: ParticleBatch,
                         rh_indices: torch.Tensor, rh_positions: torch.Tensor,
                         systems: SystemBatch, cutoffs,
                         compaction: float = 0.5) -> NeighborListChangesResult:
    """Edges removed and added when particles ``rh_indices`` (distinct) move
    to ``rh_positions`` (neighborlist.py:1424-1515; the reference's ``rh``
    is ``WithIndices(indices, data)``).  ``removed``: the current directed
    edges with an end among the moved particles; ``added``: those of the
    moved configuration; indices are the particles' own.  Each output holds
    up to ``compaction`` x the neighbor list's edge buffer for n + k
    particles; more raises CapacityError or, with ``neighborlist.fix``,
    regrows the buffer (next power of two) and recomputes."""
    lib = _lib.load()
    dev = lh.positions.device
    idx = rh_indices.to(torch.int32).contiguous()
    new = rh_positions.to(torch.float32).reshape(-1, 3).contiguous()
    k = idx.numel()
    if new.shape[0] != k:
        raise ValueError(f"{k} moved particles but {new.shape[0]} new positions")
    sys_off, pos, excl, cut = _inputs(lh, systems, cutoffs)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    avg = neighborlist.avg_edges
    while True:
        cap = max(int(avg * (lh.n + k) * compaction), 1)
        a_ij = torch.empty((cap, 2), dtype=torch.int32, device=dev)
        a_sh = torch.empty((cap, 3), dtype=torch.int32, device=dev)
        r_ij = torch.empty((cap, 2), dtype=torch.int32, device=dev)
        r_sh = torch.empty((cap, 3), dtype=torch.int32, device=dev)
        _lib.check(lib.kupsmi_neighborlist_changes(
            _lib.ctx(), _lib.stream_ptr(), systems.n_sys, _lib.ptr(sys_off), lh.n, _lib.ptr(pos),
            _lib.ptr(excl), _lib.ptr(systems.tril), _lib.ptr(cut), k, _lib.ptr(idx),
            _lib.ptr(new), cap, _lib.ptr(a_ij), _lib.ptr(a_sh), _lib.ptr(counts[0:1]), cap,
            _lib.ptr(r_ij), _lib.ptr(r_sh), _lib.ptr(counts[1:2])))
        n_added, n_removed = (int(v) for v in counts.tolist())
        if n_added < 0:
            raise ValueError("moved particle indices must be distinct and in range")
        if max(n_added, n_removed) <= cap:
            neighborlist.avg_edges = avg
            return NeighborListChangesResult(
                Edges(a_ij[:n_added], a_sh[:n_added].reshape(-1, 1, 3)),
                Edges(r_ij[:n_removed], r_sh[:n_removed].reshape(-1, 1, 3)))
        if not neighborlist.fix:
            raise CapacityError(
                f"neighborlist_changes needs {max(n_added, n_removed)} edges, capacity is {cap}",
                max(n_added, n_removed))
        avg = next_power_of_two(-(-max(n_added, n_removed) // max(int((lh.n + k) * compaction), 1)))
