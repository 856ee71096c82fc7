"""Neighbor lists with the reference's plugin interface.

``CellListNeighborList`` / ``DenseNearestNeighborList`` mirror
core/neighborlist.py:902-1077: ``nl(lh, rh, systems, cutoffs,
rh_index_remap)`` returns the directed ``Edges`` of every system (rh = None)
or of an rh subset against the table.  The pair set is decided on the GPU in
FP64 (``kupsmi_neighborlist_cells`` / ``kupsmi_neighborlist`` /
``kupsmi_neighborlist_query``) with the reference's semantics (strict
cutoff, image grid for narrow cells, minimum-image-only exclusion).  Capacities follow the reference: the edge
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


def _cutoffs(cutoffs, n_sys: int, dev) -> torch.Tensor:
    cut = torch.as_tensor(cutoffs, dtype=torch.float64, device=dev)
    return cut.expand(n_sys).contiguous() if cut.dim() == 0 or cut.numel() == 1 else cut.contiguous()


@dataclass
class DenseNearestNeighborList:
    """All same-system pairs (core/neighborlist.py:902-982).

    ``nl(lh, None, systems, cutoffs)`` is the full directed list of every
    system (``kupsmi_neighborlist``, every ordered pair tested);
    ``nl(lh, rh, systems, cutoffs, rh_index_remap)`` the rh-subset query
    (``kupsmi_neighborlist_query``, the reference's rules of
    core/neighborlist.py:505-539, 680-685).  ``avg_edges`` sizes the full
    list's edge buffer (``avg_edges * n``), ``avg_query_edges`` the query's
    (``avg_query_edges * 2 k``) and ``avg_changes`` the buffers of
    ``neighborlist_changes``; each regrows on its own (the reference has one
    ``max_num_edges`` per list; here a large query does not resize the full
    list's buffer)."""

    avg_edges: int = 64
    fix: bool = True
    avg_query_edges: int = 64
    avg_changes: int | None = None

    def __call__(self, lh: ParticleBatch, rh: ParticleBatch | None, systems: SystemBatch,
                 cutoffs: torch.Tensor, rh_index_remap=None) -> Edges:
        if rh is None:
            if rh_index_remap is not None:
                raise ValueError("rh_index_remap needs rh particles")
            return self._full(lh, systems, cutoffs)
        return self._query(lh, rh, systems, cutoffs, rh_index_remap)

    def _launch(self, lib, args, cap, ij, sh, count):
        _lib.check(lib.kupsmi_neighborlist(*args, cap, _lib.ptr(ij), _lib.ptr(sh),
                                           _lib.ptr(count)))

    def _full(self, lh: ParticleBatch, systems: SystemBatch, cutoffs) -> Edges:
        lib = _lib.load()
        n_sys = systems.n_sys
        dev = lh.positions.device
        sys_off, sizes = system_offsets(lh.system, n_sys)
        excl = lh.exclusion
        if excl is None:
            excl = torch.arange(lh.n, dtype=torch.int32, device=dev)
        cut = _cutoffs(cutoffs, n_sys, dev)
        # converted inputs are bound to names so they outlive the launches
        pos = lh.positions.to(torch.float32).contiguous()
        excl = excl.to(torch.int32).contiguous()
        count = torch.zeros(1, dtype=torch.int64, device=dev)
        args = self._args(lh, n_sys, sys_off, sizes, pos, excl, systems, cut)
        avg = self.avg_edges
        while True:
            cap = max(avg * lh.n, 1)
            ij = torch.empty((cap, 2), dtype=torch.int32, device=dev)
            sh = torch.empty((cap, 3), dtype=torch.int32, device=dev)
            self._launch(lib, args, cap, ij, sh, count)
            required = int(count.item())
            if required <= cap:
                self.avg_edges = avg
                return Edges(ij[:required], sh[:required].reshape(-1, 1, 3))
            if not self.fix:
                raise CapacityError(
                    f"neighbor list needs {required} edges, capacity is {cap}", required)
            avg = next_power_of_two(-(-required // max(lh.n, 1)))

    def _args(self, lh, n_sys, sys_off, sizes, pos, excl, systems, cut):
        return (_lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), lh.n, _lib.ptr(pos),
                _lib.ptr(excl), _lib.ptr(systems.tril), _lib.ptr(cut))

    def _query(self, lh: ParticleBatch, rh: ParticleBatch, systems: SystemBatch, cutoffs,
               rh_index_remap) -> Edges:
        """Edges between lh (the full table) and the rh particles
        (core/neighborlist.py:233-255): with ``rh_index_remap`` (lh index of
        every rh particle) both directions in lh indices, else (lh, rh)
        pairs with the second index into rh."""
        lib = _lib.load()
        n_sys = systems.n_sys
        dev = lh.positions.device
        sys_off, _ = system_offsets(lh.system, n_sys)
        excl = lh.exclusion
        if excl is None:
            excl = torch.arange(lh.n, dtype=torch.int32, device=dev)
        excl = excl.to(torch.int32).contiguous()
        pos = lh.positions.to(torch.float32).contiguous()
        k = rh.n
        rpos = rh.positions.to(torch.float32).reshape(-1, 3).contiguous()
        rsys = rh.system.to(torch.int32).contiguous()
        remap = None
        if rh_index_remap is not None:
            remap = torch.as_tensor(rh_index_remap, device=dev).to(torch.int32).contiguous()
            if remap.numel() != k:
                raise ValueError(f"{k} rh particles but {remap.numel()} remap entries")
        rexcl = rh.exclusion
        if rexcl is None:
            # an rh particle is excluded with the lh particle it stands for
            rexcl = remap if remap is not None else (
                torch.arange(k, dtype=torch.int32, device=dev) + lh.n)
        rexcl = rexcl.to(torch.int32).contiguous()
        cut = _cutoffs(cutoffs, n_sys, dev)
        count = torch.zeros(1, dtype=torch.int64, device=dev)
        avg = self.avg_query_edges
        while True:
            cap = max(avg * max(k, 1) * 2, 1)
            ij = torch.empty((cap, 2), dtype=torch.int32, device=dev)
            sh = torch.empty((cap, 3), dtype=torch.int32, device=dev)
            _lib.check(lib.kupsmi_neighborlist_query(
                _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), lh.n, _lib.ptr(pos),
                _lib.ptr(excl), k, _lib.ptr(rpos), _lib.ptr(rsys), _lib.ptr(rexcl),
                _lib.ptr(remap), _lib.ptr(systems.tril), _lib.ptr(cut), cap, _lib.ptr(ij),
                _lib.ptr(sh), _lib.ptr(count)))
            required = int(count.item())
            if required <= cap:
                self.avg_query_edges = avg
                return Edges(ij[:required], sh[:required].reshape(-1, 1, 3))
            if not self.fix:
                raise CapacityError(
                    f"neighbor query needs {required} edges, capacity is {cap}", required)
            avg = next_power_of_two(-(-required // max(2 * k, 1)))


@dataclass
class CellListNeighborList(DenseNearestNeighborList):
    """Cell-list neighbor list (core/neighborlist.py:985-1077).  The full
    list comes from ``kupsmi_neighborlist_cells``: atoms counting-sorted into
    cells at least rc wide, each atom testing the distinct cells among its 27
    neighbours (the reference's _cell_list_subselect, 295-364) with the same
    FP64 pair decision, so the edge set equals the dense list's; systems hold
    at most 16384 atoms.  rh queries scan the rh particle's system (small
    patches, as in MC moves)."""

    def _args(self, lh, n_sys, sys_off, sizes, pos, excl, systems, cut):
        mx = int(sizes.max()) if len(sizes) else 0
        return (_lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), lh.n, mx,
                _lib.ptr(pos), _lib.ptr(excl), _lib.ptr(systems.tril), _lib.ptr(cut))

    def _launch(self, lib, args, cap, ij, sh, count):
        _lib.check(lib.kupsmi_neighborlist_cells(*args, cap, _lib.ptr(ij), _lib.ptr(sh),
                                                 _lib.ptr(count)))


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
    up to ``compaction`` x an edge buffer of ``avg_changes`` (initially
    ``avg_edges``) per particle for n + k particles; more raises
    CapacityError or, with ``neighborlist.fix``, regrows ``avg_changes``
    (next power of two, the full list's buffer is left alone) and
    recomputes."""
    lib = _lib.load()
    dev = lh.positions.device
    idx = rh_indices.to(torch.int32).contiguous()
    new = rh_positions.to(torch.float32).reshape(-1, 3).contiguous()
    k = idx.numel()
    if new.shape[0] != k:
        raise ValueError(f"{k} moved particles but {new.shape[0]} new positions")
    sys_off, pos, excl, cut = _inputs(lh, systems, cutoffs)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    avg = neighborlist.avg_changes or neighborlist.avg_edges
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
            neighborlist.avg_changes = avg
            return NeighborListChangesResult(
                Edges(a_ij[:n_added], a_sh[:n_added].reshape(-1, 1, 3)),
                Edges(r_ij[:n_removed], r_sh[:n_removed].reshape(-1, 1, 3)))
        if not neighborlist.fix:
            raise CapacityError(
                f"neighborlist_changes needs {max(n_added, n_removed)} edges, capacity is {cap}",
                max(n_added, n_removed))
        avg = next_power_of_two(-(-max(n_added, n_removed) // max(int((lh.n + k) * compaction), 1)))
