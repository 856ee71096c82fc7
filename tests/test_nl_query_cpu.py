"""Pins of the oracle's rh-subset neighbor query (CPU only).

The query follows NearestNeighborList.__call__(lh, rh, systems, cutoffs,
rh_index_remap) (core/neighborlist.py:233-255, 505-539, 680-685).  Pinned by
the reference's own rh cases (test/core/test_neighborlist.py:402-676, with the
reference's test convention: lh exclusion ids = particle index, rh exclusion
ids = the remap) and by an identity that ties the query to the full list
(itself pinned by brute force in test_oracle.py): querying a subset of the
particles at their own positions, with the remap naming them, returns every
directed edge of the full list that has an end in the subset, once, plus a
second copy of each self-image edge of a subset particle.
"""

from collections import Counter

import numpy as np

import oracle
from kups_amd import workloads


def _q(lh_pos, lh_sys, rh_pos, rh_sys, remap, cutoffs, L=10.0, rh_excl=None, n_sys=None):
    lh_pos = np.asarray(lh_pos, float)
    lh_sys = np.asarray(lh_sys)
    n_sys = n_sys or int(lh_sys.max()) + 1
    off = np.concatenate([[0], np.cumsum(np.bincount(lh_sys, minlength=n_sys))])
    tril = np.tile([L, 0, L, 0, 0, L], (n_sys, 1)).astype(float)
    remap = None if remap is None else np.asarray(remap)
    if rh_excl is None:
        rh_excl = remap
    ij, sh = oracle.neighborlist_query(lh_pos, np.arange(len(lh_pos)), off, tril,
                                       np.asarray(cutoffs, float), rh_pos, rh_sys, rh_excl,
                                       remap)
    return [(int(i), int(j)) for i, j in ij], sh


def _valid(edges, n):
    return {e for e in edges if e[0] < n and e[1] < n}


def test_rh_positions_basic():
    # test_neighborlist.py:402-434
    e, _ = _q([[0, 0, 0], [2, 0, 0], [5, 0, 0]], [0, 0, 0], [[1, 0, 0], [3, 0, 0]], [0, 0],
              [2, 0], [1.5])
    v = _valid(e, 3)
    assert {(1, 2), (2, 1), (1, 0), (0, 1)} <= v


def test_rh_batch_boundaries():
    # test_neighborlist.py:436-502
    lh = [[0, 0, 0], [5, 0, 0], [0, 0, 0], [5, 0, 0]]
    sysm = np.array([0, 0, 1, 1])
    e, _ = _q(lh, sysm, [[1, 0, 0], [1, 0, 0]], [0, 1], [1, 3], [1.5, 1.5])
    v = _valid(e, 4)
    assert v and all(sysm[a] == sysm[b] for a, b in v)
    e, _ = _q(lh, sysm, [[0.5, 0, 0], [0.5, 0, 0]], [1, 0], [3, 1], [1.5, 1.5])
    assert all(sysm[a] == sysm[b] for a, b in _valid(e, 4))


def test_rh_positions_asymmetric_search():
    # test_neighborlist.py:504-547
    e, _ = _q([[0, 0, 0], [2, 0, 0], [4, 0, 0]], [0, 0, 0], [[1, 0, 0], [3, 0, 0]], [0, 0],
              [0, 2], [1.5])
    v = _valid(e, 3)
    assert {(1, 0), (0, 1), (1, 2), (2, 1)} <= v


def test_rh_index_remap_for_subsets():
    # test_neighborlist.py:549-593 (remap entry 4 names no lh particle)
    pos = [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]]
    e, _ = _q(pos, [0, 0, 0, 0], pos, [0, 0, 0, 0], [0, 2, 1, 4], [2.5])
    assert _valid(e, 4) == {(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1), (1, 3), (3, 1),
                            (2, 3), (3, 2)}


def test_rh_index_remap_prevents_self_interaction():
    # test_neighborlist.py:595-625
    pos = [[0, 0, 0], [1, 0, 0]]
    e, _ = _q(pos, [0, 0], pos, [0, 0], [0, 1], [1.5])
    v = _valid(e, 2)
    assert v and all(a != b for a, b in v)


def test_rh_arguments_combined():
    # test_neighborlist.py:627-676, with the lh particles sorted by system
    # (original 0, 2 in system 0, original 1 in system 1; remap [2, 1, 0]
    # becomes [1, 2, 0])
    sysm = np.array([0, 0, 1])
    e, _ = _q([[0, 0, 0]] * 3, sysm, [[1, 0, 0]] * 3, [0, 1, 0], [1, 2, 0], [1.5, 1.5])
    v = _valid(e, 3)
    assert v and all(sysm[a] == sysm[b] for a, b in v)


def test_rh_without_remap_indexes_rh():
    # rh disjoint from lh: edges (i, r) with r indexing rh, one direction
    e, sh = _q([[0, 0, 0], [4, 0, 0]], [0, 0], [[1, 0, 0], [9.5, 0, 0]], [0, 0], None,
               [1.6], rh_excl=[-1, -1])
    assert sorted(e) == [(0, 0), (0, 1)]
    assert {tuple(s) for s in sh} == {(0, 0, 0), (1, 0, 0)} or {tuple(s) for s in sh} == {
        (0, 0, 0), (-1, 0, 0)}
    # the shift of (0, rh 1): f_0 - f_r = -0.95 rounds to -1, d = r_r - r_0 + s L = -0.5
    idx = e.index((0, 1))
    assert tuple(sh[idx]) == (-1, 0, 0)


def _multiset(ij, sh):
    return Counter((int(i), int(j), tuple(int(x) for x in s)) for (i, j), s in zip(ij, sh))


def test_subset_query_equals_full_list_edges_touching_subset():
    """Identity: query(rh = lh[idx] at the same positions, remap = idx)
    == full-list edges with an end in idx, each once, plus self-image edges
    of idx particles twice (the reference appends both directions of every
    candidate, and a self image's candidate set already holds both)."""
    rng = np.random.default_rng(3)
    b = workloads.batched_lj_boxes(sizes=[40, 64, 8], seed=5)
    pos = b.positions.astype(np.float64)
    excl = np.arange(len(pos))
    cut = np.full(b.n_sys, workloads.AR_CUTOFF)
    full_ij, full_sh = oracle.neighborlist(pos, excl, b.sys_off, b.tril, cut)
    idx = np.sort(np.concatenate([rng.choice(104, 20, replace=False), [104, 107]]))
    sys_of = np.searchsorted(b.sys_off, np.arange(len(pos)), side="right") - 1
    ij, sh = oracle.neighborlist_query(pos, excl, b.sys_off, b.tril, cut, pos[idx], sys_of[idx],
                                       excl[idx], idx)
    got = _multiset(ij, sh)
    s_idx = set(idx.tolist())
    want = Counter()
    for (i, j), s in zip(full_ij, full_sh):
        if i in s_idx or j in s_idx:
            key = (int(i), int(j), tuple(int(x) for x in s))
            want[key] += 2 if i == j else 1
    assert got == want
    # the 8-atom box is narrower than rc: its self images must appear
    assert any(i == j for (i, j, _) in want)
