"""Hop targets, group ids and the saddle list from densities and neighbour rows -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by the product).

A vectorised NumPy restatement of two stages of the reference's fortran/compute_adaptahop.f90,
given every particle's density and its first nhop neighbours (0-based ids in ascending
distance, as omp_find_nearest_parts returns them, f90:2841-2962):

* ``hop_targets`` -- find_local_maxima's hop rule (f90:774-795): an active particle
  (density > rho_threshold) hops to the densest of itself and its first nhop neighbours;
  a strictly greater density replaces, an equal one keeps the smaller index, so the target is
  the argmax of density with ties going to the smallest index.  Inactive particles get -1.
* ``group_ids`` -- the roots (target == self) numbered 1.. in ascending particle index
  (f90:797-803), every active particle labelled with the root its pointer chain reaches
  (f90:819-845; the chain strictly climbs in (density, -index), so it ends at a root).
* ``saddle_pairs`` -- dev_compute_saddle_list (f90:2491-2688): for every active particle i of
  group g1 and each of its first nhop neighbours j in a group g2 other than 0 and g1, the
  directed saddle s(g1->g2) is the largest min(rho_i, rho_j); a pair survives only if the
  reverse pair exists, both directions then holding the smaller of the two.  Returned as
  (g1, g2, s) sorted by (g1, g2): the multiset per group that the reference's per-group
  quicksort (indexx, f90:2702-2791) orders by decreasing s.

Pinned in tests/test_oracle.py against the C restatement (orc_maxima, orc_saddle), which is
itself pinned bit for bit to the compiled reference, on full inputs with ties and duplicate
positions.  These functions are for inputs too large for the per-particle C restatement of the
saddle stage (which re-runs a kNN query for every member twice, as the reference does).
"""
from __future__ import annotations

import numpy as np


def hop_targets(rho, nbr, rho_threshold):
    """0-based hop target of every particle (-1 when density <= rho_threshold).  nbr: (N, m)
    0-based neighbour ids; the first nhop = m columns are scanned."""
    rho = np.asarray(rho, dtype=np.float64)
    nbr = np.asarray(nbr)
    n = rho.shape[0]
    hop = np.full(n, -1, dtype=np.int64)
    act = np.nonzero(rho > rho_threshold)[0]
    if act.size == 0:
        return hop.astype(np.int32)
    for lo in range(0, act.size, 1 << 20):
        a = act[lo:lo + (1 << 20)]
        cand = np.concatenate([a[:, None], nbr[a].astype(np.int64)], axis=1)
        rc = rho[cand]
        best = rc.max(axis=1)
        # among the candidates at the maximum density, the smallest index
        hop[a] = np.where(rc == best[:, None], cand, np.iinfo(np.int64).max).min(axis=1)
    return hop.astype(np.int32)


def group_ids(hop):
    """1-based group id of every particle (0 for hop == -1) and the number of groups."""
    hop = np.asarray(hop, dtype=np.int64)
    act = hop >= 0
    root = hop.copy()
    # pointer doubling: every active particle reaches its chain's root
    while True:
        nxt = np.where(act, root[np.maximum(root, 0)], -1)
        if np.array_equal(nxt, root):
            break
        root = nxt
    roots = np.nonzero(act & (hop == np.arange(hop.shape[0])))[0]
    label = np.zeros(hop.shape[0], dtype=np.int64)
    label[roots] = np.arange(1, roots.size + 1)
    gid = np.where(act, label[np.maximum(root, 0)], 0)
    return gid.astype(np.int32), int(roots.size)


def saddle_pairs(rho, gid, nbr):
    """Symmetrised saddle pairs: (g1, g2, s) arrays sorted by (g1, g2), both directions
    present.  gid: 1-based group ids (0 = none); nbr (N, nhop) 0-based neighbour ids."""
    rho = np.asarray(rho, dtype=np.float64)
    gid = np.asarray(gid, dtype=np.int64)
    nbr = np.asarray(nbr)
    ng = int(gid.max()) if gid.size else 0
    keys, vals = [], []
    members = np.nonzero(gid > 0)[0]
    for lo in range(0, members.size, 1 << 20):
        i = members[lo:lo + (1 << 20)]
        j = nbr[i].astype(np.int64)
        g1 = np.broadcast_to(gid[i][:, None], j.shape)
        g2 = gid[j]
        keep = (g2 > 0) & (g2 != g1)
        if not keep.any():
            continue
        s = np.minimum(np.broadcast_to(rho[i][:, None], j.shape), rho[j])
        keys.append(g1[keep] * (ng + 1) + g2[keep])
        vals.append(s[keep])
    if not keys:
        e = np.zeros(0, dtype=np.int64)
        return e.astype(np.int32), e.astype(np.int32), np.zeros(0)
    key = np.concatenate(keys)
    val = np.concatenate(vals)
    order = np.lexsort((val, key))
    key, val = key[order], val[order]
    last = np.r_[key[1:] != key[:-1], True]
    key, val = key[last], val[last]          # the largest s of each directed pair
    g1, g2 = key // (ng + 1), key % (ng + 1)
    rev = g2 * (ng + 1) + g1
    pos = np.searchsorted(key, rev)
    pos_c = np.minimum(pos, key.size - 1)
    has = key[pos_c] == rev
    s = np.minimum(val, val[pos_c])
    return g1[has].astype(np.int32), g2[has].astype(np.int32), s[has]


def csr_as_pairs(offset, isad, rho):
    """A per-group saddle CSR (0-based offsets, 1-based neighbour ids, densities) as (g1, g2, s)
    sorted by (g1, g2), for comparison with :func:`saddle_pairs`."""
    offset = np.asarray(offset, dtype=np.int64)
    g1 = np.repeat(np.arange(1, offset.shape[0], dtype=np.int64), np.diff(offset))
    g2 = np.asarray(isad, dtype=np.int64)
    order = np.lexsort((g2, g1))
    return g1[order].astype(np.int32), g2[order].astype(np.int32), np.asarray(rho)[order]
