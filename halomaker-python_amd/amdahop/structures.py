"""Structures from the drop-in's outputs: the MSM selection and the member grouping.

The reference turns the node tables of ``neikdtree`` into halos and subhalos in Python,
``select_with_MS_method_1319`` (src/compute_neiKDtree_mod.py:260-464) or, for method HOP,
``select_halos_1317`` (:228-252), and then groups the
particles by structure (compute_adaptahop_131, :109-119).  Here the selection works on the
node tables (a few thousand nodes: host work) as whole-array NumPy passes: the sister chains
become (owner, subnode) pairs in chain order, own masses are a segmented left fold, the main
branches are a segmented first maximum taken level by level from the deepest nodes up, the
structures are propagated level by level from the top down, and the structure tree's links
come from the sorted (mother structure, structure) pairs.  Only the outcomes the reference's
loop orders decide are kept from it: the subtraction order of truemass, the first subnode in
sister order on equal truemass, structures numbered by their first node, halos and daughters
listed in structure order.  The per-particle part (relabelling every particle, then the
argsort / np.unique grouping) runs on the device through ``ahop_group_members_device``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

# node record of the reference (halo_defs.py:555-560); real_table holds the 8 float fields and
# integer_table the 6 integer fields, in this order
NODE_DTYPE = np.dtype([
    ("rho_saddle", "f8"), ("density", "f8"), ("densmax", "f8"), ("radius", "f8"),
    ("truemass", "f8"), ("px", "f8"), ("py", "f8"), ("pz", "f8"),
    ("mass", "i4"), ("level", "i4"), ("mother", "i4"), ("firstchild", "i4"),
    ("nsisters", "i4"), ("sister", "i4"),
])
_REAL = ["rho_saddle", "density", "densmax", "radius", "truemass", "px", "py", "pz"]
_INT = ["mass", "level", "mother", "firstchild", "nsisters", "sister"]


def node_table(real_table, integer_table):
    """The reference's ``node_0`` (compute_neiKDtree_mod.py:55-60): nodes 1..nnodes of the
    (8, nnodes) and (6, nnodes) tables, entry 0 unused."""
    rt = np.asarray(real_table, dtype=np.float64)
    it = np.asarray(integer_table, dtype=np.int32)
    if rt.ndim != 2 or it.ndim != 2 or rt.shape[0] != 8 or it.shape[0] != 6 \
            or rt.shape[1] != it.shape[1]:
        raise ValueError("real_table must be (8, nnodes) and integer_table (6, nnodes)")
    node_0 = np.zeros(rt.shape[1] + 1, dtype=NODE_DTYPE)
    for k, name in enumerate(_REAL):
        node_0[name][1:] = rt[k]
    for k, name in enumerate(_INT):
        node_0[name][1:] = it[k]
    return node_0


@dataclass
class MSMStructures:
    """Result of the MSM selection.  ``node_to_struct[inode - 1]``: 1-based structure of node
    ``inode``; per structure (0-based arrays over 1-based ids, -1 = none): ``mother``
    (0 for a halo), ``first_sister``, ``first_daughter``, ``level`` (1 for a halo)."""
    node_to_struct: np.ndarray
    nb_halos: int
    nb_sub: int
    mother: np.ndarray
    first_sister: np.ndarray
    first_daughter: np.ndarray
    level: np.ndarray

    @property
    def nstruct(self):
        return self.nb_halos + self.nb_sub


def _chain_positions(nxt):
    """Heads of the linked lists given by ``nxt`` (1-based successor of each entry, 0 = end;
    index 0 unused) and every entry's distance from its head, by pointer doubling on the
    predecessor links."""
    n1 = nxt.shape[0]
    ids = np.arange(n1)
    pred = ids.copy()                   # a head points to itself
    has = nxt > 0
    pred[nxt[has]] = ids[has]
    dist = (pred != ids).astype(np.int64)
    while True:
        far = pred[pred]
        if np.array_equal(far, pred):
            return pred, dist
        dist = dist + dist[pred]
        pred = far


def _depths(up):
    """Depth of every entry of the forest given by ``up`` (1-based parent, 0 = top; index 0
    unused): 0 at the top, by relaxation from the top down."""
    depth = np.zeros(up.shape[0], dtype=np.int64)
    inner = up > 0
    for _ in range(up.shape[0]):
        nd = np.where(inner, depth[up] + 1, 0)
        if np.array_equal(nd, depth):
            return depth
        depth = nd
    raise ValueError("the node links contain a cycle")


def _children(node_0, nnodes):
    """Every node's subnodes in the order of its sister chain (firstchild, then sister
    links, as the reference walks them): (owner, child) pairs sorted by owner and chain
    position.  Top-level nodes, which hang off the unexported root, have no owner."""
    fc = node_0["firstchild"].astype(np.int64)
    sis = node_0["sister"].astype(np.int64)
    if np.any(fc[1:] > nnodes) or np.any(sis[1:] > nnodes) or np.any(sis[1:] < 0):
        raise ValueError("firstchild / sister links outside the node table")
    head, pos = _chain_positions(np.where(np.arange(nnodes + 1) > 0, sis, 0))
    owner_of_head = np.zeros(nnodes + 1, dtype=np.int64)
    starts = np.nonzero((fc > 0) & (np.arange(nnodes + 1) > 0))[0]
    owner_of_head[fc[starts]] = starts
    owner = owner_of_head[head]
    owner[0] = 0
    kid = np.nonzero(owner > 0)[0]
    order = np.lexsort((pos[kid], owner[kid]))
    return owner[kid][order], kid[order], owner


def select_with_ms_method(node_0, strict=False):
    """MSM selection (select_with_MS_method_1319, compute_neiKDtree_mod.py:260-464) on a
    ``node_0`` table (see :func:`node_table`), updated in place with the reference's results:
    each node first keeps only its own particles (its inclusive mass and truemass less its
    subnodes', :266-280; truemass subtracted subnode by subnode in sister order, which fixes
    the rounding), then, from the deepest nodes up, absorbs its most massive subnode's
    main-branch mass and truemass and takes its densmax (:300-319; the first of the largest
    truemass in sister order).  A structure starts at every halo (mother <= 0) and at every
    subnode that is not its mother's most massive; the others continue their mother's
    structure (:321-346).  Structures are numbered by the node id of their first node, and the
    structure tree lists halos and each structure's daughters in that order (:398-429).
    ``strict`` adds a count check (:349-353).  Raises ValueError where the reference does (a
    node left with no particles of its own).  The table must list every mother before its
    subnodes, as create_group_tree's deterministic reorder does (levels ascending)."""
    node_0 = node_0.view(NODE_DTYPE) if node_0.dtype != NODE_DTYPE else node_0
    nnodes = len(node_0) - 1
    if nnodes < 1:
        raise ValueError("no nodes")
    owner_k, kid, owner = _children(node_0, nnodes)
    mother = node_0["mother"].astype(np.int64)
    mother[0] = 0
    ids = np.arange(nnodes + 1)
    if np.any(owner_k >= kid) or np.any((mother > 0) & (mother >= ids)) or np.any(mother > nnodes):
        raise ValueError("node tables must list every mother before its subnodes")

    # own particles: [inclusive, subnode 1, subnode 2, ...] folded left by subtraction per
    # owner (ufunc reduceat folds left, the order of the reference's in-place updates)
    owners = np.unique(owner_k)
    seg_val_m = np.concatenate([node_0["mass"][owners].astype(np.int64),
                                node_0["mass"][kid].astype(np.int64)])
    seg_val_t = np.concatenate([node_0["truemass"][owners], node_0["truemass"][kid]])
    seg_key = np.concatenate([owners, owner_k])
    seg_rank = np.concatenate([np.full(owners.shape[0], -1), np.arange(kid.shape[0])])
    seg = np.lexsort((seg_rank, seg_key))
    first = np.r_[0, np.nonzero(np.diff(seg_key[seg]))[0] + 1] if seg.size else np.zeros(0, int)
    own_m = node_0["mass"].astype(np.int64)
    own_t = node_0["truemass"].astype(np.float64)
    if owners.size:
        own_m[owners] = np.subtract.reduceat(seg_val_m[seg], first)
        own_t[owners] = np.subtract.reduceat(seg_val_t[seg], first)
        bad = owners[(own_m[owners] <= 0) | (own_t[owners] <= 0.0)]
        if bad.size:
            b = int(bad[0])
            raise ValueError(f"node {b} has mass {own_m[b]}, truemass {own_t[b]} after "
                             f"removing its subnodes")

    # main branches, deepest owners first: a node's main-branch truemass is its own plus that
    # of its most massive subnode (the first maximum in sister order), the same for mass,
    # and densmax is taken from that subnode
    tot_m, tot_t = own_m.copy(), own_t.copy()
    dmax = node_0["densmax"].astype(np.float64)
    best = np.zeros(nnodes + 1, dtype=np.int64)
    depth = _depths(owner)
    kid_depth = depth[kid]
    for d in range(int(depth.max()) if nnodes else 0, 0, -1):
        sel = np.nonzero(kid_depth == d)[0]          # sister-ordered subnodes at depth d
        if sel.size == 0:
            continue
        o, c = owner_k[sel], kid[sel]
        bounds = np.r_[0, np.nonzero(np.diff(o))[0] + 1]
        seg_max = np.maximum.reduceat(tot_t[c], bounds)
        lens = np.diff(np.r_[bounds, sel.size])
        at_max = tot_t[c] == np.repeat(seg_max, lens)
        pick = np.minimum.reduceat(np.where(at_max, np.arange(sel.size), sel.size), bounds)
        mo, mc = o[bounds], c[pick]
        best[mo] = mc
        tot_m[mo] = own_m[mo] + tot_m[mc]
        tot_t[mo] = own_t[mo] + tot_t[mc]
        dmax[mo] = dmax[mc]
    node_0["mass"][1:] = tot_m[1:]
    node_0["truemass"][1:] = tot_t[1:]
    node_0["densmax"][1:] = dmax[1:]

    # structures: numbered by their first node; a node that is its mother's most massive
    # subnode continues its mother's structure, every other node starts one
    is_halo = mother <= 0
    continues = (~is_halo) & (best[np.where(is_halo, 0, mother)] == ids)
    continues[0] = False
    starts_s = ~continues
    starts_s[0] = False
    sid = np.cumsum(starts_s) * starts_s
    struct = np.where(starts_s, sid, 0)
    mdepth = _depths(mother)
    for d in range(1, int(mdepth.max()) + 1):
        lvl = np.nonzero((mdepth == d) & continues)[0]
        struct[lvl] = struct[mother[lvl]]
    node_to_struct = struct[1:].astype(np.int32)
    nb_halos = int(np.count_nonzero(is_halo[1:]))
    nb_sub = int(np.count_nonzero(starts_s)) - nb_halos
    nstruct = nb_halos + nb_sub
    if np.any(node_to_struct <= 0) or (strict and nstruct != int(node_to_struct.max())):
        raise ValueError("error in the structure count")

    # structure tree from the (mother structure, structure) pairs, in structure order:
    # halos chained one after the other, daughters listed under their mother
    first_node = np.nonzero(starts_s)[0]                 # ascending, so structure order
    smother = np.where(is_halo[first_node], 0, struct[np.where(is_halo, 0, mother)][first_node])
    smother = smother.astype(np.int64)
    up = np.r_[0, smother]
    level = (_depths(up)[1:] + 1).astype(np.int32)
    first_sister = np.full(nstruct, -1, dtype=np.int32)
    first_daughter = np.full(nstruct, -1, dtype=np.int32)
    order = np.lexsort((np.arange(1, nstruct + 1), smother))
    sm_sorted = smother[order]
    same = sm_sorted[1:] == sm_sorted[:-1]
    first_sister[order[:-1][same]] = (order[1:][same] + 1).astype(np.int32)
    lead = np.r_[True, ~same]
    heads = order[lead]
    sub_heads = heads[smother[heads] > 0]
    first_daughter[smother[sub_heads] - 1] = (sub_heads + 1).astype(np.int32)
    return MSMStructures(node_to_struct, nb_halos, nb_sub, smother.astype(np.int32),
                         first_sister, first_daughter, level)


def select_halos(node_0):
    """HOP selection (select_halos_1317, compute_neiKDtree_mod.py:228-252): every node
    belongs to its top-level ancestor (the node reached by following mothers up to one with
    mother 0), whose node id labels its particles.  Returns (node_to_halo (nnodes,) int32,
    nb_halos), nb_halos counting the nodes with mother <= 0.  A label outside 1..nb_halos is
    a ValueError, as in the reference's relabelling (:246-249)."""
    node_0 = node_0.view(NODE_DTYPE) if node_0.dtype != NODE_DTYPE else node_0
    nnodes = len(node_0) - 1
    mother = node_0["mother"].astype(np.int64)
    mother[0] = 0
    if np.any(mother < 0) or np.any(mother > nnodes):
        raise ValueError("mother links outside the node table")
    depth = _depths(mother)
    top = np.arange(nnodes + 1)
    for d in range(1, int(depth.max()) + 1 if nnodes else 1):
        lvl = np.nonzero(depth == d)[0]
        top[lvl] = top[mother[lvl]]
    node_to_halo = top[1:].astype(np.int32)
    nb_halos = int(np.count_nonzero(node_0["mother"][1:] <= 0))
    if np.any(node_to_halo > nb_halos) or np.any(node_to_halo <= 0):
        raise ValueError("ihalo > nb_of_halos or nil")
    return node_to_halo, nb_halos


def group_members_device(node, labels, pids, offsets, node_to_struct=None, nnodes=None,
                         nstruct=None):
    """Device form of :func:`group_members`: every argument a CUDA tensor already in place
    (``node``, ``labels``, ``pids`` (N,) int32; ``offsets`` (nlab + 2,) int64, where nlab is
    ``nstruct`` with a map and ``nnodes`` without).  Writes the labels, the grouped particle
    indices and ``offsets[l]`` = first grouped position of label l (so label l holds
    ``pids[offsets[l]:offsets[l + 1]]``); returns the device time in ms."""
    n = int(node.shape[0])
    nn = int(node_to_struct.shape[0]) if node_to_struct is not None else int(nnodes)
    nlab = int(nstruct) if node_to_struct is not None else nn
    for t, name, dt, shape in ((node, "node", "int32", (n,)), (labels, "labels", "int32", (n,)),
                               (pids, "pids", "int32", (n,)),
                               (offsets, "offsets", "int64", (nlab + 2,))):
        if not t.is_cuda or str(t.dtype) != f"torch.{dt}" or tuple(t.shape) != shape \
                or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous CUDA {dt} tensor of shape {shape}")
    import torch
    # the library runs on its own stream: finish torch's work on the inputs first
    torch.cuda.synchronize(node.device)
    ms = ctypes.c_float(0.0)
    check(lib().ahop_group_members_device(
        ctypes.c_void_p(node.data_ptr()), n,
        ctypes.c_void_p(node_to_struct.data_ptr() if node_to_struct is not None else None),
        nn, nlab, ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(pids.data_ptr()),
        ctypes.c_void_p(offsets.data_ptr()), ctypes.byref(ms)))
    return ms.value


def group_members(node_of_part, node_to_struct=None, nnodes=None):
    """Relabel particles by structure and group them (ahop_group_members_device).

    ``node_of_part`` (N,) int32: liste_parts (1-based node ids, 0 = in no node), host or
    device.  ``node_to_struct`` (nnodes,) int32 from :func:`select_with_ms_method`, or None to
    group by node id (``nnodes`` then gives the label range).  Returns host arrays in the
    reference's layout (compute_neiKDtree_mod.py:109-119): the relabelled ``whereIam_parts``,
    ``pids0_groupsorted`` (particle indices grouped by label; ascending within a label, where
    the reference's unstable argsort leaves the order open), and the present labels with their
    first index and count (``np.unique(..., return_index, return_counts)``)."""
    import torch

    dev = torch.device("cuda")
    node = torch.as_tensor(node_of_part).to(device=dev, dtype=torch.int32).contiguous()
    n = int(node.shape[0])
    if node.dim() != 1 or n < 1:
        raise ValueError("node_of_part must be a non-empty (N,) array")
    if node_to_struct is not None:
        mp = torch.as_tensor(node_to_struct).to(device=dev, dtype=torch.int32).contiguous()
        nn = int(mp.shape[0])
        nlab = int(mp.max().item()) if nn else 0
    else:
        mp = None
        if nnodes is None:
            nnodes = int(node.max().item())
        nn = nlab = int(nnodes)
    labels = torch.empty(n, dtype=torch.int32, device=dev)
    pids = torch.empty(n, dtype=torch.int32, device=dev)
    offsets = torch.empty(nlab + 2, dtype=torch.int64, device=dev)
    ms = group_members_device(node, labels, pids, offsets, node_to_struct=mp, nnodes=nn,
                              nstruct=nlab)
    off = offsets.cpu().numpy()
    counts = np.diff(off)[:nlab + 1]
    present = np.nonzero(counts > 0)[0]
    return (labels.cpu().numpy(), pids.cpu().numpy(), present.astype(np.int32),
            off[present].astype(np.int32), counts[present].astype(np.int32), ms)
