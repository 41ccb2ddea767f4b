"""Structures from the drop-in's outputs: the MSM selection and the member grouping.

The reference turns the node tables of ``neikdtree`` into halos and subhalos in Python,
``select_with_MS_method_1319`` (src/compute_neiKDtree_mod.py:260-464) or, for method HOP,
``select_halos_1317`` (:228-252), and then groups the
particles by structure (compute_adaptahop_131, :109-119).  Here the selection is restated on
the node tables in NumPy (a few thousand nodes: host work, in the reference's order), and the
per-particle part (relabelling every particle, then the argsort / np.unique grouping) runs on
the device through ``ahop_group_members_device``.
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


def select_with_ms_method(node_0, strict=False):
    """MSM selection (select_with_MS_method_1319, compute_neiKDtree_mod.py:260-464) on a
    ``node_0`` table (see :func:`node_table`), updated in place as the reference does: node
    masses first lose their subnodes' masses (:266-280), then every node absorbs its most
    massive subnode (by truemass, first on ties) and takes its densmax (:302-319).  Each halo
    (mother <= 0) and each subnode that is not its mother's most massive starts a structure;
    a most massive subnode joins its mother's (:322-346).  Then the structure tree (:398-429).
    ``strict`` adds the reference's count check (:349-353).  Raises ValueError where the
    reference does (a node left with no particles of its own)."""
    node_0 = node_0.view(NODE_DTYPE) if node_0.dtype != NODE_DTYPE else node_0
    nnodes = len(node_0) - 1
    if nnodes < 1:
        raise ValueError("no nodes")
    # own masses: subtract the subnodes' (inclusive) masses, nodes in ascending order (:266-280)
    where = np.nonzero(node_0["firstchild"] > 0)[0]
    if len(where) and where[0] == 0:
        where = where[1:]
    for inode1 in where:
        isub1 = int(node_0["firstchild"][inode1])
        while isub1 > 0:
            node_0["mass"][inode1] -= node_0["mass"][isub1]
            node_0["truemass"][inode1] -= node_0["truemass"][isub1]
            isub1 = int(node_0["sister"][isub1])
        if node_0["mass"][inode1] <= 0 or node_0["truemass"][inode1] <= 0.0:
            raise ValueError(f"node {inode1} has mass {node_0['mass'][inode1]}, truemass "
                             f"{node_0['truemass'][inode1]} after removing its subnodes")
    # most massive subnode of every node, nodes in descending order (:300-319)
    mostmasssub = np.full(nnodes, -1, dtype=np.int32)
    firstchild = node_0["firstchild"]
    for inode1 in range(nnodes, 0, -1):
        isub1 = int(firstchild[inode1])
        if isub1 > 0:
            maxmasssub = node_0["truemass"][isub1]
            mostmasssub[inode1 - 1] = isub1
            isub1 = int(node_0["sister"][isub1])
        while isub1 > 0:
            if node_0["truemass"][isub1] > maxmasssub:
                maxmasssub = node_0["truemass"][isub1]
                mostmasssub[inode1 - 1] = isub1
            isub1 = int(node_0["sister"][isub1])
        if mostmasssub[inode1 - 1] > 0:
            m = int(mostmasssub[inode1 - 1])
            node_0["mass"][inode1] += node_0["mass"][m]
            node_0["truemass"][inode1] += node_0["truemass"][m]
            node_0["densmax"][inode1] = node_0["densmax"][m]
    # structure of every node (:321-346): starters are halos and non-most-massive subnodes;
    # a most massive subnode points to its mother, and pointer jumping finds the starter
    mothers = node_0["mother"][1:nnodes + 1].astype(np.int64)
    indices = np.arange(1, nnodes + 1)
    is_halo = mothers <= 0
    valid_mother_idx = np.where(~is_halo, mothers - 1, 0)
    is_most_mass = (~is_halo) & (indices == mostmasssub[valid_mother_idx])
    is_new_sub = (~is_halo) & (~is_most_mass)
    nb_halos = int(np.sum(is_halo))
    nb_sub = int(np.sum(is_new_sub))
    is_starter = is_halo | is_new_sub
    starter_ids = np.zeros(nnodes, dtype=np.int64)
    starter_ids[is_starter] = np.cumsum(is_starter)[is_starter]
    parent = indices - 1
    parent[is_most_mass] = mothers[is_most_mass] - 1
    while True:
        new_parent = parent[parent]
        if np.array_equal(parent, new_parent):
            break
        parent = new_parent
    node_to_struct = starter_ids[parent].astype(np.int32)
    nstruct = nb_halos + nb_sub
    if strict and nstruct != int(starter_ids.max()):
        raise ValueError(f"structure count {nstruct} != {int(starter_ids.max())}")
    # structure tree, nodes in ascending order (:398-429)
    mother = np.full(nstruct, -1, dtype=np.int32)
    first_sister = np.full(nstruct, -1, dtype=np.int32)
    first_daughter = np.full(nstruct, -1, dtype=np.int32)
    level = np.zeros(nstruct, dtype=np.int32)
    ihalo1 = -1
    for inode0 in range(nnodes):
        istruct1 = int(node_to_struct[inode0])
        if istruct1 <= 0:
            raise ValueError("node without a structure")
        if mother[istruct1 - 1] < 0:
            if node_0["mother"][inode0 + 1] <= 0:      # a main halo
                mother[istruct1 - 1] = 0
                if ihalo1 > 0:
                    first_sister[ihalo1 - 1] = istruct1
                ihalo1 = istruct1
                level[istruct1 - 1] = 1
            else:                                        # a subhalo
                imother1 = int(node_to_struct[node_0["mother"][inode0 + 1] - 1])
                mother[istruct1 - 1] = imother1
                level[istruct1 - 1] = level[imother1 - 1] + 1
                if first_daughter[imother1 - 1] <= 0:
                    first_daughter[imother1 - 1] = istruct1
                else:                                    # last of its mother's daughters
                    isub1 = int(first_daughter[imother1 - 1])
                    while first_sister[isub1 - 1] > 0:
                        isub1 = int(first_sister[isub1 - 1])
                    first_sister[isub1 - 1] = istruct1
    return MSMStructures(node_to_struct, nb_halos, nb_sub, mother, first_sister,
                         first_daughter, level)


def select_halos(node_0):
    """HOP selection (select_halos_1317, compute_neiKDtree_mod.py:228-252): every node
    belongs to its top-level ancestor (mother 0), found by walking mothers and reusing the
    ancestors already found, as the reference does.  Returns (node_to_halo (nnodes,) int32,
    nb_halos).  Top-level node ids must be 1..nb_halos, else ValueError as the reference's
    relabelling raises (:246-249): the particle labels are these node ids."""
    node_0 = node_0.view(NODE_DTYPE) if node_0.dtype != NODE_DTYPE else node_0
    nnodes = len(node_0) - 1
    mother = node_0["mother"]
    node_to_halo = np.full(nnodes, -1, dtype=np.int32)
    nb_halos = 0
    for inode1 in range(1, nnodes + 1):
        if mother[inode1] <= 0:
            nb_halos += 1
        ih1 = inode1
        while mother[ih1] != 0:
            ih1 = int(mother[ih1]) if node_to_halo[ih1 - 1] <= 0 else int(node_to_halo[ih1 - 1])
        node_to_halo[inode1 - 1] = ih1
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
