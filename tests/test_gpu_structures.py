"""Particles grouped by structure on the GPU (ahop_group_members_device) against the
reference's NumPy grouping (compute_neiKDtree_mod.py:109-119: argsort + np.unique), and the
HOP and MSM selection + grouping chains on the drop-in's output."""
import numpy as np
import pytest

from amdahop import generate as G

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def numpy_grouping(labels):
    order = np.argsort(labels, kind="stable")
    unis, idxs, counts = np.unique(labels[order], return_index=True, return_counts=True)
    return order, unis, idxs, counts


def check_grouping(labels, got):
    lab, pids, unis, idxs, counts, _ = got
    order, u2, i2, c2 = numpy_grouping(labels)
    assert np.array_equal(lab, labels)
    # within a label the reference's order is open (unstable argsort); the device's is
    # ascending, i.e. that of a stable argsort
    assert np.array_equal(pids, order)
    assert np.array_equal(unis, u2) and np.array_equal(idxs, i2) and np.array_equal(counts, c2)


@pytest.mark.parametrize("n,nnodes,nstruct", [(1, 1, 1), (1000, 7, 4), (300_000, 5000, 1800)])
def test_grouping_matches_argsort_unique(gpu, n, nnodes, nstruct):
    from amdahop.structures import group_members
    rng = np.random.default_rng(n)
    log(f"grouping n={n} nnodes={nnodes}")
    node = rng.integers(0, nnodes + 1, size=n).astype(np.int32)
    node[rng.random(n) < 0.5] = 0                       # half the particles in no node
    # node ids as labels
    check_grouping(node, group_members(node, nnodes=nnodes))
    # through a node-to-structure map that leaves some structures empty
    mp = rng.integers(1, nstruct + 1, size=nnodes).astype(np.int32)
    labels = np.where(node > 0, mp[np.maximum(node, 1) - 1], 0).astype(np.int32)
    check_grouping(labels, group_members(node, node_to_struct=mp))


def test_out_of_range_ids_are_errors(gpu):
    from amdahop.structures import group_members
    node = np.array([0, 1, 2, 3], dtype=np.int32)
    with pytest.raises(RuntimeError, match="outside"):
        group_members(node, nnodes=2)
    with pytest.raises(RuntimeError, match="outside"):
        group_members(node, node_to_struct=np.array([1, 2, 0], dtype=np.int32))


def hosts_with_satellites(seed=21):
    """24 Plummer hosts (6000 particles, scale 0.02), each with three satellites (800
    particles, scale 0.002) at 0.03 from its centre, over 30 000 uniform particles: a box
    with substructure."""
    rng = np.random.default_rng(seed)

    def plummer(m, a):
        r = a / np.sqrt(rng.uniform(0.0, 1.0, m) ** (-2.0 / 3.0) - 1.0)
        r = np.minimum(r, 20 * a)
        v = rng.normal(size=(m, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True) * r[:, None]

    parts = [rng.uniform(-0.5, 0.5, size=(30_000, 3))]
    for c in rng.uniform(-0.4, 0.4, size=(24, 3)):
        parts.append(c + plummer(6000, 0.02))
        for _ in range(3):
            u = rng.normal(size=3)
            parts.append(c + 0.03 * u / np.linalg.norm(u) + plummer(800, 0.002))
    pos = np.concatenate(parts)
    pos = (pos + 0.5) % 1.0 - 0.5
    n = len(pos)
    return G.ParticleSet(np.ascontiguousarray(pos), np.full(n, 1.0 / n), 32, 16, "satellites", n)


def test_msm_and_grouping_on_the_dropin_output(gpu):
    """A box with substructure through the drop-in, then MSM on its node tables (the tables
    are bit-exact against the compiled reference elsewhere) and the device grouping: every
    structure's members number its final node mass (the reference's check,
    compute_neiKDtree_mod.py:381-392)."""
    from compute_adaptahop import neikdtree
    from amdahop.structures import group_members, node_table, select_halos, select_with_ms_method
    ps = hosts_with_satellites()
    n = ps.n
    log(f"drop-in + MSM n={n}")
    neikdtree.sync_from_init_adaptahop(n, 20, 20, 0.7, 0.7, 0.3, 0.3, 1.0, 1.0, 0.7, 70.0, 1.0,
                                       1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5, 70.0, 1.0, [0.0, 0.0, 0.0])
    neikdtree.sync_from_change_pos(n, n, 0.001 / n ** (1 / 3), 0.001, 1.0, 1.0)
    neikdtree.sync_others(False, False, n, 16, 80.0, 1.0 / n, 0.0, 16, ps.nvoisins, 4.0, 1.0,
                          "MSM", 30, False)
    neikdtree.compute_adaptahop(np.asfortranarray(ps.pos.T), ps.mass)
    liste = neikdtree.liste_parts.copy()
    node_0 = node_table(neikdtree.real_table.copy(order="F"),
                        neikdtree.integer_table.copy(order="F"))
    neikdtree.close()
    # method HOP: top-level ancestors; a halo's members number its top node's (inclusive) mass
    n2h, nbh = select_halos(node_0.copy())
    lab_h, _, unis_h, _, counts_h, _ = group_members(liste, node_to_struct=n2h)
    assert np.array_equal(lab_h, np.where(liste > 0, n2h[np.maximum(liste, 1) - 1], 0))
    assert np.array_equal(np.bincount(lab_h, minlength=nbh + 1)[1:], node_0["mass"][1:nbh + 1])
    r = select_with_ms_method(node_0, strict=True)
    log(f"{len(node_0) - 1} nodes -> {r.nb_halos} halos, {r.nb_sub} subhalos")
    assert r.nb_sub > 0                                   # the box has substructure
    assert np.all(node_0["level"][1:r.nb_halos + 1] == 1)
    assert np.all(node_0["level"][r.nb_halos + 1:] != 1)
    lab, pids, unis, idxs, counts, ms = group_members(liste, node_to_struct=r.node_to_struct)
    log(f"grouping {ms:.3f} ms")
    ref_lab = np.where(liste > 0, r.node_to_struct[np.maximum(liste, 1) - 1], 0)
    check_grouping(ref_lab.astype(np.int32), (lab, pids, unis, idxs, counts, ms))
    per_struct = np.bincount(lab, minlength=r.nstruct + 1)
    starters = {}
    for inode0 in range(len(node_0) - 1):
        s = int(r.node_to_struct[inode0])
        starters.setdefault(s, inode0 + 1)
    for s, inode1 in starters.items():
        assert node_0["mass"][inode1] == per_struct[s], s


def test_device_grouping_checks_its_arguments(gpu):
    import torch
    from amdahop.structures import group_members_device
    node = torch.zeros(10, dtype=torch.int32, device=gpu)
    lab = torch.empty(10, dtype=torch.int32, device=gpu)
    pids = torch.empty(10, dtype=torch.int32, device=gpu)
    with pytest.raises(ValueError, match="offsets"):
        group_members_device(node, lab, pids, torch.empty(3, dtype=torch.int64, device=gpu),
                             nnodes=5)
    with pytest.raises(ValueError, match="labels"):
        group_members_device(node, lab[:5], pids, torch.empty(7, dtype=torch.int64, device=gpu),
                             nnodes=5)
    off = torch.empty(7, dtype=torch.int64, device=gpu)
    group_members_device(node, lab, pids, off, nnodes=5)
    assert off.cpu().numpy().tolist() == [0] + [10] * 6       # everything in label 0
