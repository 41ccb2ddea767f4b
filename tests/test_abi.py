"""C ABI: the library loads without a GPU, exports every declared symbol, and its host
logic (argument checks, the structure-tree stage) is correct.  No device call here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_golden, golden_input
from amdahop._lib import LIB_PATH, SIGNATURES, check, lib, ptr

HEADER = os.path.join(ROOT, "include", "amdahop.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ahop_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_binding_covers_header():
    assert sorted(SIGNATURES) == declared_functions()


def test_dropin_module_surface():
    from compute_adaptahop import neikdtree
    for f in ("compute_adaptahop", "sync_from_change_pos", "sync_from_init_adaptahop",
              "sync_others", "close"):
        assert callable(getattr(neikdtree, f))
    assert neikdtree.liste_parts is None and neikdtree.density is None
    assert int(neikdtree.nnodes) == 0 and neikdtree.nnodes.dtype == np.int32


def test_bad_shape_is_an_error_not_a_stop():
    """The reference STOPs on pos_in not shaped (3,npart) (f90:509-512)."""
    from compute_adaptahop import neikdtree
    with pytest.raises(RuntimeError, match=r"shape \(3,npart\)"):
        neikdtree.compute_adaptahop(np.zeros((4, 10)), np.zeros(10))
    pos = np.zeros((3, 10), order="F")
    status = lib().ahop_compute_adaptahop(ptr(pos), 2, 10, ptr(np.zeros(10)), 10)
    assert status != 0 and b"shape" in lib().ahop_last_error()


def test_size_mismatch_is_an_error():
    from compute_adaptahop import neikdtree
    neikdtree.sync_from_change_pos(100, 100, 0.1, 0.001, 1.0, 1.0)
    with pytest.raises(RuntimeError, match="do not match npart"):
        neikdtree.compute_adaptahop(np.zeros((3, 50), order="F"), np.zeros(50))


def test_parameter_validation():
    assert lib().ahop_set_precision(2) != 0
    assert lib().ahop_set_bucket(0) != 0 and lib().ahop_set_bucket(65) != 0
    check(lib().ahop_set_bucket(32))
    check(lib().ahop_set_precision(1))


def _run_tree(ps, g, eps):
    n = ps.n
    ng = int(g["ngroups"])
    lp = np.empty(n, dtype=np.int32)
    nn = ctypes.c_int32(0)
    rt = np.zeros(8 * 2 * max(ng, 1))
    it = np.zeros(6 * 2 * max(ng, 1), dtype=np.int32)
    check(lib().ahop_create_group_tree(
        n, ptr(ps.pos), ptr(ps.mass), ptr(g["density"]), ptr(g["gid"]), ng, ptr(g["densityg"]),
        ptr(np.ascontiguousarray(g["saddle_offset"], dtype=np.int32)),
        ptr(np.ascontiguousarray(g["saddle_isad"], dtype=np.int32)),
        ptr(np.ascontiguousarray(g["saddle_rho"])), 80.0, 4.0, 1.0, eps, 20, 1.0, ptr(lp),
        ctypes.byref(nn), ptr(rt), ptr(it)))
    k = nn.value
    return lp, k, rt[:8 * k].reshape(8, k, order="F"), it[:6 * k].reshape(6, k, order="F")


@pytest.mark.parametrize("name", ["clumpy_16k", "mixed_32k"])
def test_host_structure_tree_matches_reference(name):
    """create_group_tree (f90:885-983) on the reference's own stage outputs reproduces
    the reference's node ids and node tables bit for bit."""
    g = load_golden(name)
    ps = golden_input(g)
    eps = 0.001 * 1.0 / ps.n ** (1.0 / 3.0)
    lp, k, rt, it = _run_tree(ps, g, eps)
    assert k == int(g["nnodes"])
    assert np.array_equal(lp, g["liste_parts"])
    assert np.array_equal(rt, g["real_table"])
    assert np.array_equal(it, g["integer_table"])


def test_host_structure_tree_sensitive_to_threshold():
    """Planted change: a higher membership threshold must change the tree."""
    g = load_golden("clumpy_16k")
    ps = golden_input(g)
    eps = 0.001 / ps.n ** (1.0 / 3.0)
    n = ps.n
    ng = int(g["ngroups"])
    lp = np.empty(n, dtype=np.int32)
    nn = ctypes.c_int32(0)
    rt = np.zeros(16 * ng)
    it = np.zeros(12 * ng, dtype=np.int32)
    check(lib().ahop_create_group_tree(
        n, ptr(ps.pos), ptr(ps.mass), ptr(g["density"]), ptr(g["gid"]), ng, ptr(g["densityg"]),
        ptr(np.ascontiguousarray(g["saddle_offset"], dtype=np.int32)),
        ptr(np.ascontiguousarray(g["saddle_isad"], dtype=np.int32)),
        ptr(np.ascontiguousarray(g["saddle_rho"])), 80.0, 4.0, 1.0, eps, 2000, 1.0, ptr(lp),
        ctypes.byref(nn), ptr(rt), ptr(it)))
    assert nn.value < int(g["nnodes"])


def test_sync_calls_accept_the_bridge_argument_lists():
    """The three sync_* calls with the argument lists compute_neiKDtree_mod.py passes
    (init_adaptahop_130 :183-197, change_pos_1310 :207-210, sync_fortran :740-747)."""
    from compute_adaptahop import neikdtree
    neikdtree.sync_from_init_adaptahop(1000, 20, 20, 0.7, 0.7, 0.3, 0.3, 1.0, 1.0, 0.7, 70.0, 1.0,
                                       1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5, 70.0, 1.0,
                                       np.zeros(3))
    neikdtree.sync_from_change_pos(np.int64(1000), 1000, 1e-4, 0.001, 1.0, 1.0)
    neikdtree.sync_others(False, False, 1000, 16, 80.0, 1e-3, 0.0, 16, 32, 4.0, 1.0, "MSM", 30, False)
    neikdtree.close()


def _sha(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def table_case_stages(g):
    """The stage outputs the host structure tree starts from, regenerated with the oracle for
    a node-table golden and checked against the reference's digests: densities, group ids,
    group peak densities and the saddle CSR (per group in decreasing saddle density)."""
    import oracle as O
    from oracle import hop_saddle as HS
    ps = golden_input(g)
    o = O.Oracle(ps.pos, ps.mass, int(g["nvoisins"]), nthreads=8)
    rho, _, nb = o.density(with_nbrs=True)
    o.close()
    hop = HS.hop_targets(rho, nb, 80.0)
    gid, ng = HS.group_ids(hop)
    dg = np.zeros(ng)
    roots = np.nonzero(hop == np.arange(ps.n))[0]
    dg[gid[roots] - 1] = rho[roots]
    assert _sha(rho, gid, dg) == str(g["stage_digest"])
    g1, g2, s = HS.saddle_pairs(rho, gid, nb)
    assert _sha(g1, g2, s) == str(g["saddle_digest"])
    order = np.lexsort((g2, -s, g1))
    off = np.zeros(ng + 1, dtype=np.int32)
    np.add.at(off, g1.astype(np.int64), 1)
    off = np.cumsum(off).astype(np.int32)
    return ps, rho, gid, ng, dg, off, g2[order].astype(np.int32), s[order]


@pytest.mark.parametrize("name", ["c1_tables", "c2_tables"])
def test_host_structure_tree_at_size(name):
    """create_group_tree (f90:885-1131, :1245-1628) at 262,144 and 2,097,152 particles, on
    stage outputs equal to the reference's: node ids and node tables bit for bit with the
    reference's full run, and the parallel branches (the component taskloop and the subtree
    tasks) taken."""
    g = load_golden(name)
    ps, rho, gid, ng, dg, off, isad, srho = table_case_stages(g)
    eps = 0.001 / ps.n ** (1.0 / 3.0)
    lp = np.empty(ps.n, dtype=np.int32)
    nn = ctypes.c_int32(0)
    rt = np.zeros(8 * 2 * ng)
    it = np.zeros(6 * 2 * ng, dtype=np.int32)
    check(lib().ahop_create_group_tree(ps.n, ptr(ps.pos), ptr(ps.mass), ptr(rho), ptr(gid), ng,
                                       ptr(dg), ptr(off), ptr(isad), ptr(srho), 80.0, 4.0, 1.0,
                                       eps, 20, 1.0, ptr(lp), ctypes.byref(nn), ptr(rt), ptr(it)))
    k = nn.value
    assert k == int(g["nnodes"])
    assert np.array_equal(lp, g["liste_parts"])
    assert np.array_equal(rt[:8 * k].reshape(8, k, order="F"), g["real_table"])
    assert np.array_equal(it[:6 * k].reshape(6, k, order="F"), g["integer_table"])
    st = np.zeros(4, dtype=np.int64)
    check(lib().ahop_tree_stats(ptr(st), None))
    print(f"{name}: {k} nodes, taskloop levels {st[0]} ({st[1]} components), "
          f"subtree tasks {st[2]}")
    assert st[3] == k and st[0] >= 1 and st[1] >= 16 and st[2] >= 1


def _set_optimize_nodes(on, n):
    from compute_adaptahop import neikdtree
    neikdtree.sync_others(False, False, n, 16, 80.0, 1.0 / n, 0.0, 16, 32, 4.0, 1.0, "MSM", 30, on)


@pytest.mark.parametrize("name", ["c1_tables", "c2_tables"])
def test_host_structure_tree_optimize_nodes(name):
    """The moment-prefix node creation (optimize_nodes, f90:954-958, :1632-2311) on stage
    outputs equal to the reference's: node ids and both tables bit for bit with the
    reference's own optimize_nodes run, whose real tables differ from the default path's
    (position and radius of components spanning the periodic edge, rounding elsewhere)."""
    g = load_golden(name)
    opt = load_golden(name + "_opt")
    ps, rho, gid, ng, dg, off, isad, srho = table_case_stages(g)
    eps = 0.001 / ps.n ** (1.0 / 3.0)

    def run():
        lp = np.empty(ps.n, dtype=np.int32)
        nn = ctypes.c_int32(0)
        rt = np.zeros(8 * 2 * ng)
        it = np.zeros(6 * 2 * ng, dtype=np.int32)
        check(lib().ahop_create_group_tree(ps.n, ptr(ps.pos), ptr(ps.mass), ptr(rho), ptr(gid),
                                           ng, ptr(dg), ptr(off), ptr(isad), ptr(srho), 80.0, 4.0,
                                           1.0, eps, 20, 1.0, ptr(lp), ctypes.byref(nn), ptr(rt),
                                           ptr(it)))
        k = nn.value
        return lp, k, rt[:8 * k].reshape(8, k, order="F"), it[:6 * k].reshape(6, k, order="F")

    _set_optimize_nodes(True, ps.n)
    try:
        lp, k, rt, it = run()
    finally:
        _set_optimize_nodes(False, ps.n)
    assert k == int(opt["nnodes"])
    assert np.array_equal(lp, g["liste_parts"])
    assert np.array_equal(it, opt["integer_table"])
    assert np.array_equal(rt, opt["real_table"])
    assert not np.array_equal(rt, g["real_table"])
    # the switch is back off: the default path's tables again
    lp, k, rt, it = run()
    assert np.array_equal(rt, g["real_table"]) and np.array_equal(it, g["integer_table"])
