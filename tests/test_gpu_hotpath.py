"""HIP path vs the oracle / the compiled reference's golden fixtures, through the C ABI.

Bit-exact bars: densities (FP64 values computed in the reference's arithmetic order),
hop targets, group ids, node ids and node tables.  The saddle list is compared as a
multiset per (group, saddle density) because the reference sorts it with an unstable
quicksort.  Both precision modes must meet the same bars: the mixed path certifies
its FP32 neighbour ranking and re-ranks in FP64, so its outputs are exact too.
"""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden_input, load_golden, saddle_as_sets
from amdahop import generate as G

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dropin_run(ps, mode="mixed", bucket=64, nhop=16, rho_threshold=80.0, xlong=1.0,
               nvoisins=None, optimize_nodes=False):
    """Drive the drop-in module exactly like compute_neiKDtree_mod.py does."""
    from compute_adaptahop import neikdtree
    n = ps.n
    k = nvoisins or ps.nvoisins
    neikdtree.set_precision(mode)
    neikdtree.set_bucket(bucket)
    neikdtree.sync_from_init_adaptahop(n, 20, 20, 0.7, 0.7, 0.3, 0.3, 1.0, 1.0, 0.7, 70.0, xlong,
                                       xlong, xlong, xlong, xlong, xlong / 2, xlong / 2, xlong / 2,
                                       70.0, 1.0, [0.0, 0.0, 0.0])
    neikdtree.sync_from_change_pos(n, n, 0.001 * xlong / n ** (1 / 3), 0.001, xlong, xlong)
    neikdtree.sync_others(False, False, n, 16, rho_threshold, 1.0 / n, 0.0, nhop, k, 4.0, 1.0,
                          "MSM", 30, optimize_nodes)
    neikdtree.compute_adaptahop(np.asfortranarray(ps.pos.T), ps.mass)
    out = dict(density=neikdtree.density.copy(), liste_parts=neikdtree.liste_parts.copy(),
               nnodes=int(neikdtree.nnodes), real_table=neikdtree.real_table.copy(order="F"),
               integer_table=neikdtree.integer_table.copy(order="F"))
    out["gid"], out["hop"] = neikdtree.groups()
    out["saddle"] = neikdtree.saddle()
    neikdtree.close()
    neikdtree.set_precision("mixed")
    neikdtree.set_bucket(64)
    if optimize_nodes:
        neikdtree.sync_others(False, False, n, 16, rho_threshold, 1.0 / n, 0.0, nhop, k, 4.0, 1.0,
                              "MSM", 30, False)
    return out


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["clumpy_16k", "mixed_32k"])
def test_dropin_matches_reference_golden(gpu, name, mode):
    g = load_golden(name)
    ps = golden_input(g)
    log(f"golden {name} mode={mode} n={ps.n}")
    r = dropin_run(ps, mode)
    assert np.array_equal(r["density"], g["density"])
    assert np.array_equal(r["gid"], g["gid"])
    assert r["nnodes"] == int(g["nnodes"])
    assert np.array_equal(r["liste_parts"], g["liste_parts"])
    assert np.array_equal(r["real_table"], g["real_table"])
    assert np.array_equal(r["integer_table"], g["integer_table"])
    assert saddle_as_sets(*r["saddle"]) == saddle_as_sets(g["saddle_offset"], g["saddle_isad"],
                                                          g["saddle_rho"])


@pytest.fixture(scope="module")
def c1_oracle():
    ps = G.config("c1")
    log(f"oracle on c1 n={ps.n}")
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins)
    rho, nact = o.density()
    hop, gid, dg, ng = o.maxima(rho)
    sad = o.saddle(rho, gid, ng)
    o.close()
    return ps, rho, hop, gid, ng, sad


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_c1_full_size_matches_oracle(gpu, c1_oracle, mode):
    ps, rho, hop, gid, ng, sad = c1_oracle
    log(f"c1 mode={mode}")
    r = dropin_run(ps, mode)
    assert np.array_equal(r["density"], rho)
    assert np.array_equal(r["hop"], hop)
    assert np.array_equal(r["gid"], gid)
    assert saddle_as_sets(*r["saddle"]) == saddle_as_sets(*sad)


@pytest.mark.parametrize("bucket", [1, 7, 32])
def test_tree_shape_does_not_change_results(gpu, c1_oracle, bucket):
    ps, rho, hop, gid, ng, sad = c1_oracle
    log(f"c1 bucket={bucket}")
    r = dropin_run(ps, "mixed", bucket=bucket)
    assert np.array_equal(r["density"], rho)
    assert np.array_equal(r["gid"], gid)


@pytest.mark.parametrize("k", [1, 7, 16, 20, 40, 63, 100, 128])
def test_every_heap_shape(gpu, k):
    """k from 1 to 128 covers every heap capacity of the kNN kernel (k <= 16: the whole mixed
    heap in registers; 32, 64, 128: register entries with groups of LDS children) and k values
    that leave groups partly used.  Densities and neighbour rows bit-exact, both precisions."""
    import torch
    from amdahop import knn_device
    ps = G.config("c1", n=6000, nclumps=5, seed=11)
    log(f"heap shapes k={k}")
    o = O.Oracle(ps.pos, ps.mass, k, nhop=min(16, k))
    rho, _ = o.density()
    rows = {int(i): o.knn(int(i))[1]
            for i in np.random.default_rng(k).choice(ps.n, 100, replace=False)}
    o.close()
    pos = torch.from_numpy(ps.pos).cuda()
    mass = torch.from_numpy(ps.mass).cuda()
    for mixed in (True, False):
        dens, nbr, _ = knn_device(pos, mass, k, mixed=mixed)
        assert np.array_equal(dens.cpu().numpy(), rho), f"mixed={mixed}"
        nb = nbr.cpu().numpy()
        for i, ids in rows.items():
            assert np.array_equal(nb[i], ids), f"mixed={mixed} particle {i}"


def _device_run(ps, mixed, saddle=True, **kw):
    import torch
    from amdahop import run_device
    pos = torch.from_numpy(ps.pos).cuda()
    mass = torch.from_numpy(ps.mass).cuda()
    dens, gid, info = run_device(pos, mass, ps.nvoisins, mixed=mixed, saddle=saddle, **kw)
    return dens.cpu().numpy(), gid.cpu().numpy(), info


def test_c2_mixed_equals_fp64_and_oracle_sample(gpu):
    ps = G.config("c2")
    log(f"c2 n={ps.n}")
    d64, g64, i64 = _device_run(ps, False)
    dmx, gmx, imx = _device_run(ps, True)
    log(f"c2 fallback queries {imx.nfallback}, groups {imx.ngroups}")
    assert np.array_equal(d64, dmx)
    assert np.array_equal(g64, gmx)
    assert i64.ngroups == imx.ngroups and i64.nactive == imx.nactive
    assert imx.nfallback < ps.n // 100
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins)
    idx = np.random.default_rng(9).choice(ps.n, size=20000, replace=False)
    assert np.array_equal(o.density_at(idx), dmx[idx])


@pytest.fixture(scope="module")
def c3_dropin():
    """The headline box (16,777,216 particles, k=64, mixed families) through the drop-in in
    both precision modes."""
    ps = G.config("c3")
    runs = {}
    for mode in ("mixed", "fp64"):
        log(f"c3 drop-in mode={mode} n={ps.n}")
        runs[mode] = dropin_run(ps, mode)
    return ps, runs


def _assert_saddle_decreasing(off, rho):
    off = np.asarray(off, dtype=np.int64)
    d = np.diff(rho)
    inner = np.ones(d.shape[0], dtype=bool)
    inner[off[1:-1][(off[1:-1] > 0) & (off[1:-1] < rho.shape[0])] - 1] = False
    assert np.all(d[inner] <= 0), "saddle list not decreasing within a group"


@pytest.mark.timeout(900)
def test_c3_mixed_equals_fp64_everywhere(gpu, c3_dropin):
    """Every output of the mixed path equals the FP64 path's over all 16.8M particles:
    densities (so the relative density error is 0), group ids, hop targets, node ids, node
    tables and the saddle multisets."""
    from oracle import hop_saddle as HS
    ps, runs = c3_dropin
    m, f = runs["mixed"], runs["fp64"]
    assert np.array_equal(m["density"], f["density"])
    assert np.array_equal(m["gid"], f["gid"])
    assert np.array_equal(m["hop"], f["hop"])
    assert m["nnodes"] == f["nnodes"] and m["nnodes"] > 1000
    assert np.array_equal(m["liste_parts"], f["liste_parts"])
    assert np.array_equal(m["real_table"], f["real_table"])
    assert np.array_equal(m["integer_table"], f["integer_table"])
    for r in (m, f):
        _assert_saddle_decreasing(r["saddle"][0], r["saddle"][2])
    pm, pf = HS.csr_as_pairs(*m["saddle"]), HS.csr_as_pairs(*f["saddle"])
    for a, b in zip(pm, pf):
        assert np.array_equal(a, b)
    log(f"c3: {int(m['gid'].max())} groups, {pm[0].shape[0]} saddle entries, "
        f"{m['nnodes']} nodes, all equal between modes")


@pytest.mark.timeout(900)
def test_c3_full_size_matches_oracle(gpu, c3_dropin):
    """Over all 16.8M particles: densities bit-exact against the oracle (the C restatement
    pinned to the compiled reference), and hop targets, group ids and saddle multisets equal
    to the reference's rules (oracle.hop_saddle, pinned to the restatement) applied to the
    oracle's own neighbour rows and densities."""
    from oracle import hop_saddle as HS
    ps, runs = c3_dropin
    r = runs["mixed"]
    log("c3 oracle: density and first-16 neighbour rows of every particle")
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins)
    rho, nact, nb = o.density(with_nbrs=True)
    o.close()
    assert np.array_equal(r["density"], rho)
    log(f"c3 oracle: densities equal ({nact} active); hop, groups and saddles")
    hop = HS.hop_targets(rho, nb, 80.0)
    assert np.array_equal(r["hop"], hop)
    gid, ng = HS.group_ids(hop)
    assert np.array_equal(r["gid"], gid)
    pairs = HS.saddle_pairs(rho, gid, nb)
    got = HS.csr_as_pairs(*r["saddle"])
    for a, b in zip(got, pairs):
        assert np.array_equal(a, b)
    log(f"c3 oracle: {ng} groups, {pairs[0].shape[0]} saddle entries equal")


def test_saddle_symmetric_and_bounded_c2(gpu):
    import torch
    from amdahop import run_device, device_saddle
    ps = G.config("c2")
    pos = torch.from_numpy(ps.pos).cuda()
    mass = torch.from_numpy(ps.mass).cuda()
    dens, gid, info = run_device(pos, mass, ps.nvoisins, saddle=True)
    off, isad, rs = device_saddle(info)
    rho = dens.cpu().numpy()
    g = gid.cpu().numpy()
    dg = np.zeros(info.ngroups + 1)
    np.maximum.at(dg, g, rho)
    g1 = np.repeat(np.arange(1, info.ngroups + 1), np.diff(off))
    fwd = dict(zip(zip(g1.tolist(), isad.tolist()), rs.tolist()))
    assert len(fwd) == len(rs)
    for (a, b), s in fwd.items():
        assert fwd[(b, a)] == s
        assert 80.0 < s <= min(dg[a], dg[b])


def test_small_and_edge_inputs(gpu):
    # tiny uniform box, k close to N
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.5, 0.5, size=(150, 3))
    ps = G.ParticleSet(pos, np.full(150, 1 / 150), 100, 16, "tiny", 150)
    o = O.Oracle(ps.pos, ps.mass, 100)
    rho, _ = o.density()
    for mode in ("fp64", "mixed"):
        r = dropin_run(ps, mode, rho_threshold=0.5)
        assert np.array_equal(r["density"], rho)
    # particles exactly on the lower box face and duplicated positions
    ps = G.config("c1", n=6000, nclumps=4, seed=3)
    ps.pos[:20, 0] = -0.5
    ps.pos[100:110] = ps.pos[200:210]
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins)
    rho, _ = o.density()
    hop, gid, _, ng = o.maxima(rho)
    r = dropin_run(ps, "mixed")
    assert np.array_equal(r["density"], rho)
    assert np.array_equal(r["gid"], gid)
    # nothing above threshold: no groups, no nodes
    r = dropin_run(ps, "mixed", rho_threshold=1e12)
    assert r["nnodes"] == 0 and not r["gid"].any() and not r["liste_parts"].any()
    # nhop == nvoisins
    r = dropin_run(ps, "fp64", nhop=32)
    hop2, gid2, _, _ = O.Oracle(ps.pos, ps.mass, 32, nhop=32).maxima(rho)
    assert np.array_equal(r["gid"], gid2)


def test_non_unit_box(gpu):
    ps = G.config("c1", n=8000, nclumps=5, seed=8)
    ps.pos *= 2.0
    ps.mass /= 8.0
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins, xlong=2.0)
    rho, _ = o.density()
    r = dropin_run(ps, "mixed", xlong=2.0)
    assert np.array_equal(r["density"], rho)


def test_out_of_box_position_is_an_error(gpu):
    ps = G.config("c1", n=4000, nclumps=3, seed=2)
    ps.pos[7, 1] = 0.5
    with pytest.raises(RuntimeError, match="inside"):
        dropin_run(ps, "mixed")
    ps.pos[7, 1] = np.nan
    with pytest.raises(RuntimeError, match="inside"):
        dropin_run(ps, "mixed")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["c1_tables", "c2_tables"])
def test_dropin_node_tables_at_size(gpu, name, mode):
    """The whole drop-in call at 262,144 and 2,097,152 particles (150 and 1,084 nodes):
    node ids and node tables bit for bit with the reference's own full run, the host tree's
    parallel branches taken."""
    from amdahop._lib import check, lib, ptr
    g = load_golden(name)
    ps = golden_input(g)
    log(f"drop-in tables {name} mode={mode} n={ps.n}")
    r = dropin_run(ps, mode)
    assert r["nnodes"] == int(g["nnodes"])
    assert np.array_equal(r["liste_parts"], g["liste_parts"])
    assert np.array_equal(r["real_table"], g["real_table"])
    assert np.array_equal(r["integer_table"], g["integer_table"])
    st = np.zeros(4, dtype=np.int64)
    check(lib().ahop_tree_stats(ptr(st), None))
    log(f"{name}: taskloop levels {st[0]}, subtree tasks {st[2]}")
    assert st[0] >= 1 and st[2] >= 1


@pytest.mark.timeout(600)
def test_dropin_optimize_nodes_tables(gpu):
    """The drop-in call with optimize_nodes on at 2,097,152 particles: node ids and both
    tables bit for bit with the reference's own optimize_nodes run (its moment-prefix node
    creation, f90:1632-2311)."""
    g = load_golden("c2_tables")
    opt = load_golden("c2_tables_opt")
    ps = golden_input(g)
    log(f"drop-in optimize_nodes c2 n={ps.n}")
    r = dropin_run(ps, "mixed", optimize_nodes=True)
    assert r["nnodes"] == int(opt["nnodes"])
    assert np.array_equal(r["liste_parts"], g["liste_parts"])
    assert np.array_equal(r["integer_table"], opt["integer_table"])
    assert np.array_equal(r["real_table"], opt["real_table"])
