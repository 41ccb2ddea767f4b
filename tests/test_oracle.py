"""The checker itself: C restatement vs the compiled reference, brute force and closed forms."""
import numpy as np
import pytest

import oracle as O
from conftest import golden_input, load_golden, saddle_as_sets
from amdahop import generate as G

CASES = ["clumpy_16k", "mixed_32k"]


@pytest.fixture(scope="module", params=CASES)
def case(request):
    g = load_golden(request.param)
    ps = golden_input(g)
    o = O.Oracle(ps.pos, ps.mass, int(g["nvoisins"]), nthreads=8)
    rho, nact = o.density()
    hop, gid, dg, ng = o.maxima(rho)
    yield g, ps, o, rho, hop, gid, dg, ng
    o.close()


def test_density_matches_reference_bitwise(case):
    g, ps, o, rho, *_ = case
    assert np.array_equal(rho, g["density"])


def test_knn_matches_reference(case):
    g, ps, o, *_ = case
    for i, d2r, idr in zip(g["knn_sample"], g["knn_d2"], g["knn_ids"]):
        d2, ids = o.knn(int(i))
        assert np.array_equal(d2, d2r)
        assert np.array_equal(ids, idr)


def test_groups_match_reference(case):
    g, ps, o, rho, hop, gid, dg, ng = case
    assert ng == int(g["ngroups"])
    assert np.array_equal(gid, g["gid"])
    assert np.array_equal(dg, g["densityg"])


def test_saddle_matches_reference(case):
    g, ps, o, rho, hop, gid, dg, ng = case
    off, isad, rs = o.saddle(rho, gid, ng)
    assert np.array_equal(off, g["saddle_offset"])
    # the restatement runs the same quicksort, so even tie order agrees
    assert np.array_equal(isad, g["saddle_isad"])
    assert np.array_equal(rs, g["saddle_rho"])


def test_hop_invariants(case):
    """Every hop goes to a lexicographically larger (rho, -index); roots = local maxima;
    group ids number the roots in ascending particle index (f90:797-803)."""
    g, ps, o, rho, hop, gid, dg, ng = case
    act = rho > 80.0
    assert np.all(hop[~act] == -1)
    i = np.nonzero(act)[0]
    t = hop[i]
    assert np.all((rho[t] > rho[i]) | ((rho[t] == rho[i]) & (t <= i)))
    roots = i[t == i]
    assert len(roots) == ng
    assert np.array_equal(gid[roots], np.arange(1, ng + 1))
    assert np.all(gid[act] > 0) and np.all(gid[~act] == 0)


def test_saddle_invariants(case):
    g, ps, o, rho, hop, gid, dg, ng = case
    off, isad, rs = o.saddle(rho, gid, ng)
    pairs = {}
    for g1 in range(ng):
        for j in range(off[g1], off[g1 + 1]):
            pairs[(g1 + 1, int(isad[j]))] = rs[j]
    for (a, b), s in pairs.items():
        assert pairs[(b, a)] == s                       # symmetric
        assert s > 80.0 and s <= min(dg[a - 1], dg[b - 1])


def test_knn_equals_brute_force():
    ps = G.clumpy_box(3000, 3, seed=7, nvoisins=24)
    o = O.Oracle(ps.pos, ps.mass, 24, nthreads=4)
    for i in range(0, 3000, 97):
        d2, ids = o.knn(i)
        bd, bi = O.knn_brute(ps.pos, i, 24)
        assert np.array_equal(d2, bd)
        assert np.array_equal(ids, bi)


def test_knn_brute_catches_wrong_periodic_wrap():
    """A point at the box edge: its nearest neighbour is across the periodic boundary."""
    pos = np.full((200, 3), 0.0)
    rng = np.random.default_rng(1)
    pos[:] = rng.uniform(-0.2, 0.2, size=(200, 3))
    pos[0] = [-0.4999, 0.0, 0.0]
    pos[1] = [0.4999, 0.0, 0.0]
    d2, ids = O.knn_brute(pos, 0, 1)
    assert ids[0] == 1 and abs(d2[0] - 0.0002 ** 2) < 1e-15
    o = O.Oracle(pos, np.full(200, 1 / 200), 1, nhop=1, nthreads=1)
    d2o, ido = o.knn(0)
    assert ido[0] == 1 and d2o[0] == d2[0]


def _lattice(n):
    a = 1.0 / n
    g = (np.arange(n) + 0.5) * a - 0.5
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1), a


def test_lattice_density_closed_form():
    """Simple cubic lattice, k=32: the first 32 neighbours are the shells at a, sqrt2 a,
    sqrt3 a and 2a (6+12+8+6), so h = a and the 2a shell has W(2) = 0:
    rho = (1 + 6 W(1) + 12 W(sqrt2) + 8 W(sqrt3)) / pi_single."""
    pos, a = _lattice(16)
    n = pos.shape[0]
    o = O.Oracle(pos, np.full(n, 1.0 / n), 32, nthreads=4)
    rho, _ = o.density()
    W = lambda q: 1 - 1.5 * q * q + 0.75 * q ** 3 if q <= 1 else 0.25 * (2 - q) ** 3
    expect = (1 + 6 * W(1.0) + 12 * W(np.sqrt(2)) + 8 * W(np.sqrt(3))) / float(np.float32(np.pi))
    assert np.allclose(rho, expect, rtol=1e-12, atol=0)
    # a dropped self term or a missing shell would miss by > 1 %
    assert abs(rho.mean() / expect - 1) < 1e-12


def test_indexx_sorts():
    rng = np.random.default_rng(3)
    for m in (1, 2, 7, 8, 50, 333):
        a = rng.integers(0, 10, size=m).astype(float)
        ix = O.indexx(a)
        assert sorted(ix.tolist()) == list(range(m))
        assert np.all(np.diff(a[ix]) >= 0)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
def test_restatement_equals_live_reference_new_seed():
    ps = G.config("c3", n=20000, nclumps=12, seed=11)
    R = O.Reference(ps.pos, ps.mass, ps.nvoisins, nthreads=8)
    st = R.staged()
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins, nthreads=8)
    rho, _ = o.density()
    hop, gid, dg, ng = o.maxima(rho)
    off, isad, rs = o.saddle(rho, gid, ng)
    assert np.array_equal(rho, st["density"])
    assert np.array_equal(gid, st["gid"])
    assert saddle_as_sets(off, isad, rs) == saddle_as_sets(st["saddle_offset"], st["saddle_isad"],
                                                           st["saddle_rho"])


def _numpy_rules(o, rho):
    from oracle import hop_saddle as HS
    _, _, nb = o.density(with_nbrs=True)
    hop = HS.hop_targets(rho, nb, o.rho_thr)
    gid, ng = HS.group_ids(hop)
    return nb, hop, gid, ng, HS.saddle_pairs(rho, gid, nb)


def test_numpy_hop_saddle_rules_equal_restatement(case):
    """oracle.hop_saddle (the full-size checker of the hop, group and saddle stages) equals the
    per-particle C restatement, which equals the compiled reference."""
    from oracle import hop_saddle as HS
    g, ps, o, rho, hop, gid, dg, ng = case
    nb, hop2, gid2, ng2, pairs = _numpy_rules(o, rho)
    assert np.array_equal(hop2, hop)
    assert ng2 == ng and np.array_equal(gid2, gid)
    ref = HS.csr_as_pairs(g["saddle_offset"], g["saddle_isad"], g["saddle_rho"])
    for a, b in zip(pairs, ref):
        assert np.array_equal(a, b)


def test_numpy_hop_saddle_rules_on_ties():
    """Equal densities (duplicated positions and masses) exercise the min-index tie rule."""
    from oracle import hop_saddle as HS
    ps = G.config("c1", n=6000, nclumps=4, seed=3)
    ps.pos[100:140] = ps.pos[200:240]
    o = O.Oracle(ps.pos, ps.mass, ps.nvoisins, nthreads=4)
    rho, _ = o.density()
    hop, gid, dg, ng = o.maxima(rho)
    assert len(np.unique(rho[rho > 80])) < int((rho > 80).sum())   # ties are present
    nb, hop2, gid2, ng2, pairs = _numpy_rules(o, rho)
    assert np.array_equal(hop2, hop) and np.array_equal(gid2, gid)
    ref = HS.csr_as_pairs(*o.saddle(rho, gid, ng))
    for a, b in zip(pairs, ref):
        assert np.array_equal(a, b)
    # planted mistakes in the rules are caught: ties to the larger index, max instead of min
    # in the symmetrisation
    act = np.nonzero(rho > 80)[0]
    cand = np.concatenate([act[:, None], nb[act]], axis=1)
    rc = rho[cand]
    wrong = np.where(rc == rc.max(axis=1)[:, None], cand, -1).max(axis=1)
    assert not np.array_equal(wrong, hop[act])
    # directed maxima before the symmetrisation: some pairs differ between the two
    # directions, so symmetrising with max instead of min would change them
    i = np.nonzero(gid > 0)[0]
    j = nb[i]
    ga, gb = np.broadcast_to(gid[i][:, None], j.shape), gid[j]
    keep = (gb > 0) & (gb != ga)
    sv = np.minimum(np.broadcast_to(rho[i][:, None], j.shape), rho[j])[keep]
    directed = {}
    for a, b, v in zip(ga[keep].tolist(), gb[keep].tolist(), sv.tolist()):
        directed[(a, b)] = max(directed.get((a, b), -1.0), v)
    g1, g2, s = pairs
    with_max = np.array([max(directed[(a, b)], directed[(b, a)])
                         for a, b in zip(g1.tolist(), g2.tolist())])
    assert not np.array_equal(with_max, ref[2])


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
def test_reference_optimize_nodes_tables():
    """The reference's moment-prefix create-nodes path (optimize_nodes, f90:954-958,
    :1632-2311), run live on C1: the stored optimize_nodes tables are its output; its node ids
    and integer tables are the default path's, its real tables are not.  On C1 they differ by
    the rounding of the prefix sums (relative 1e-11 at most: the radius, a square root of a
    difference of moments, moves most); on C2 positions and radii of components that span the
    periodic edge differ by far more (each group's moments are unwrapped about its own
    densest member)."""
    g = load_golden("c1_tables")
    opt = load_golden("c1_tables_opt")
    ps = golden_input(g)
    R = O.Reference(ps.pos, ps.mass, int(g["nvoisins"]), nthreads=8)
    r = R.full(optimize_nodes=True)
    assert r["nnodes"] == int(g["nnodes"]) == int(opt["nnodes"])
    assert np.array_equal(r["liste_parts"], g["liste_parts"])
    assert np.array_equal(r["integer_table"], g["integer_table"])
    assert np.array_equal(r["real_table"], opt["real_table"])
    rt, gt = r["real_table"], g["real_table"]
    assert not np.array_equal(rt, gt)            # a different summation, not the same path
    assert np.array_equal(rt[[0, 2, 4]], gt[[0, 2, 4]])     # saddle, densmax, truemass
    assert np.allclose(rt, gt, rtol=1e-11, atol=0)
