"""MSM selection (amdahop.structures.select_with_ms_method) on node tables: a hand-worked tree,
the tie rule, random forests against an independent walk-up definition and the reference's
own particle-count check (compute_neiKDtree_mod.py:381-392), and the error case."""
import numpy as np
import pytest

from amdahop.structures import NODE_DTYPE, node_table, select_with_ms_method


def make_nodes(mother, own_count, truemass_own=None, densmax=None):
    """node_0 for a forest given by 1-based mothers (0 = halo; children after their mother),
    with each node's own particle count: inclusive masses, first child / sister links in
    index order, levels."""
    nn = len(mother)
    node_0 = np.zeros(nn + 1, dtype=NODE_DTYPE)
    mother = np.asarray(mother)
    own = np.asarray(own_count)
    tm_own = np.asarray(truemass_own if truemass_own is not None else own, dtype=np.float64)
    mass = own.astype(np.int64).copy()
    tm = tm_own.copy()
    for i in range(nn - 1, -1, -1):          # children come after mothers
        if mother[i] > 0:
            mass[mother[i] - 1] += mass[i]
            tm[mother[i] - 1] += tm[i]
    node_0["mass"][1:] = mass
    node_0["truemass"][1:] = tm
    node_0["mother"][1:] = mother
    node_0["densmax"][1:] = densmax if densmax is not None else np.arange(1, nn + 1)
    level = np.zeros(nn, dtype=np.int32)
    last_child = {}
    for i in range(nn):
        m = mother[i]
        level[i] = 1 if m <= 0 else level[m - 1] + 1
        if m > 0:
            if node_0["firstchild"][m] == 0:
                node_0["firstchild"][m] = i + 1
            else:
                node_0["sister"][last_child[m]] = i + 1
            last_child[m] = i + 1
    node_0["level"][1:] = level
    return node_0


def test_hand_worked_tree():
    # halo 1 with subs 3 (inclusive 40, holding 5) and 4; halo 2 alone
    mother = [0, 0, 1, 1, 3]
    own = [35, 30, 30, 25, 10]
    node_0 = make_nodes(mother, own, densmax=[5.0, 3.0, 9.0, 7.0, 11.0])
    assert list(node_0["mass"][1:]) == [100, 30, 40, 25, 10]
    r = select_with_ms_method(node_0)
    # node 3 is node 1's most massive subnode (40 > 25) and node 5 is node 3's: both join
    # halo 1; node 4 starts structure 3
    assert list(r.node_to_struct) == [1, 2, 1, 3, 1]
    assert (r.nb_halos, r.nb_sub) == (2, 1)
    assert list(r.mother) == [0, 0, 1]
    assert list(r.first_sister) == [2, -1, -1]
    assert list(r.first_daughter) == [3, -1, -1]
    assert list(r.level) == [1, 1, 2]
    # node 1: own 35 + node 3 (own 30 + node 5's 10) = 75, densmax from node 5 through node 3
    assert node_0["mass"][1] == 75 and node_0["truemass"][1] == 75.0
    assert node_0["densmax"][1] == 11.0 and node_0["densmax"][3] == 11.0
    assert node_0["mass"][4] == 25 and node_0["mass"][2] == 30


def test_equal_truemass_keeps_the_first_subnode():
    node_0 = make_nodes([0, 1, 1, 1], [10, 5, 7, 7])
    r = select_with_ms_method(node_0)
    # subnodes 3 and 4 tie at 7: the first in sister order (3) joins the halo
    assert list(r.node_to_struct) == [1, 2, 1, 3]
    assert r.nb_sub == 2 and list(r.mother) == [0, 1, 1]
    assert list(r.first_daughter) == [2, -1, -1] and list(r.first_sister) == [-1, 3, -1]


def test_node_without_own_particles_is_an_error():
    node_0 = make_nodes([0, 1], [3, 4])
    node_0["mass"][1] = 4        # the halo's inclusive mass equals its subnode's
    with pytest.raises(ValueError, match="after removing"):
        select_with_ms_method(node_0)


def random_forest(rng, nn):
    nh = max(1, nn // 5)
    mother = [0] * nh
    for i in range(nh, nn):
        mother.append(int(rng.integers(1, i + 1)))
    own = rng.integers(1, 50, size=nn)
    tm = own * rng.uniform(0.5, 2.0, size=nn)
    return mother, own, tm


@pytest.mark.parametrize("seed", range(6))
def test_random_forests(seed):
    rng = np.random.default_rng(seed)
    nn = int(rng.integers(5, 300))
    mother, own, tm = random_forest(rng, nn)
    node_0 = make_nodes(mother, own, truemass_own=tm, densmax=rng.uniform(1, 10, size=nn))
    orig = node_0.copy()
    r = select_with_ms_method(node_0)
    nh = sum(1 for m in mother if m == 0)
    assert r.nb_halos == nh and r.nstruct == r.nb_halos + r.nb_sub
    assert list(r.node_to_struct[:nh]) == list(range(1, nh + 1))
    # independent definition, recursive: a node's main-branch truemass is its own truemass
    # plus its heaviest subnode's main-branch truemass (first in sister order on ties), and a
    # node belongs to its mother's structure iff it is that heaviest subnode; walk up
    import functools

    @functools.lru_cache(maxsize=None)
    def branch(m):
        c, kids = orig["firstchild"][m], []
        while c > 0:
            kids.append(int(c))
            c = orig["sister"][c]
        best = None
        for c in kids:
            if best is None or branch(c)[0] > branch(best)[0]:
                best = c
        own_tm = float(tm[m - 1])
        return (own_tm + (branch(best)[0] if best else 0.0), best)

    def most_massive_child(m):
        return branch(m)[1]

    top = np.empty(nn, dtype=np.int64)
    for i in range(1, nn + 1):
        j = i
        while mother[j - 1] > 0 and most_massive_child(mother[j - 1]) == j:
            j = mother[j - 1]
        top[i - 1] = j
    for i in range(nn):
        for j in range(nn):
            assert (top[i] == top[j]) == (r.node_to_struct[i] == r.node_to_struct[j])
    # the reference's count check: particles carried by each structure = its top node's mass
    counts = np.zeros(r.nstruct + 1, dtype=np.int64)
    np.add.at(counts, r.node_to_struct, own)
    for s in range(1, r.nstruct + 1):
        t = top[np.nonzero(r.node_to_struct == s)[0][0]]
        assert node_0["mass"][t] == counts[s]
    # structure tree: mothers, levels and daughter lists agree
    for s in range(1, r.nstruct + 1):
        m = r.mother[s - 1]
        assert (m == 0) == (r.level[s - 1] == 1)
        if m > 0:
            assert r.level[s - 1] == r.level[m - 1] + 1
    for m in range(1, r.nstruct + 1):
        daughters, d = [], r.first_daughter[m - 1]
        while d > 0:
            daughters.append(d)
            d = r.first_sister[d - 1]
        assert sorted(daughters) == [s for s in range(1, r.nstruct + 1) if r.mother[s - 1] == m]


def test_node_table_layout():
    rt = np.arange(8 * 3, dtype=np.float64).reshape(8, 3)
    it = np.arange(6 * 3, dtype=np.int32).reshape(6, 3)
    n0 = node_table(rt, it)
    assert len(n0) == 4 and n0["truemass"][2] == rt[4, 1] and n0["sister"][3] == it[5, 2]
    with pytest.raises(ValueError):
        node_table(rt[:7], it)


def test_hop_selection_takes_the_top_level_ancestor():
    from amdahop.structures import select_halos
    # halos 1, 2; node 3 under 1, node 4 under 3, node 5 under 2
    node_0 = make_nodes([0, 0, 1, 3, 2], [5, 5, 5, 5, 5])
    n2h, nb = select_halos(node_0)
    assert nb == 2 and list(n2h) == [1, 2, 1, 1, 2]
    # a top-level node numbered after a subnode: its id exceeds the halo count
    node_0 = make_nodes([0, 1, 0], [5, 5, 5])
    with pytest.raises(ValueError, match="nb_of_halos"):
        select_halos(node_0)


@pytest.mark.parametrize("seed", range(3))
def test_hop_selection_on_random_forests(seed):
    from amdahop.structures import select_halos
    rng = np.random.default_rng(seed)
    nn = int(rng.integers(5, 200))
    mother, own, tm = random_forest(rng, nn)
    n2h, nb = select_halos(make_nodes(mother, own))
    for i in range(nn):                       # walk up independently
        j = i + 1
        while mother[j - 1] > 0:
            j = mother[j - 1]
        assert n2h[i] == j
    assert nb == sum(1 for m in mother if m == 0)


def test_mothers_must_precede_their_subnodes():
    node_0 = make_nodes([0, 1, 1], [5, 5, 5])
    # renumber so that a subnode comes before its mother
    bad = node_0.copy()
    bad["mother"][1:] = [3, 3, 0]
    bad["firstchild"][1:] = [0, 0, 1]
    bad["sister"][1:] = [2, 0, 0]
    with pytest.raises(ValueError, match="before its subnodes"):
        select_with_ms_method(bad)


@pytest.mark.parametrize("name", ["c1_tables", "c2_tables"])
def test_msm_on_reference_tables(name):
    """On the reference's own node tables and node ids (262,144 and 2,097,152 particles): the
    reference's particle-count checks (compute_neiKDtree_mod.py:281-296 and :381-392) hold --
    every node's own mass is the count of its particles, and every structure's top node
    carries the particles relabelled to it -- and every structure's main branch follows the
    heaviest subnode."""
    from conftest import load_golden
    g = load_golden(name)
    node_0 = node_table(g["real_table"], g["integer_table"])
    orig = node_0.copy()
    lp = g["liste_parts"]
    nn = len(node_0) - 1
    own_count = np.bincount(lp, minlength=nn + 1)
    r = select_with_ms_method(node_0)
    assert r.nstruct == r.nb_halos + r.nb_sub and r.nb_halos == int(np.sum(orig["mother"][1:] <= 0))
    labels = np.where(lp > 0, r.node_to_struct[np.maximum(lp, 1) - 1], 0)
    per_struct = np.bincount(labels, minlength=r.nstruct + 1)
    for s in range(1, r.nstruct + 1):
        members = np.nonzero(r.node_to_struct == s)[0] + 1
        top = members.min()
        assert node_0["mass"][top] == per_struct[s]
        assert np.sum(own_count[members]) == per_struct[s]
