"""Pins of the frozen-neighbour SPH density and its FP64 gradient (oracle/sph_adjoint.py):
the density against the C oracle (bit-exact, same rows), the gradient against central
finite differences of that density."""
import numpy as np
import pytest

import oracle as O
from oracle import sph_adjoint as SA


def rows_from_oracle(o, n):
    return np.array([o.knn(i)[1] for i in range(n)], dtype=np.int64)


def small_system(seed, n, edge=False):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, size=(n, 3))
    # a clump so that densities span a range
    pos[: n // 3] = 0.05 * rng.normal(size=(n // 3, 3)) + (0.49 if edge else 0.1)
    pos = np.mod(pos + 0.5, 1.0) - 0.5
    mass = np.exp(0.3 * rng.normal(size=n)) / n
    return pos, mass


def test_frozen_density_equals_the_c_oracle():
    pos, mass = small_system(0, 3000)
    o = O.Oracle(pos, mass, 16)
    rho, _ = o.density()
    nbr = rows_from_oracle(o, pos.shape[0])
    assert nbr.min() >= 0 and nbr.max() < pos.shape[0]
    assert np.array_equal(SA.density_frozen(pos, mass, nbr), rho)
    o.close()


@pytest.mark.parametrize("edge", [False, True])
def test_gradient_equals_central_differences(edge):
    n, k = 400, 8
    pos, mass = small_system(1, n, edge)
    o = O.Oracle(pos, mass, k, nhop=k)
    nbr = rows_from_oracle(o, n)
    o.close()
    rng = np.random.default_rng(2)
    g = rng.normal(size=n)
    gpos, gmass = SA.density_grad(pos, mass, nbr, g)
    h = 1e-7
    for p in rng.choice(n, size=25, replace=False):
        aff = SA.affected(nbr, p)

        def loss(P, M):
            return np.sum(g[aff] * SA.density_frozen(P, M, nbr, idx=aff))
        for ax in range(3):
            Pp, Pm = pos.copy(), pos.copy()
            Pp[p, ax] += h
            Pm[p, ax] -= h
            fd = (loss(Pp, mass) - loss(Pm, mass)) / (2 * h)
            assert fd == pytest.approx(gpos[p, ax], rel=2e-6, abs=1e-6 * np.abs(gpos).max())
        hm = 1e-4 * mass[p]
        Mp, Mm = mass.copy(), mass.copy()
        Mp[p] += hm
        Mm[p] -= hm
        fd = (loss(pos, Mp) - loss(pos, Mm)) / (2 * hm)
        assert fd == pytest.approx(gmass[p], rel=1e-8, abs=1e-9 * np.abs(gmass).max())


def test_gradient_of_a_translated_system_is_unchanged():
    # moving every particle by the same vector across the box edge changes nothing
    n, k = 300, 8
    pos, mass = small_system(3, n)
    o = O.Oracle(pos, mass, k, nhop=k)
    nbr = rows_from_oracle(o, n)
    o.close()
    g = np.random.default_rng(4).normal(size=n)
    a = SA.density_grad(pos, mass, nbr, g)
    p2 = np.mod(pos + np.array([0.37, -0.21, 0.45]) + 0.5, 1.0) - 0.5
    b = SA.density_grad(p2, mass, nbr, g)
    assert np.allclose(a[0], b[0], rtol=1e-8, atol=1e-10 * np.abs(a[0]).max())
    assert np.allclose(a[1], b[1], rtol=1e-10)
    # and the gradient with respect to a uniform translation sums to zero
    assert np.abs(a[0].sum(axis=0)).max() < 1e-9 * np.abs(a[0]).max()
