"""Frozen-neighbour SPH density adjoint on the GPU (ahop_knn_device +
ahop_density_adjoint_device) against the FP64 NumPy gradient and central finite
differences, to a relative 1e-5."""
import numpy as np
import pytest

import oracle as O
from oracle import sph_adjoint as SA
from amdahop import generate as G

pytestmark = pytest.mark.gpu

TOL = 1e-5


def log(msg):
    print(msg, flush=True)


def to_dev(ps, gpu):
    import torch
    return torch.from_numpy(ps.pos).to(gpu), torch.from_numpy(ps.mass).to(gpu)


@pytest.mark.parametrize("mixed", [True, False])
def test_knn_rows_and_density_match_the_oracle(gpu, mixed):
    from amdahop import knn_device
    ps = G.config("c1", n=20000, nclumps=12, seed=5)
    log(f"knn rows c1 n=20000 mixed={mixed}")
    pos, mass = to_dev(ps, gpu)
    dens, nbr, info = knn_device(pos, mass, 32, mixed=mixed)
    o = O.Oracle(ps.pos, ps.mass, 32)
    rho, _ = o.density()
    assert np.array_equal(dens.cpu().numpy(), rho)
    nb = nbr.cpu().numpy()
    for i in np.random.default_rng(0).choice(ps.n, 300, replace=False):
        d2, ids = o.knn(i)
        assert np.array_equal(nb[i], ids), i
    o.close()
    # the rows reproduce the density through the frozen-row restatement
    assert np.array_equal(SA.density_frozen(ps.pos, ps.mass, nb.astype(np.int64)), rho)


def gradient_error(g_dev, g_ref):
    """max over particles of |g_dev - g_ref| / max(|g_ref|, 1e-3 * rms(g_ref))."""
    ref = np.asarray(g_ref).reshape(len(g_ref), -1)
    dev = np.asarray(g_dev).reshape(len(g_ref), -1)
    norm = np.sqrt((ref ** 2).sum(1))
    floor = 1e-3 * np.sqrt(np.mean(norm ** 2))
    err = np.sqrt(((dev - ref) ** 2).sum(1)) / np.maximum(norm, floor)
    return float(err.max()), float(np.median(err))


def test_adjoint_matches_the_fp64_gradient(gpu):
    import torch
    from amdahop import density_adjoint, knn_device
    ps = G.config("c1")
    log(f"adjoint c1 n={ps.n} k=32")
    pos, mass = to_dev(ps, gpu)
    dens, nbr, _, order = knn_device(pos, mass, 32, return_order=True)
    # the Morton order is a permutation of the particles
    assert np.array_equal(np.sort(order.cpu().numpy()), np.arange(ps.n))
    w = np.random.default_rng(3).normal(size=ps.n)
    rho = dens.cpu().numpy()
    grad_rho = w / rho
    ref_p, ref_m = SA.density_grad(ps.pos, ps.mass, nbr.cpu().numpy().astype(np.int64), grad_rho)
    # per-thread global atomics, and per-workgroup LDS sums over a window of the order
    for o in (None, order):
        gp, gm, ms = density_adjoint(pos, mass, nbr, torch.from_numpy(grad_rho).to(gpu), order=o)
        log(f"adjoint device ms {ms:.3f} (Morton order: {o is not None})")
        ep = gradient_error(gp.cpu().numpy(), ref_p)
        em = gradient_error(gm.cpu().numpy(), ref_m)
        log(f"position gradient error max {ep[0]:.2e} median {ep[1]:.2e}; "
            f"mass gradient error max {em[0]:.2e} median {em[1]:.2e}")
        assert ep[0] < TOL and em[0] < TOL


def test_adjoint_matches_central_differences_on_1000_particles(gpu):
    import torch
    from amdahop import density_adjoint, knn_device
    ps = G.config("c5")
    log(f"adjoint vs finite differences c5 n={ps.n} k=32")
    pos, mass = to_dev(ps, gpu)
    dens, nbr, _ = knn_device(pos, mass, 32)
    nb = nbr.cpu().numpy().astype(np.int64)
    rng = np.random.default_rng(3)
    w = rng.normal(size=ps.n)
    rho = dens.cpu().numpy()
    gp, gm, _ = density_adjoint(pos, mass, nbr, torch.from_numpy(w / rho).to(gpu))
    gp, gm = gp.cpu().numpy(), gm.cpu().numpy()
    # reverse rows: which particles list p as a neighbour
    flat = nb.ravel()
    order = np.argsort(flat, kind="stable")
    starts = np.searchsorted(flat[order], np.arange(ps.n + 1))
    h = 1e-7
    worst = 0.0
    for p in rng.choice(ps.n, size=1000, replace=False):
        aff = np.union1d([p], order[starts[p]:starts[p + 1]] // nb.shape[1])

        def loss(P, M):
            return np.sum(w[aff] * np.log(SA.density_frozen(P, M, nb, idx=aff)))
        fd = np.empty(4)
        for ax in range(3):
            x0 = ps.pos[p, ax]
            ps.pos[p, ax] = x0 + h
            lp = loss(ps.pos, ps.mass)
            ps.pos[p, ax] = x0 - h
            lm = loss(ps.pos, ps.mass)
            ps.pos[p, ax] = x0
            fd[ax] = (lp - lm) / (2 * h)
        m0, hm = ps.mass[p], 1e-4 * ps.mass[p]
        ps.mass[p] = m0 + hm
        lp = loss(ps.pos, ps.mass)
        ps.mass[p] = m0 - hm
        lm = loss(ps.pos, ps.mass)
        ps.mass[p] = m0
        fd[3] = (lp - lm) / (2 * hm)
        ep = np.linalg.norm(gp[p] - fd[:3]) / max(np.linalg.norm(fd[:3]), 1e-300)
        em = abs(gm[p] - fd[3]) / max(abs(fd[3]), 1e-300)
        worst = max(worst, ep, em)
    log(f"worst relative difference to central differences: {worst:.2e}")
    assert worst < TOL


def test_autograd_function_uses_the_adjoint(gpu):
    import torch
    from amdahop import density_adjoint, knn_device, sph_density
    ps = G.config("c1", n=50000, nclumps=30, seed=6)
    pos, mass = to_dev(ps, gpu)
    pos.requires_grad_(True)
    mass.requires_grad_(True)
    rho, nbr = sph_density(pos, mass, 32)
    w = torch.from_numpy(np.random.default_rng(7).normal(size=ps.n)).to(gpu)
    loss = (w * torch.log(rho)).sum()
    loss.backward()
    d2, nb2, _ = knn_device(pos.detach(), mass.detach(), 32)
    assert torch.equal(rho.detach(), d2) and torch.equal(nbr, nb2)
    gp, gm, _ = density_adjoint(pos.detach(), mass.detach(), nbr, w / rho.detach())
    # the adjoint adds with atomics, so two calls agree to rounding, not bit for bit
    assert torch.allclose(pos.grad, gp, rtol=0, atol=1e-12 * float(gp.abs().max()))
    assert torch.allclose(mass.grad, gm, rtol=0, atol=1e-12 * float(gm.abs().max()))


def test_bad_neighbour_rows_are_rejected(gpu):
    import torch
    from amdahop import density_adjoint
    ps = G.config("c1", n=5000, nclumps=4, seed=8)
    pos, mass = to_dev(ps, gpu)
    nbr = torch.zeros((ps.n, 8), dtype=torch.int32, device=gpu)
    nbr[7, 3] = ps.n                       # one id past the end
    with pytest.raises(RuntimeError, match="outside"):
        density_adjoint(pos, mass, nbr, torch.ones(ps.n, dtype=torch.float64, device=gpu))
    # an order with an id outside the particles, and one that repeats an id
    nbr[7, 3] = 0
    order = torch.arange(ps.n, dtype=torch.int32, device=gpu)
    order[11] = -1
    with pytest.raises(RuntimeError, match="permutation"):
        density_adjoint(pos, mass, nbr, torch.ones(ps.n, dtype=torch.float64, device=gpu),
                        order=order)
    order[11] = 12
    with pytest.raises(RuntimeError, match="permutation"):
        density_adjoint(pos, mass, nbr, torch.ones(ps.n, dtype=torch.float64, device=gpu),
                        order=order)
