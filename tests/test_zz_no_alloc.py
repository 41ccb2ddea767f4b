"""The allocation-free adjoint (ahop_density_adjoint_device_noalloc) against the allocating
entry point (ahop_density_adjoint_device), both called through the C ABI on the same frozen
neighbour rows: two calls in a row, the second on a smaller box, with and without the Morton
order, and the size limit.  Both add the neighbours' terms with FP64 atomics in an order the
hardware chooses, so the two agree to rounding: every gradient component within 1e-11 of the
largest component of its array (FP64 sums of a few hundred terms)."""
import ctypes
import sys

import numpy as np
import pytest

from amdahop import generate as G

pytestmark = pytest.mark.gpu

TOL = 1e-11


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(n, nclumps, seed, gpu):
    import torch
    from amdahop import knn_device
    ps = G.config("c1", n=n, nclumps=nclumps, seed=seed)
    pos = torch.from_numpy(ps.pos).to(gpu)
    mass = torch.from_numpy(ps.mass).to(gpu)
    dens, nbr, _, order = knn_device(pos, mass, ps.nvoisins, return_order=True)
    w = torch.from_numpy(np.random.default_rng(seed).normal(size=n)).to(gpu)
    return pos, mass, nbr.contiguous(), order.contiguous(), (w / dens).contiguous()


def _call(name, pos, mass, nbr, order, g, gpos, gmass):
    import torch
    from amdahop._lib import lib
    torch.cuda.synchronize()
    n, k = int(mass.shape[0]), int(nbr.shape[1])
    args = [ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()), n, k,
            ctypes.c_void_p(nbr.data_ptr()),
            ctypes.c_void_p(order.data_ptr() if order is not None else None),
            ctypes.c_void_p(g.data_ptr()), 1.0, 1.0, 1.0, ctypes.c_void_p(gpos.data_ptr()),
            ctypes.c_void_p(gmass.data_ptr())]
    if name == "ahop_density_adjoint_device":
        args.append(None)
    return getattr(lib(), name)(*args)


def _compare(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    err = np.abs(a - b).max() / np.abs(b).max()
    assert err <= TOL, err
    return err


@pytest.mark.parametrize("with_order", [True, False])
def test_noalloc_adjoint_equals_the_allocating_one(gpu, with_order):
    import torch
    from amdahop._lib import lib
    for n, ncl, seed in ((262144, 150, 21), (100000, 60, 22)):
        log(f"no-alloc adjoint n={n} order={with_order}")
        pos, mass, nbr, order, g = _case(n, ncl, seed, gpu)
        o = order if with_order else None
        ref_p = torch.empty((n, 3), dtype=torch.float64, device=gpu)
        ref_m = torch.empty(n, dtype=torch.float64, device=gpu)
        got_p = torch.full((n, 3), float("nan"), dtype=torch.float64, device=gpu)
        got_m = torch.full((n,), float("nan"), dtype=torch.float64, device=gpu)
        assert _call("ahop_density_adjoint_device", pos, mass, nbr, o, g, ref_p, ref_m) == 0, \
            lib().ahop_last_error()
        assert _call("ahop_density_adjoint_device_noalloc", pos, mass, nbr, o, g, got_p,
                     got_m) == 0, lib().ahop_last_error()
        torch.cuda.synchronize()
        ep, em = _compare(got_p, ref_p), _compare(got_m, ref_m)
        log(f"  max differences relative to the largest component: pos {ep:.2e}, mass {em:.2e}")


def test_noalloc_adjoint_errors(gpu):
    """Sizes beyond the fixed scratch and a broken order are errors, and the entry point
    still works after them."""
    import torch
    from amdahop._lib import lib
    pos, mass, nbr, order, g = _case(20000, 10, 23, gpu)
    n = int(mass.shape[0])
    gp = torch.empty((n, 3), dtype=torch.float64, device=gpu)
    gm = torch.empty(n, dtype=torch.float64, device=gpu)
    # a size above the limit is rejected on the host, before any device access
    args = [ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()), 4194305,
            int(nbr.shape[1]), ctypes.c_void_p(nbr.data_ptr()), None,
            ctypes.c_void_p(g.data_ptr()), 1.0, 1.0, 1.0, ctypes.c_void_p(gp.data_ptr()),
            ctypes.c_void_p(gm.data_ptr())]
    assert lib().ahop_density_adjoint_device_noalloc(*args) != 0
    assert b"at most 4194304" in lib().ahop_last_error()
    bad = order.clone()
    bad[5] = bad[6]                                   # not a permutation
    assert _call("ahop_density_adjoint_device_noalloc", pos, mass, nbr, bad, g, gp, gm) != 0
    assert b"permutation" in lib().ahop_last_error()
    ref_p = torch.empty_like(gp)
    ref_m = torch.empty_like(gm)
    assert _call("ahop_density_adjoint_device", pos, mass, nbr, order, g, ref_p, ref_m) == 0
    assert _call("ahop_density_adjoint_device_noalloc", pos, mass, nbr, order, g, gp, gm) == 0
    _compare(gp, ref_p)
    _compare(gm, ref_m)
