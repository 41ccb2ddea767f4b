import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "halomaker-python_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_input(g):
    from amdahop import generate as G
    ps = G.config(str(g["config"]), n=int(g["n"]), nclumps=int(g["nclumps"]))
    assert G.digest(ps.pos, ps.mass) == str(g["input_digest"]), "generator output changed"
    return ps


def saddle_as_sets(off, isad, rho):
    """Per group: list of (rho, sorted neighbour ids with that rho) in decreasing rho.

    The reference sorts each group's saddle list with an unstable quicksort, so only the
    multiset per saddle value is a contract (its consumer only needs components)."""
    out = []
    for g in range(len(off) - 1):
        s, e = int(off[g]), int(off[g + 1])
        r = rho[s:e]
        if e > s:
            assert np.all(np.diff(r) <= 0), "saddle list not decreasing"
        d = {}
        for v, i in zip(r.tolist(), isad[s:e].tolist()):
            d.setdefault(v, []).append(i)
        out.append(sorted((v, tuple(sorted(ids))) for v, ids in d.items()))
    return out


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
