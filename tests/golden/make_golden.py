"""Writes the golden fixtures of the AdaptaHOP hot path from the compiled reference.

The reference Fortran (fortran/compute_adaptahop.f90) is built from the reference
source by ``make -f oracle/Makefile ref`` into oracle/_ref/libref_ahop.so and driven
through oracle/ref_driver.f90: stage by stage for the kNN lists, densities, group ids
and saddle CSR, and through its public compute_adaptahop entry for the final node ids
and node tables.  Inputs are regenerated from amdahop.generate; their SHA-256 is stored
so a generator change is detected.  Run from the repository root:

    python tests/golden/make_golden.py [--tables-only | --opt-only]

``--opt-only`` writes only the node tables of the reference's moment-prefix create-nodes
path (optimize_nodes) at the C1 and C2 sizes, ``<case>_opt.npz``.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "halomaker-python_amd"))

import oracle as O  # noqa: E402
from oracle.hop_saddle import csr_as_pairs  # noqa: E402
from amdahop import generate as G  # noqa: E402

CASES = {
    # name: (config, n, nclumps)
    "clumpy_16k": ("c1", 16384, 10),
    "mixed_32k": ("c3", 32768, 18),
}
PARAMS = dict(nhop=16, rho_threshold=80.0, fudge=4.0, alphap=1.0, fudgepsilon=0.001,
              nmembers=20, nlevelmax=30)


# Node ids and node tables of the reference at the C1 and C2 sizes (262,144 and 2,097,152
# particles, k=32).  The stage outputs the host structure tree starts from (densities, group
# ids, saddle pairs) are not stored: a test regenerates them with the oracle and checks their
# SHA-256 against the reference's, stored here (the saddle pairs as (g1, g2, s) sorted by
# (g1, g2), the multiset the reference's unstable per-group sort leaves as the contract).
TABLE_CASES = {
    "c1_tables": ("c1", 262144, 150),
    "c2_tables": ("c2", 2097152, 1200),
}


def _sha(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def opt_tables(out_dir):
    """The same inputs through the public entry with optimize_nodes on (f90:954-958): node ids
    and integer tables equal the default path's (checked here), the real tables are stored."""
    for name, (cfg, n, ncl) in TABLE_CASES.items():
        ps = G.config(cfg, n=n, nclumps=ncl)
        R = O.Reference(ps.pos, ps.mass, ps.nvoisins, **PARAMS, nthreads=8)
        full = R.full(optimize_nodes=True)
        base = np.load(os.path.join(out_dir, f"{name}.npz"))
        if str(base["input_digest"]) != G.digest(ps.pos, ps.mass):
            raise SystemExit(f"{name}: inputs differ from the default-path golden")
        if not (np.array_equal(full["liste_parts"], base["liste_parts"])
                and np.array_equal(full["integer_table"], base["integer_table"])):
            raise SystemExit(f"{name}: optimize_nodes changed the node ids or integer tables")
        np.savez_compressed(os.path.join(out_dir, f"{name}_opt.npz"), nnodes=full["nnodes"],
                            real_table=full["real_table"],
                            integer_table=full["integer_table"])
        ndiff = int(np.sum(full["real_table"] != base["real_table"]))
        print(name, "optimize_nodes: nnodes", full["nnodes"], "real entries differing", ndiff)


def main():
    out_dir = os.path.dirname(os.path.abspath(__file__))
    if "--opt-only" in sys.argv:
        opt_tables(out_dir)
        return
    for name, (cfg, n, ncl) in TABLE_CASES.items():
        ps = G.config(cfg, n=n, nclumps=ncl)
        R = O.Reference(ps.pos, ps.mass, ps.nvoisins, **PARAMS, nthreads=8)
        st = R.staged()
        full = R.full()
        if not np.array_equal(st["density"], full["density"]):
            raise SystemExit("staged and full reference runs disagree")
        np.savez_compressed(
            os.path.join(out_dir, f"{name}.npz"),
            config=cfg, n=n, nclumps=ncl, nvoisins=ps.nvoisins,
            input_digest=G.digest(ps.pos, ps.mass),
            stage_digest=_sha(st["density"], st["gid"], st["densityg"]),
            saddle_digest=_sha(*csr_as_pairs(st["saddle_offset"], st["saddle_isad"],
                                             st["saddle_rho"])),
            ngroups=st["ngroups"], nsaddle=len(st["saddle_isad"]),
            liste_parts=full["liste_parts"], nnodes=full["nnodes"],
            real_table=full["real_table"], integer_table=full["integer_table"])
        print(name, "ngroups", st["ngroups"], "nnodes", full["nnodes"],
              "saddle entries", len(st["saddle_isad"]))
    opt_tables(out_dir)
    if "--tables-only" in sys.argv:
        return
    for name, (cfg, n, ncl) in CASES.items():
        ps = G.config(cfg, n=n, nclumps=ncl)
        R = O.Reference(ps.pos, ps.mass, ps.nvoisins, **PARAMS, nthreads=4)
        rng = np.random.default_rng(123)
        sample = np.sort(rng.choice(n, size=32, replace=False))
        st = R.staged(knn_ids=sample)
        full = R.full()
        if not np.array_equal(st["density"], full["density"]):
            raise SystemExit("staged and full reference runs disagree")
        np.savez_compressed(
            os.path.join(out_dir, f"{name}.npz"),
            config=cfg, n=n, nclumps=ncl, nvoisins=ps.nvoisins,
            input_digest=G.digest(ps.pos, ps.mass),
            density=st["density"], gid=st["gid"], ngroups=st["ngroups"], densityg=st["densityg"],
            saddle_offset=st["saddle_offset"], saddle_isad=st["saddle_isad"],
            saddle_rho=st["saddle_rho"],
            knn_sample=sample, knn_d2=np.array([d for d, _ in st["knn"]]),
            knn_ids=np.array([i for _, i in st["knn"]]),
            liste_parts=full["liste_parts"], nnodes=full["nnodes"],
            real_table=full["real_table"], integer_table=full["integer_table"])
        print(name, "ngroups", st["ngroups"], "nnodes", full["nnodes"],
              "saddle entries", len(st["saddle_isad"]))


if __name__ == "__main__":
    main()
