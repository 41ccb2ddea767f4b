"""Checker for the AdaptaHOP hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package.  It wraps two CPU libraries:

* ``_build/liboracle_ahop.so`` -- our C restatement (ahop_oracle.c) of
  reference fortran/compute_adaptahop.f90 (kNN :2841-2962, tree :2967-3221,
  density :682-744/:2816-2838/:3270-3287, local maxima :748-881, saddle list
  :2433-2813);
* ``_ref/libref_ahop.so`` -- the reference Fortran itself, compiled from the
  reference source by oracle/Makefile, driven through ref_driver.f90.

Pinning (each checked in tests/test_oracle.py): the restatement equals the
compiled reference bit for bit on the committed golden fixtures and live when
_ref is built; kNN equals an O(N^2) brute force on small inputs; density on a
perfect periodic lattice equals its closed form; hop/group invariants and
saddle symmetry hold.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle_ahop.so")
REF_SO = os.path.join(HERE, "_ref", "libref_ahop.so")

_i64, _i32, _dbl, _vp = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p


def build(ref=True):
    """Compile the restatement (and the reference when its source is present)."""
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"), "oracle"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"), "ref"], check=True)


def _ptr(a):
    return a.ctypes.data_as(_vp)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_create.restype = _vp
        L.orc_create.argtypes = [_i64, _vp, _vp, ctypes.c_int, ctypes.c_int, _dbl, _dbl,
                                 ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [_vp]
        L.orc_ncells.restype = _i64
        L.orc_ncells.argtypes = [_vp]
        L.orc_knn.argtypes = [_vp, _i64, _vp, _vp]
        L.orc_density.restype = _i64
        L.orc_density.argtypes = [_vp, _vp, _vp]
        L.orc_density_at.argtypes = [_vp, _i64, _vp, _vp]
        L.orc_maxima.restype = _i32
        L.orc_maxima.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.orc_saddle.restype = _i64
        L.orc_saddle.argtypes = [_vp, _vp, _vp, _i32, _vp, _vp, _vp, _i64]
        L.orc_knn_brute.argtypes = [_i64, _vp, _dbl, _i64, ctypes.c_int, _vp, _vp]
        L.orc_indexx.argtypes = [_i32, _vp, _vp]
        _lib = L
    return _lib


class Oracle:
    """Stage-by-stage restatement on one particle set (positions (N,3), box side xlong)."""

    def __init__(self, pos, mass, nvoisins, nhop=16, rho_threshold=80.0, xlong=1.0,
                 nlevelmax=30, nthreads=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.n = self.mass.shape[0]
        self.k, self.nhop, self.rho_thr = int(nvoisins), int(nhop), float(rho_threshold)
        nthreads = nthreads or min(os.cpu_count() or 1, 16)
        self._c = lib().orc_create(self.n, _ptr(self.pos), _ptr(self.mass), self.k, self.nhop,
                                   self.rho_thr, float(xlong), int(nlevelmax), int(nthreads))
        if not self._c:
            raise ValueError("oracle: invalid sizes (need N > 100, 1 <= nvoisins < N, nhop <= nvoisins)")

    def close(self):
        if self._c:
            lib().orc_destroy(self._c)
            self._c = None

    __del__ = close

    def knn(self, i):
        d2 = np.empty(self.k)
        ids = np.empty(self.k, dtype=np.int32)
        lib().orc_knn(self._c, int(i), _ptr(d2), _ptr(ids))
        return d2, ids

    def density(self, with_nbrs=False):
        rho = np.empty(self.n)
        nb = np.empty((self.n, self.nhop), dtype=np.int32) if with_nbrs else None
        nact = lib().orc_density(self._c, _ptr(rho), _ptr(nb) if with_nbrs else None)
        return (rho, int(nact), nb) if with_nbrs else (rho, int(nact))

    def density_at(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        out = np.empty(idx.shape[0])
        lib().orc_density_at(self._c, idx.shape[0], _ptr(idx), _ptr(out))
        return out

    def maxima(self, rho):
        hop = np.empty(self.n, dtype=np.int32)
        gid = np.empty(self.n, dtype=np.int32)
        dg = np.empty(self.n)
        ng = lib().orc_maxima(self._c, _ptr(rho), _ptr(hop), _ptr(gid), _ptr(dg))
        return hop, gid, dg[:ng].copy(), int(ng)

    def saddle(self, rho, gid, ngroups):
        off = np.empty(ngroups + 1, dtype=np.int32)
        tot = lib().orc_saddle(self._c, _ptr(rho), _ptr(gid), ngroups, _ptr(off), None, None, 0)
        isad = np.empty(max(tot, 1), dtype=np.int32)
        rs = np.empty(max(tot, 1))
        lib().orc_saddle(self._c, _ptr(rho), _ptr(gid), ngroups, _ptr(off), _ptr(isad), _ptr(rs), tot)
        return off, isad[:tot], rs[:tot]


def knn_brute(pos, i, k, xlong=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    d2 = np.empty(k)
    ids = np.empty(k, dtype=np.int32)
    lib().orc_knn_brute(pos.shape[0], _ptr(pos), float(xlong), int(i), int(k), _ptr(d2), _ptr(ids))
    return d2, ids


def indexx(arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    ix = np.empty(arr.shape[0], dtype=np.int32)
    lib().orc_indexx(arr.shape[0], _ptr(arr), _ptr(ix))
    return ix


# --------------------------------------------------------------------------
# The compiled reference (present only where oracle/_ref was built).
# --------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_SO)
        L.ref_configure.argtypes = [_i64, ctypes.c_int, ctypes.c_int, _dbl, _dbl, _dbl, _dbl,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, _dbl,
                                    ctypes.c_int, ctypes.c_int]
        for f in ("ref_run_full", "ref_stage_begin"):
            getattr(L, f).argtypes = [_i64, _vp, _vp]
        L.ref_knn.argtypes = [ctypes.c_int, _vp, _vp]
        L.ref_get_density.argtypes = [_vp, _i64]
        L.ref_get_liste_parts.argtypes = [_vp, _i64]
        L.ref_get_densityg.argtypes = [_vp, ctypes.c_int]
        L.ref_get_saddle.argtypes = [_vp, _vp, _vp]
        L.ref_get_tables.argtypes = [_vp, _vp]
        for f in ("ref_ngroups", "ref_nnodes", "ref_nactive", "ref_saddle_total"):
            getattr(L, f).restype = ctypes.c_int
        _ref = L
    return _ref


class Reference:
    """The reference Fortran run on one particle set, the way compute_neiKDtree_mod.py
    drives it (positions passed as the (3,N) Fortran array pos.T)."""

    def __init__(self, pos, mass, nvoisins, nhop=16, rho_threshold=80.0, xlong=1.0,
                 fudge=4.0, alphap=1.0, fudgepsilon=0.001, nmembers=20, nlevelmax=30,
                 nthreads=None):
        self.pos_f = np.asfortranarray(np.asarray(pos, dtype=np.float64).T)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.n = self.mass.shape[0]
        self.k = int(nvoisins)
        self.args = (self.n, int(nvoisins), int(nhop), float(rho_threshold), float(fudge),
                     float(alphap), float(fudgepsilon), int(nmembers), int(nlevelmax),
                     int(nthreads or min(os.cpu_count() or 1, 16)), float(xlong))

    def staged(self, knn_ids=None, saddle=True):
        """tree -> density -> maxima [-> saddle]; returns a dict of stage outputs."""
        L = ref_lib()
        L.ref_configure(*self.args, 0, 0)
        L.ref_stage_begin(self.n, _ptr(self.pos_f), _ptr(self.mass))
        out = {}
        L.ref_stage_tree()
        L.ref_stage_density()
        rho = np.empty(self.n)
        L.ref_get_density(_ptr(rho), self.n)
        out["density"] = rho
        if knn_ids is not None:
            ks = []
            for i in knn_ids:
                d2 = np.empty(self.k + 1)
                ids = np.empty(self.k, dtype=np.int32)
                L.ref_knn(int(i) + 1, _ptr(d2), _ptr(ids))
                ks.append((d2[1:].copy(), ids - 1))
            out["knn"] = ks
        L.ref_stage_maxima()
        gid = np.empty(self.n, dtype=np.int32)
        L.ref_get_liste_parts(_ptr(gid), self.n)
        ng = L.ref_ngroups()
        dg = np.empty(max(ng, 1))
        L.ref_get_densityg(_ptr(dg), ng)
        out.update(gid=gid, ngroups=ng, densityg=dg[:ng].copy())
        if saddle:
            L.ref_stage_saddle()
            tot = L.ref_saddle_total()
            off = np.empty(ng + 1, dtype=np.int32)
            isad = np.empty(max(tot, 1), dtype=np.int32)
            rs = np.empty(max(tot, 1))
            L.ref_get_saddle(_ptr(off), _ptr(isad), _ptr(rs))
            out.update(saddle_offset=off - 1, saddle_isad=isad[:tot].copy(), saddle_rho=rs[:tot].copy())
        L.ref_close()
        return out

    def time_knn_density_hop(self):
        """Wall seconds of the reference's tree + density + local-maxima stages
        (f90:2967, :682, :748) -- its kNN+density+hop path as shipped."""
        import time
        L = ref_lib()
        L.ref_configure(*self.args, 0, 0)
        L.ref_stage_begin(self.n, _ptr(self.pos_f), _ptr(self.mass))
        t0 = time.perf_counter()
        L.ref_stage_tree()
        L.ref_stage_density()
        L.ref_stage_maxima()
        t1 = time.perf_counter()
        L.ref_close()
        return t1 - t0

    def timed_full(self, optimize_nodes=False):
        """The public compute_adaptahop entry (f90:495-543) with the reference's verbose
        prints on: its own per-stage seconds (system_clock, printed to stderr with 10 ms
        resolution: tree :3021, density :742, local maxima :879, saddle list :902, create
        nodes :961) and the wall seconds of the whole call.  optimize_nodes selects the
        reference's blocked-prefix create-nodes path (opt_create_nodes_omp_root, f90:955-958)."""
        import re
        import sys
        import tempfile
        import time
        L = ref_lib()
        L.ref_configure(*self.args, 1, 1 if optimize_nodes else 0)
        sys.stderr.flush()
        with tempfile.TemporaryFile() as tf:
            saved = os.dup(2)
            os.dup2(tf.fileno(), 2)
            try:
                t0 = time.perf_counter()
                L.ref_run_full(self.n, _ptr(self.pos_f), _ptr(self.mass))
                wall = time.perf_counter() - t0
            finally:
                os.dup2(saved, 2)
                os.close(saved)
            tf.seek(0)
            text = tf.read().decode(errors="replace")
        nn = L.ref_nnodes()
        L.ref_close()
        names = {"create the tree structure": "tree", "compute mean density": "density",
                 "find local maxima": "maxima", "saddle_list": "saddle",
                 "create nodes": "create_nodes"}
        out = {"wall": wall, "nnodes": nn}
        for m in re.finditer(r"-->\s*([0-9.]+) seconds to (.+)", text):
            key = names.get(m.group(2).strip())
            if key:
                out[key] = float(m.group(1))
        return out

    def full(self, optimize_nodes=False):
        """The public compute_adaptahop entry (f90:495-543): final node ids and tables.
        optimize_nodes selects the blocked-prefix create-nodes path (f90:955-958)."""
        L = ref_lib()
        L.ref_configure(*self.args, 0, 1 if optimize_nodes else 0)
        L.ref_run_full(self.n, _ptr(self.pos_f), _ptr(self.mass))
        lp = np.empty(self.n, dtype=np.int32)
        L.ref_get_liste_parts(_ptr(lp), self.n)
        rho = np.empty(self.n)
        L.ref_get_density(_ptr(rho), self.n)
        nn = L.ref_nnodes()
        rt = np.zeros((8, max(nn, 1)), order="F")
        it = np.zeros((6, max(nn, 1)), dtype=np.int32, order="F")
        L.ref_get_tables(_ptr(rt), _ptr(it))
        L.ref_close()
        return dict(liste_parts=lp, density=rho, nnodes=nn, real_table=rt[:, :nn].copy(order="F"),
                    integer_table=it[:, :nn].copy(order="F"))
