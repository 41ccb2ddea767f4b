"""Drop-in replacement of the reference's f2py extension ``compute_adaptahop``.

``src/compute_neiKDtree_mod.py:32-36`` does ``from compute_adaptahop import neikdtree``
and then calls exactly the surface declared in ``fortran/compute_adaptahop.pyf:1-70``:

* ``neikdtree.sync_from_init_adaptahop(...)``  (compute_neiKDtree_mod.py:183-197)
* ``neikdtree.sync_from_change_pos(...)``      (:207-210)
* ``neikdtree.sync_others(...)``               (:740-747)
* ``neikdtree.compute_adaptahop(pos.T, mass)`` (:48-50)
* reads ``liste_parts``, ``density``, ``nnodes``, ``real_table``, ``integer_table`` (:54-60)
* ``neikdtree.close()``                        (:65)

This module offers the same names, argument order, dtypes, shapes and Fortran order,
backed by libamdahop.so (MI355X kernels for kNN/density/hop/saddle, host C++ for the
structure tree).  Put ``halomaker-python_amd/`` on ``sys.path`` ahead of the f2py build
and the reference bridge runs unchanged.  Differences: a failed invariant raises
``RuntimeError`` instead of stopping the interpreter, and the array attributes are
views that stay valid until the next ``compute_adaptahop``/``close`` (as f2py's are).
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_pkg_root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _pkg_root not in sys.path:
    sys.path.insert(0, _pkg_root)

from amdahop._lib import lib, check, ptr  # noqa: E402

__all__ = ["neikdtree"]


def _view(p, shape, ctype, order="C"):
    if not p:
        return None
    count = int(np.prod(shape))
    if count == 0:
        return np.empty(shape, dtype=ctype, order=order)
    buf = (ctypes.c_byte * (count * np.dtype(ctype).itemsize)).from_address(p)
    a = np.frombuffer(buf, dtype=ctype, count=count)
    return a.reshape(shape, order=order)


class _NeiKDtree:
    """The ``neikdtree`` module object: five callables and five readable arrays."""

    # -- callables ---------------------------------------------------------------------
    def compute_adaptahop(self, pos_in, mass_in):
        pos = np.asarray(pos_in, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[0] != 3:
            raise RuntimeError("ERROR: pos_in should have shape (3,npart)")
        pos = np.asfortranarray(pos)       # (3,N) Fortran order == (N,3) row-major bytes
        mass = np.ascontiguousarray(np.asarray(mass_in, dtype=np.float64).ravel())
        check(lib().ahop_compute_adaptahop(ptr(pos), 3, pos.shape[1], ptr(mass), mass.shape[0]))

    def sync_from_change_pos(self, npart_in, nusedpart_in, epsilon_in, fudgepsilon_in, xlong_in,
                             boxsize2_in):
        check(lib().ahop_sync_from_change_pos(int(npart_in), int(nusedpart_in), float(epsilon_in),
                                              float(fudgepsilon_in), float(xlong_in),
                                              float(boxsize2_in)))

    def sync_from_init_adaptahop(self, npart_in, nmembthresh_in, nmembers_in, omegal_in,
                                 omega_lambda_f_in, omega0_in, omega_f_in, aexp_max_in, af_in,
                                 hubble_in, h_f_in, boxsize2_in, lf_in, xlong_in, ylong_in,
                                 zlong_in, xlongs2_in, ylongs2_in, zlongs2_in, hub_pt_in, aexp_in,
                                 pos_shift_in):
        shift = np.ascontiguousarray(np.asarray(pos_shift_in, dtype=np.float64).ravel()[:3])
        if shift.shape[0] != 3:
            raise RuntimeError("pos_shift_in must have 3 elements")
        vals = [float(v) for v in (omegal_in, omega_lambda_f_in, omega0_in, omega_f_in,
                                   aexp_max_in, af_in, hubble_in, h_f_in, boxsize2_in, lf_in,
                                   xlong_in, ylong_in, zlong_in, xlongs2_in, ylongs2_in,
                                   zlongs2_in, hub_pt_in, aexp_in)]
        check(lib().ahop_sync_from_init_adaptahop(int(npart_in), int(nmembthresh_in),
                                                  int(nmembers_in), *vals, ptr(shift)))

    def sync_others(self, verbose_in, megaverbose_in, npart_in, nbpes_in, rho_threshold_in,
                    massp_in, boxsize_in, nhop_in, nvoisins_in, fudge_in, alphap_in, method_in,
                    nlevelmax_in, optimize_nodes_in):
        m = method_in.encode() if isinstance(method_in, str) else bytes(method_in)
        check(lib().ahop_sync_others(int(bool(verbose_in)), int(bool(megaverbose_in)),
                                     int(npart_in), int(nbpes_in), float(rho_threshold_in),
                                     float(massp_in), float(boxsize_in), int(nhop_in),
                                     int(nvoisins_in), float(fudge_in), float(alphap_in), m,
                                     len(m), int(nlevelmax_in), int(bool(optimize_nodes_in))))

    def close(self):
        check(lib().ahop_close())

    # -- readable module arrays --------------------------------------------------------
    @property
    def liste_parts(self):
        n = ctypes.c_int64(0)
        p = lib().ahop_liste_parts(ctypes.byref(n))
        return _view(p, (n.value,), np.int32)

    @property
    def density(self):
        n = ctypes.c_int64(0)
        p = lib().ahop_density(ctypes.byref(n))
        return _view(p, (n.value,), np.float64)

    @property
    def nnodes(self):
        return np.array(lib().ahop_nnodes(), dtype=np.int32)

    @property
    def real_table(self):
        n = ctypes.c_int64(0)
        p = lib().ahop_real_table(ctypes.byref(n))
        return _view(p, (8, n.value), np.float64, order="F")

    @property
    def integer_table(self):
        n = ctypes.c_int64(0)
        p = lib().ahop_integer_table(ctypes.byref(n))
        return _view(p, (6, n.value), np.int32, order="F")

    # -- extensions (not in the f2py module) -------------------------------------------
    def set_precision(self, mode):
        """'mixed' (default): FP32 box-relative search + certified FP64 re-rank; 'fp64'."""
        check(lib().ahop_set_precision({"mixed": 1, "fp64": 0}[mode]))

    def set_bucket(self, bucket):
        check(lib().ahop_set_bucket(int(bucket)))

    def groups(self):
        """(group ids, hop targets) of the local-maxima stage, original order."""
        n = ctypes.c_int64(0)
        lib().ahop_density(ctypes.byref(n))
        gid = np.empty(n.value, dtype=np.int32)
        hop = np.empty(n.value, dtype=np.int32)
        check(lib().ahop_get_groups(ptr(gid), ptr(hop), n.value))
        return gid, hop

    def saddle(self):
        """Saddle CSR (offset 0-based, neighbour group ids 1-based, saddle densities)."""
        ng = lib().ahop_ngroups()
        ne = lib().ahop_nsaddle()
        off = np.empty(ng + 1, dtype=np.int32)
        isad = np.empty(max(ne, 1), dtype=np.int32)
        rho = np.empty(max(ne, 1), dtype=np.float64)
        check(lib().ahop_get_saddle(ptr(off), ptr(isad), ptr(rho)))
        return off, isad[:ne], rho[:ne]


neikdtree = _NeiKDtree()
