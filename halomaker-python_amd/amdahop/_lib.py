"""ctypes binding of libamdahop.so (C ABI declared in include/amdahop.h).

The library is built in-tree (``make -C halomaker-python_amd/csrc``) and loaded from
this directory.  There is no fallback: if the library is missing the import fails
loudly, so a run can never silently compute on the CPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AMDAHOP_LIB selects another build of the same library (e.g. one built with different
# compile-time options); by default the in-tree build is loaded
LIB_PATH = os.environ.get("AMDAHOP_LIB") or os.path.join(HERE, "libamdahop.so")

_i64, _i32, _dbl, _vp, _f32 = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p, ctypes.c_float


class HotParams(ctypes.Structure):
    _fields_ = [("nvoisins", _i32), ("nhop", _i32), ("rho_threshold", _dbl),
                ("xlong", _dbl), ("ylong", _dbl), ("zlong", _dbl),
                ("mixed", _i32), ("bucket", _i32), ("saddle", _i32)]


class HaloParams(ctypes.Structure):
    _fields_ = [("lbox", _dbl), ("hub_pt", _dbl), ("gravconst", _dbl), ("vir_overdens", _dbl),
                ("rho_mean", _dbl), ("cdm", _i32), ("nshell", _i32), ("direct_pe_nmax", _i32),
                ("tile", _i32), ("workgroup", _i32), ("sc", _i32), ("dpmmc", _i32),
                ("dcell_min", _dbl), ("eps_sc", _dbl), ("light", _i32),
                ("virial_pe_profile", _i32), ("virial_pe_norm_policy", _i32),
                ("virial_shell_self_mode", _i32), ("ang_mom_nshells", _i32),
                ("d_ang_mom", ctypes.c_void_p), ("d_ang_axes", ctypes.c_void_p)]


class HaloInfo(ctypes.Structure):
    _fields_ = [("tile", _i32), ("workgroup", _i32), ("trials", _i32), ("ms", _f32),
                ("ms_direct_pe", _f32), ("ms_virial", _f32)]


class RunInfo(ctypes.Structure):
    _fields_ = [("nactive", _i64), ("ngroups", _i32), ("nfallback", _i64), ("nsaddle", _i64),
                ("t_sort", _f32), ("t_tree", _f32), ("t_knn", _f32), ("t_fallback", _f32),
                ("t_hop", _f32), ("t_saddle", _f32), ("t_total", _f32),
                ("knn_nodes", _i64), ("knn_buckets", _i64), ("knn_candidates", _i64),
                ("knn_insert_steps", _i64), ("knn_inserts", _i64), ("knn_waves", _i64),
                ("knn_log_max", _i64), ("knn_log_overflow", _i64), ("knn_misorder", _i64)]


# name -> (restype, argtypes); must list every function declared in include/amdahop.h
SIGNATURES = {
    "ahop_sync_from_change_pos": (ctypes.c_int, [_i64, _i32, _dbl, _dbl, _dbl, _dbl]),
    "ahop_sync_from_init_adaptahop": (ctypes.c_int, [_i64, _i32, _i32] + [_dbl] * 18 + [_vp]),
    "ahop_sync_others": (ctypes.c_int, [_i32, _i32, _i32, _i32, _dbl, _dbl, _dbl, _i32, _i32,
                                        _dbl, _dbl, ctypes.c_char_p, _i32, _i32, _i32]),
    "ahop_compute_adaptahop": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64]),
    "ahop_close": (ctypes.c_int, []),
    "ahop_liste_parts": (_vp, [ctypes.POINTER(_i64)]),
    "ahop_density": (_vp, [ctypes.POINTER(_i64)]),
    "ahop_nnodes": (_i32, []),
    "ahop_real_table": (_vp, [ctypes.POINTER(_i64)]),
    "ahop_integer_table": (_vp, [ctypes.POINTER(_i64)]),
    "ahop_last_error": (ctypes.c_char_p, []),
    "ahop_set_precision": (ctypes.c_int, [_i32]),
    "ahop_set_bucket": (ctypes.c_int, [_i32]),
    "ahop_ngroups": (_i32, []),
    "ahop_get_groups": (ctypes.c_int, [_vp, _vp, _i64]),
    "ahop_nsaddle": (_i64, []),
    "ahop_get_saddle": (ctypes.c_int, [_vp, _vp, _vp]),
    "ahop_run_device": (ctypes.c_int, [_vp, _vp, _i64, ctypes.POINTER(HotParams), _vp, _vp,
                                       ctypes.POINTER(RunInfo)]),
    "ahop_device_saddle": (ctypes.c_int, [_vp, _vp, _vp]),
    "ahop_device_hop": (ctypes.c_int, [_vp, _i64]),
    "ahop_tree_stats": (ctypes.c_int, [_vp, _vp]),
    "ahop_knn_device": (ctypes.c_int, [_vp, _vp, _i64, ctypes.POINTER(HotParams), _vp, _vp, _vp,
                                       ctypes.POINTER(RunInfo)]),
    "ahop_density_adjoint_device": (ctypes.c_int, [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _dbl,
                                                   _dbl, _dbl, _vp, _vp, ctypes.POINTER(_f32)]),
    "ahop_density_adjoint_device_noalloc": (ctypes.c_int, [_vp, _vp, _i64, _i32, _vp, _vp, _vp,
                                                           _dbl, _dbl, _dbl, _vp, _vp]),
    "ahop_ramses_count": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _i32, _i32, _i32, _i32,
                                          _i32, _vp]),
    "ahop_ramses_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _i32, _i32, _i32, _i32,
                                         _i32, _i64, _i64, _i64, _dbl, _dbl, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp, _vp, _vp]),
    "ahop_ramses_descriptor": (ctypes.c_int, [ctypes.c_char_p, _vp, _vp, _vp]),
    "ahop_halo_stellar_device": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                 _i32, _vp, ctypes.POINTER(HaloParams), _dbl, _vp,
                                                 _vp, ctypes.POINTER(_f32)]),
    "ahop_halo_profile_radii_device": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp,
                                                       _vp, _vp, _vp, _vp, ctypes.POINTER(_f32)]),
    "ahop_halo_profile_fits_device": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                      _vp, _vp, ctypes.POINTER(_f32)]),
    "ahop_group_members_device": (ctypes.c_int, [_vp, _i64, _vp, _i32, _i32, _vp, _vp, _vp,
                                                 ctypes.POINTER(_f32)]),
    "ahop_halo_props_device": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _vp,
                                              ctypes.POINTER(HaloParams), _vp,
                                              ctypes.POINTER(HaloInfo)]),
    "ahop_create_group_tree": (ctypes.c_int, [_i64, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp,
                                              _dbl, _dbl, _dbl, _dbl, _i32, _dbl, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """Load libamdahop.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libamdahop.so not built at {LIB_PATH}: run "
                              "`make -C halomaker-python_amd/csrc` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Turn a nonzero ABI status into RuntimeError carrying ahop_last_error()."""
    if status != 0:
        msg = lib().ahop_last_error()
        raise RuntimeError((msg or b"amdahop error").decode())


def ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)
