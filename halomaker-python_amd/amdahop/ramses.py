"""RAMSES particle files read natively: the particle part of read_ramses_new_101
(src/input_output.py:763-990) with its per-file reader _read_ramses_new_1010 (:624-760), as
one counting sweep and one reading sweep over the cpu files in C++ (ahop_ramses_count,
ahop_ramses_read), one OpenMP thread per file instead of a process pool."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import check, lib, ptr

FAMILIES = {"all": 0, "dm": 1, "star": 2}
VERSIONS = {"Ra3": 3, "Ra4": 4}
# the elements a Ra4 field descriptor may list as chem_<element> (src/chem_species.py:1)
CHEM_ELEMENTS = ("H", "O", "Fe", "Mg", "C", "N", "Si", "S", "D")


def descriptor(repository):
    """What a Ra4 output's part_file_descriptor.txt provides for stars: (present, has initial
    masses, tuple of the chem_<element> elements it lists).  Raises the reference's errors
    for an initial_mass field without or before metallicity."""
    present, has_m0, mask = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    check(lib().ahop_ramses_descriptor(repository.encode(), ctypes.byref(present),
                                       ctypes.byref(has_m0), ctypes.byref(mask)))
    chem = tuple(e for i, e in enumerate(CHEM_ELEMENTS) if mask.value >> i & 1)
    return bool(present.value), bool(has_m0.value), chem


def _nchar(repository):
    ipos = repository.find("output_")             # :785-786
    if ipos < 0:
        raise ValueError("repository must contain 'output_NNNNN'")
    return repository[ipos + 7:ipos + 12]


def _ncpu(repository, nchar):
    path = os.path.join(repository, f"part_{nchar}.out00001")
    with open(path, "rb") as f:
        head = f.read(8)
    if len(head) < 8:
        raise RuntimeError(f"{path}: too short")
    return int(np.frombuffer(head[4:8], dtype=np.int32)[0])


def _info_scales(repository, nchar):
    path = os.path.join(repository, f"info_{nchar}.txt")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if "=" in line:
                    k, v = (x.strip() for x in line.split("=", 1))
                    if k in ("scale_l", "scale_t"):
                        out[k] = float(v)
    return out


def read_particles(repository, rver="Ra3", family="all", stellar=True, scale_l=None,
                   scale_t=None, nthreads=0):
    """Selected particles of a RAMSES output directory in dense-id order: dict with ``pos``
    (N, 3) shifted by -0.5, ``vel`` (N, 3) in km/s, ``mass`` (N,), ``source_pid`` (N,) int64
    (families 'dm' / 'star'), ``birth`` / ``metal`` (nstar,) for the stellar payload (raw birth
    epochs: the age is snapshot_age - epoch_to_age(birth) with the caller's cosmology), the
    counts ``ndm``, ``nstar``, ``npart``, and for a Ra4 output whose field descriptor lists
    them, ``m0`` (nstar,) initial masses and ``chem`` (nstar, 9) abundances in CHEM_ELEMENTS
    order (NaN where the files give none, as the reference allocates them), else None.
    scale_l / scale_t default to the output's info file."""
    if family not in FAMILIES or rver not in VERSIONS:
        raise ValueError("family must be all / dm / star and rver Ra3 / Ra4")
    nchar = _nchar(repository)
    ncpu = _ncpu(repository, nchar)
    sc = _info_scales(repository, nchar)
    scale_l = sc.get("scale_l", 1.0) if scale_l is None else scale_l
    scale_t = sc.get("scale_t", 1.0) if scale_t is None else scale_t
    rep, nch = repository.encode(), nchar.encode()
    fam, rv = FAMILIES[family], VERSIONS[rver]
    counts = np.zeros((ncpu, 3), dtype=np.int64)
    check(lib().ahop_ramses_count(rep, nch, ncpu, rv, fam, 1 if stellar else 0, int(nthreads),
                                  ptr(counts)))
    ndm, nstar_masked, nselected = (int(x) for x in counts.sum(axis=0))
    expected = {"all": ndm + nstar_masked, "dm": ndm, "star": nstar_masked}[family]
    if nselected != expected:                      # :896-904
        raise RuntimeError(f"Selected-count mismatch for family={family}: counted={nselected}, "
                           f"expected={expected}")
    ndm_out = ndm if family in ("all", "dm") else 0
    nstar = nstar_masked if family in ("all", "star") else 0
    if nselected == 0:
        raise RuntimeError(f"No particles selected for family={family}")
    pos = np.zeros((nselected, 3))
    vel = np.zeros((nselected, 3))
    mass = np.zeros(nselected)
    spid = np.full(nselected, -1, dtype=np.int64) if family != "all" else None
    birth = metal = m0 = chem = None
    if stellar and nstar > 0:
        birth = np.full(nstar, np.nan)
        metal = np.full(nstar, np.nan)
        if rver == "Ra4":
            _, has_m0, elems = descriptor(repository)
            m0 = np.full(nstar, np.nan) if has_m0 else None
            chem = np.full((nstar, len(CHEM_ELEMENTS)), np.nan) if elems else None

    def p(a):
        return ptr(a) if a is not None else None

    check(lib().ahop_ramses_read(
        rep, nch, ncpu, rv, fam, 1 if stellar else 0, int(nthreads), ndm_out, nselected, nstar,
        float(scale_l), float(scale_t), ptr(pos), ptr(vel), ptr(mass), p(spid), p(birth),
        p(metal), p(m0), p(chem), None))
    return dict(pos=pos, vel=vel, mass=mass, source_pid=spid, birth=birth, metal=metal,
                m0=m0, chem=chem, ndm=ndm_out, nstar=nstar, npart=nselected)


def write_descriptor(repository, stellar_fields):
    """Write a Ra4 part_file_descriptor.txt in RAMSES's layout: the fixed fields up to the tag,
    then ``stellar_fields`` (names, in record order; each a float64 record per particle)."""
    fixed = [("position_x", "d"), ("position_y", "d"), ("position_z", "d"),
             ("velocity_x", "d"), ("velocity_y", "d"), ("velocity_z", "d"), ("mass", "d"),
             ("identity", "i"), ("levelp", "i"), ("family", "b"), ("tag", "b")]
    rows = fixed + [(f, "d") for f in stellar_fields]
    with open(os.path.join(repository, "part_file_descriptor.txt"), "w") as f:
        f.write("# version:  1\n# ivar, variable_name, variable_type\n")
        for i, (name, t) in enumerate(rows, start=1):
            f.write(f"{i:3d}, {name}, {t}\n")


def write_part_file(path, ncpu, pos, vel, mass, ids, rver="Ra3", birth=None, metal=None,
                    family=None, stellar_records=None):
    """Write one RAMSES part file in the record layout _read_ramses_new_1010 reads (header
    records, positions, velocities, masses, ids, levels, the Ra4 family and tag bytes, then
    the stellar records when any birth epoch is non-zero: birth epochs and metallicities, or
    with ``stellar_records`` that list of per-particle arrays in the descriptor's order).
    For tests and benchmarks: positions in [0, 1), ids int32, family int8 (Ra4)."""
    from scipy.io import FortranFile
    n = len(mass)
    nstar = 0 if birth is None else int(np.count_nonzero(birth))
    with FortranFile(path, "w") as f:
        f.write_record(np.array([ncpu], np.int32))
        f.write_record(np.array([3], np.int32))
        f.write_record(np.array([n], np.int32))
        f.write_record(np.arange(4, dtype=np.int32))                 # localseed
        f.write_record(np.array([nstar], np.int32))
        f.write_record(np.array([0.0]))                              # mstar_tot
        f.write_record(np.array([0.0]))                              # mstar_lost
        f.write_record(np.array([0], np.int32))                      # nsink
        for d in range(3):
            f.write_record(np.ascontiguousarray(pos[:, d], dtype=np.float64))
        for d in range(3):
            f.write_record(np.ascontiguousarray(vel[:, d], dtype=np.float64))
        f.write_record(np.ascontiguousarray(mass, dtype=np.float64))
        f.write_record(np.asarray(ids).astype(np.int32))
        f.write_record(np.full(n, 7, np.int32))                      # level
        if rver == "Ra4":
            f.write_record(np.asarray(family).astype(np.int8))
            f.write_record(np.zeros(n, np.int8))                     # tag
        if nstar > 0:
            recs = stellar_records if stellar_records is not None else [birth, metal]
            for r in recs:
                f.write_record(np.ascontiguousarray(r, dtype=np.float64))
