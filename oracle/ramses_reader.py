"""NumPy/SciPy restatement of the RAMSES particle reader -- TEST INFRASTRUCTURE ONLY.

Follows _read_ramses_new_1010 (src/input_output.py:624-760) and the counting / reading
sweeps of read_ramses_new_101 (:842-990) for the particle arrays, reading the Fortran records
with scipy.io.FortranFile as the reference does (Ra3 and Ra4 layouts without a field
descriptor; birth epochs returned raw).  Pinned by tests/test_ramses_reader.py against files
written record by record, including hand-made ids and families.
"""
from __future__ import annotations

import numpy as np
from scipy.io import FortranFile


def _skip(f, n=1):
    for _ in range(n):
        f.read_record(np.uint8)


def read_file(path, rver, family, stellar, count_only, ndm_total=0, nselected=0,
              scale_l=1.0, scale_t=1.0):
    """One file: (ndm, nstar, nselected) and, reading, the scatter as dict ind -> values."""
    with FortranFile(path, "r") as f:
        f.read_ints()
        ndim, = f.read_ints()
        npart, = f.read_ints()
        _skip(f, 1)
        nstar, = f.read_ints()
        _skip(f, 2)
        nsink, = f.read_ints()
        if count_only:
            _skip(f, 2 * ndim + 1)
        else:
            tmpp = np.empty((npart, 3))
            tmpv = np.empty((npart, 3))
            for d in range(ndim):
                tmpp[:, d] = f.read_reals()
            for d in range(ndim):
                tmpv[:, d] = f.read_reals()
            tmpm = f.read_reals()
        idp = f.read_ints()
        _skip(f, 1)
        tmpt = tmpmetal = None
        if rver == "Ra4":
            fam = f.read_ints(dtype=np.int8)
            _skip(f, 1)
        if nstar > 0 or nsink > 0:
            keep = stellar and not count_only and family in {"all", "star"}
            if rver == "Ra4" and keep:
                tmpt = f.read_reals()
                tmpmetal = f.read_reals()
            elif rver != "Ra4":
                tmpt = f.read_reals()
            if rver != "Ra4" and keep:
                tmpmetal = f.read_reals()
    if rver == "Ra4":
        dmmask = fam == 1
        starmask = fam == 2
    else:
        if tmpt is None:
            tmpt = np.zeros(npart)
        dmmask = (idp > 0) & (tmpt == 0)
        starmask = ((tmpt < 0) & (idp > 0)) | ((tmpt != 0) & (idp < 0))
    mask = {"all": dmmask | starmask, "dm": dmmask, "star": starmask}[family]
    counts = (int(dmmask.sum()), int(starmask.sum()), int(mask.sum()))
    if count_only:
        return counts, None
    source_ids = np.abs(idp)
    local = np.where(starmask, source_ids + ndm_total, source_ids) if family == "all" else source_ids
    ind = local[mask] - 1
    if np.any(ind < 0) or np.any(ind >= nselected) or len(np.unique(ind)) != len(ind):
        raise RuntimeError("Non-dense or duplicate particle IDs")
    out = dict(ind=ind, pos=tmpp[mask] - 0.5, vel=tmpv[mask] * scale_l / scale_t * 1e-5,
               mass=tmpm[mask], spid=source_ids[mask])
    if stellar and family in {"all", "star"} and starmask.any() and tmpt is not None:
        out["star_slots"] = source_ids[starmask] - 1
        out["birth"] = tmpt[starmask]
        out["metal"] = tmpmetal[starmask] if tmpmetal is not None else None
    return counts, out


def read_particles(repository, nchar, ncpu, rver="Ra3", family="all", stellar=True,
                   scale_l=1.0, scale_t=1.0):
    paths = [f"{repository}/part_{nchar}.out{i:05d}" for i in range(1, ncpu + 1)]
    c = np.array([read_file(p, rver, family, stellar, True)[0] for p in paths])
    ndm, nst, nsel = (int(x) for x in c.sum(axis=0))
    ndm_out = ndm if family in ("all", "dm") else 0
    nstar = nst if family in ("all", "star") else 0
    pos = np.zeros((nsel, 3))
    vel = np.zeros((nsel, 3))
    mass = np.zeros(nsel)
    spid = np.full(nsel, -1, dtype=np.int64) if family != "all" else None
    birth = np.full(nstar, np.nan) if stellar and nstar > 0 else None
    metal = np.full(nstar, np.nan) if stellar and nstar > 0 else None
    for p in paths:
        _, o = read_file(p, rver, family, stellar, False, ndm_out, nsel, scale_l, scale_t)
        pos[o["ind"]] = o["pos"]
        vel[o["ind"]] = o["vel"]
        mass[o["ind"]] = o["mass"]
        if spid is not None:
            spid[o["ind"]] = o["spid"]
        if birth is not None and "star_slots" in o:
            birth[o["star_slots"]] = o["birth"]
            if o["metal"] is not None:
                metal[o["star_slots"]] = o["metal"]
    return dict(pos=pos, vel=vel, mass=mass, source_pid=spid, birth=birth, metal=metal,
                ndm=ndm_out, nstar=nstar, npart=nsel)
