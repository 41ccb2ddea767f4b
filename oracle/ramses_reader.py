"""NumPy/SciPy restatement of the RAMSES particle reader -- TEST INFRASTRUCTURE ONLY.

Follows _read_ramses_new_1010 (src/input_output.py:624-760) and the counting / reading
sweeps of read_ramses_new_101 (:842-990) for the particle arrays, reading the Fortran records
with scipy.io.FortranFile as the reference does (Ra3 and Ra4 layouts; for Ra4 the stellar
records in the order of the output's field descriptor, _parse_ra4_part_descriptor,
_ra4_stellar_descriptor and _read_ra4_stellar_record_block, :530-620; birth epochs returned
raw).  Pinned by tests/test_ramses_reader.py against files written record by record,
including hand-made ids, families and descriptor orders.
"""
from __future__ import annotations

import os

import numpy as np
from scipy.io import FortranFile

CHEM_ELEMENTS = ("H", "O", "Fe", "Mg", "C", "N", "Si", "S", "D")


def ra4_descriptor(repository):
    """The descriptor's field names (None without the file) and what it provides for stars:
    (fields, index of initial_mass past metallicity or None, {element: field index})."""
    path = os.path.join(repository, "part_file_descriptor.txt")
    if not os.path.exists(path):
        return None, None, {}
    fields = []
    with open(path) as stream:
        for line in stream:
            if line.lstrip().startswith("#"):
                continue
            cols = [c.strip() for c in line.split(",")]
            if len(cols) >= 2:
                fields.append(cols[1])
    m0_skip = None
    if "initial_mass" in fields and "metallicity" not in fields:
        raise RuntimeError(f"Ra4 descriptor has initial_mass but no metallicity: {path}")
    if "initial_mass" in fields:
        im, imet = fields.index("initial_mass"), fields.index("metallicity")
        if im <= imet:
            raise RuntimeError(f"Unsupported Ra4 initial_mass record order: {path}")
        m0_skip = im - imet - 1
    chem = {e: fields.index(f"chem_{e}") for e in CHEM_ELEMENTS if f"chem_{e}" in fields}
    return fields, m0_skip, chem


def _ra4_block(f, fields, chem):
    """Birth, metallicity, initial mass and chem records after the tag, where the descriptor
    places them; other fields skipped."""
    out = {"birth_time": None, "metallicity": None, "initial_mass": None}
    chem_vals = {}
    if fields is None:
        out["birth_time"] = f.read_reals()
        out["metallicity"] = f.read_reals()
        return out, chem_vals
    wanted = {"birth_time", "metallicity", "initial_mass"} | {f"chem_{e}" for e in chem}
    where = [fields.index(w) for w in wanted if w in fields]
    if not where:
        return out, chem_vals
    first = fields.index("tag") + 1 if "tag" in fields else min(where)
    for i in range(first, max(where) + 1):
        name = fields[i]
        if name not in wanted:
            _skip(f, 1)
            continue
        v = f.read_reals()
        if name.startswith("chem_"):
            chem_vals[name[5:]] = v
        else:
            out[name] = v
    return out, chem_vals


def _skip(f, n=1):
    for _ in range(n):
        f.read_record(np.uint8)


def read_file(path, rver, family, stellar, count_only, ndm_total=0, nselected=0,
              scale_l=1.0, scale_t=1.0, desc=(None, None, {})):
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
        tmpt = tmpmetal = tmpm0 = None
        tmpchem = {}
        if rver == "Ra4":
            fam = f.read_ints(dtype=np.int8)
            _skip(f, 1)
        if nstar > 0 or nsink > 0:
            keep = stellar and not count_only and family in {"all", "star"}
            if rver == "Ra4" and keep:
                blk, tmpchem = _ra4_block(f, desc[0], desc[2])
                tmpt, tmpmetal, tmpm0 = blk["birth_time"], blk["metallicity"], blk["initial_mass"]
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
    if stellar and family in {"all", "star"} and starmask.any():
        out["star_slots"] = source_ids[starmask] - 1
        out["birth"] = tmpt[starmask]
        out["metal"] = tmpmetal[starmask]
        out["m0"] = tmpm0[starmask] if tmpm0 is not None else None
        out["chem"] = {e: v[starmask] for e, v in tmpchem.items()}
    return counts, out


def read_particles(repository, nchar, ncpu, rver="Ra3", family="all", stellar=True,
                   scale_l=1.0, scale_t=1.0):
    paths = [f"{repository}/part_{nchar}.out{i:05d}" for i in range(1, ncpu + 1)]
    desc = ra4_descriptor(repository) if rver == "Ra4" else (None, None, {})
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
    m0 = np.full(nstar, np.nan) if birth is not None and desc[1] is not None else None
    chem = (np.full((nstar, len(CHEM_ELEMENTS)), np.nan)
            if birth is not None and desc[2] else None)
    for p in paths:
        _, o = read_file(p, rver, family, stellar, False, ndm_out, nsel, scale_l, scale_t, desc)
        pos[o["ind"]] = o["pos"]
        vel[o["ind"]] = o["vel"]
        mass[o["ind"]] = o["mass"]
        if spid is not None:
            spid[o["ind"]] = o["spid"]
        if birth is not None and "star_slots" in o:
            birth[o["star_slots"]] = o["birth"]
            metal[o["star_slots"]] = o["metal"]
            if m0 is not None and o["m0"] is not None:
                m0[o["star_slots"]] = o["m0"]
            for e, v in o["chem"].items():
                chem[o["star_slots"], CHEM_ELEMENTS.index(e)] = v
    return dict(pos=pos, vel=vel, mass=mass, source_pid=spid, birth=birth, metal=metal,
                m0=m0, chem=chem, ndm=ndm_out, nstar=nstar, npart=nsel)
