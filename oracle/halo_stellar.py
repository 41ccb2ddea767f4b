"""NumPy restatement of the per-halo stellar properties -- TEST INFRASTRUCTURE ONLY.

Follows src/compute_halo_props.py under mode 'full' (stellar payload on) for one halo given
its members, a star flag per member (the reference's indexps >= ndm), star ages (Gyr) and
metallicities, and the halo's total mean velocity:

  det_mass_17 :765-780        ndm, nstar, mdm, m*, mcontam
  det_center_18 :963-966      sigma_dm, sigma*
  refine_centers_1b :1569-1593  stellar centre px*, py*, pz* (densest star, then DPMMC / SC
                              over the stars)
  compute_ang_mom_19 :806-822 Lx*, Ly*, Lz*
  r_halos_1a :842-849         r*
  compute_stellar_1d :1635-1751  r50, r90, age, metal, SFR*, and vrot / sig3d / sigcyl over
                              all stars, inside r50 and inside r90 (_stellar_kinematics_1d
                              :1604-1633), and the optional chemistry (:1671-1679).

Pinning (tests/test_halo_stellar_oracle.py): equal-mass stars at known radii (r50/r90 by
counting), weighted averages and SFR sums by hand, a cold rotating ring (vrot = v, sigcyl =
0), a cloud with no rotation, the early returns and defaults, chemistry means by hand.
"""
from __future__ import annotations

import numpy as np

from .halo_props import center_multiscale, center_sphere, radius_from_center, wrap

STELLAR_FIELDS = ["ndm", "nstar", "mdm", "m*", "mcontam", "sigma_dm", "sigma*",
                  "px*", "py*", "pz*", "Lx*", "Ly*", "Lz*", "r*", "r50", "r90", "age", "metal",
                  "SFR", "SFR_r50", "SFR_r90", "SFR10", "SFR10_r50", "SFR10_r90",
                  "vrot", "sig3d", "sigcyl", "vrot_r50", "sig3d_r50", "sigcyl_r50",
                  "vrot_r90", "sig3d_r90", "sigcyl_r90"]

# clear_halo (halo_defs.py:352-380) values of the fields that stay untouched without stars
_DEFAULTS = {"SFR": 0.0, "SFR_r50": 0.0, "SFR_r90": 0.0, "SFR10": 0.0, "SFR10_r50": 0.0,
             "SFR10_r90": 0.0, "sigma_dm": 0.0, "sigma*": 0.0}


def stellar_kinematics(mass, dvx, dvy, dvz, vphi, vrad, vzyl, select):
    """_stellar_kinematics_1d :1604-1633 -> (vrot, sig3d, sigcyl)."""
    select = select & np.isfinite(mass) & (mass > 0)
    if np.count_nonzero(select) < 2:
        return np.nan, np.nan, np.nan
    weight = mass[select]
    weight_sum = np.sum(weight)
    if not np.isfinite(weight_sum) or weight_sum <= 0:
        return np.nan, np.nan, np.nan
    sig3d2 = np.average(dvx[select] ** 2 + dvy[select] ** 2 + dvz[select] ** 2, weights=weight)
    cyl = select & np.isfinite(vphi) & np.isfinite(vrad) & np.isfinite(vzyl)
    if np.count_nonzero(cyl) < 2:
        return np.nan, np.sqrt(max(sig3d2, 0.0)), np.nan
    w = mass[cyl]
    vrot = np.average(vphi[cyl], weights=w)
    mvrad = np.average(vrad[cyl], weights=w)
    mvzyl = np.average(vzyl[cyl], weights=w)
    sigcyl2 = np.average((vrad[cyl] - mvrad) ** 2 + (vphi[cyl] - vrot) ** 2 +
                         (vzyl[cyl] - mvzyl) ** 2, weights=w)
    return vrot, np.sqrt(max(sig3d2, 0.0)), np.sqrt(max(sigcyl2, 0.0))


def stellar_props(pos, vel, mass, density, is_star, age, metal, vbar, *, lbox=1.0,
                  massp=0.0, sc=False, dpmmc=False, dcell_min=0.0, eps_sc=0.0, chem=None):
    """All STELLAR_FIELDS of one halo (dict).  ``age`` / ``metal`` are per member (read for
    stars only); ``vbar`` the halo's mean velocity (det_center_18); ``chem`` (n, nchem)
    per-member abundances, whose stellar means go to out["chem"] (:1671-1679)."""
    L = lbox
    out = dict.fromkeys(STELLAR_FIELDS, np.nan)
    out.update(_DEFAULTS)
    if chem is not None:
        out["chem"] = np.full(chem.shape[1], np.nan)
    star = np.asarray(is_star, dtype=bool)
    dm = ~star
    idm = mass[dm]                                                 # det_mass_17 :765-780
    out["mdm"] = np.sum(idm)
    out["mcontam"] = np.sum(idm[idm > massp * 1.00001])
    out["m*"] = np.sum(mass[star])
    out["ndm"] = len(idm)
    out["nstar"] = int(star.sum())
    dv2 = mass * ((vel[:, 0] - vbar[0]) ** 2 + (vel[:, 1] - vbar[1]) ** 2 +
                  (vel[:, 2] - vbar[2]) ** 2)                      # det_center_18 :957-966
    out["sigma_dm"] = np.sqrt(np.sum(dv2[dm]) / out["mdm"]) if out["mdm"] > 0 else 0
    out["sigma*"] = np.sqrt(np.sum(dv2[star]) / out["m*"]) if out["m*"] > 0 else 0
    if not star.any():
        out["Lx*"] = out["Ly*"] = out["Lz*"] = np.nan
        return out
    spos, svel, smass = pos[star], vel[star], mass[star]
    sdens = density[star]
    c = np.array(spos[int(np.argmax(sdens))], dtype=np.float64)   # refine_centers_1b :1576-1580
    r0 = radius_from_center(spos, c, L)
    if np.isfinite(r0) and r0 > 0:
        if dpmmc:
            c = center_multiscale(spos, smass, sdens, c, r0, 9, dcell_min)
            r0 = radius_from_center(spos, c, L)
        if sc:
            c = center_sphere(spos, smass, c, r0, L, dcell_min, eps_sc)
    out["px*"], out["py*"], out["pz*"] = c
    sv = [np.sum(smass * svel[:, d]) / out["m*"] for d in range(3)]   # compute_ang_mom_19
    sdr = [wrap(spos[:, d] - c[d], L) for d in range(3)]
    spv = [smass * (svel[:, d] - sv[d]) for d in range(3)]
    out["Lx*"] = np.sum(sdr[1] * spv[2] - sdr[2] * spv[1])
    out["Ly*"] = np.sum(sdr[2] * spv[0] - sdr[0] * spv[2])
    out["Lz*"] = np.sum(sdr[0] * spv[1] - sdr[1] * spv[0])
    out["r*"] = np.sqrt(np.max(sdr[0] ** 2 + sdr[1] ** 2 + sdr[2] ** 2))   # r_halos_1a
    # compute_stellar_1d :1649-1751
    sage, smetal = age[star], metal[star]
    drx, dry, drz = sdr
    dr2 = drx ** 2 + dry ** 2 + drz ** 2
    order = np.argsort(dr2)
    cum = np.cumsum(smass[order])
    total = cum[-1]
    out["r50"] = np.sqrt(dr2[order[np.searchsorted(cum, 0.5 * total)]])
    out["r90"] = np.sqrt(dr2[order[np.searchsorted(cum, 0.9 * total)]])
    out["age"] = np.average(sage, weights=smass)
    out["metal"] = np.average(smetal, weights=smass)
    if chem is not None:                                           # :1671-1679
        schem = chem[star]
        valid_w = np.isfinite(smass) & (smass > 0)
        for e in range(chem.shape[1]):
            values = schem[:, e]
            valid = valid_w & np.isfinite(values)
            if np.any(valid):
                out["chem"][e] = np.average(values[valid], weights=smass[valid])
    in50 = dr2 < out["r50"] ** 2
    in90 = dr2 < out["r90"] ** 2
    y100 = sage < 0.1
    out["SFR"] = np.sum(smass[y100]) * 1.0e3
    out["SFR_r50"] = np.sum(smass[y100 & in50]) * 1.0e3
    out["SFR_r90"] = np.sum(smass[y100 & in90]) * 1.0e3
    y10 = sage < 0.01
    out["SFR10"] = np.sum(smass[y10]) * 1.0e4
    out["SFR10_r50"] = np.sum(smass[y10 & in50]) * 1.0e4
    out["SFR10_r90"] = np.sum(smass[y10 & in90]) * 1.0e4
    stellar_mass = np.sum(smass)
    am = np.array([out["Lx*"], out["Ly*"], out["Lz*"]], dtype=np.float64)
    amn = np.linalg.norm(am)
    if len(smass) >= 2 and np.isfinite(stellar_mass) and stellar_mass > 0:
        bulk = np.average(svel, axis=0, weights=smass)
        dvx, dvy, dvz = (svel[:, d] - bulk[d] for d in range(3))
        n = len(smass)
        vphi, vrad, vzyl = np.full(n, np.nan), np.full(n, np.nan), np.full(n, np.nan)
        if np.isfinite(amn) and amn > 0:
            ax = am / amn
            zpos = drx * ax[0] + dry * ax[1] + drz * ax[2]
            rpx, rpy, rpz = drx - zpos * ax[0], dry - zpos * ax[1], drz - zpos * ax[2]
            rperp = np.sqrt(rpx ** 2 + rpy ** 2 + rpz ** 2)
            ok = rperp > np.finfo(np.float64).eps
            inv = np.zeros_like(rperp)
            inv[ok] = 1.0 / rperp[ok]
            erx, ery, erz = rpx * inv, rpy * inv, rpz * inv
            epx = ax[1] * erz - ax[2] * ery
            epy = ax[2] * erx - ax[0] * erz
            epz = ax[0] * ery - ax[1] * erx
            vrad[ok] = dvx[ok] * erx[ok] + dvy[ok] * ery[ok] + dvz[ok] * erz[ok]
            vphi[ok] = dvx[ok] * epx[ok] + dvy[ok] * epy[ok] + dvz[ok] * epz[ok]
            vzyl[ok] = dvx[ok] * ax[0] + dvy[ok] * ax[1] + dvz[ok] * ax[2]
        for suffix, sel in (("", np.ones(n, dtype=bool)), ("_r50", in50), ("_r90", in90)):
            v, s3, sc_ = stellar_kinematics(smass, dvx, dvy, dvz, vphi, vrad, vzyl, sel)
            out["vrot" + suffix], out["sig3d" + suffix], out["sigcyl" + suffix] = v, s3, sc_
    return out


def all_halos(offsets, pos, vel, mass, density, is_star, age, metal, vbars, chem=None, **kw):
    """One row of STELLAR_FIELDS per halo, followed by the chemistry columns when given."""
    rows = []
    for h in range(len(offsets) - 1):
        s, e = int(offsets[h]), int(offsets[h + 1])
        o = stellar_props(pos[s:e], vel[s:e], mass[s:e], density[s:e], is_star[s:e], age[s:e],
                          metal[s:e], vbars[h], chem=None if chem is None else chem[s:e], **kw)
        rows.append([o[k] for k in STELLAR_FIELDS] + ([] if chem is None else list(o["chem"])))
    return np.array(rows, dtype=np.float64)
