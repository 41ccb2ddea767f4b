"""NumPy/SciPy restatement of the extended profiles -- TEST INFRASTRUCTURE ONLY.

Follows compute_extended_profiles_1e (src/compute_halo_props.py:1819-1926) for one halo
whose members are all dark matter, given its final centre and rvir: the peak circular
velocity (vmaxcir, rmaxcir), the NFW concentration from a weighted fit of the binned
log-density (cNFW, cNFWerr; scipy.optimize.curve_fit with bounds, as the reference calls it)
and the inner log-slope from a line fit to per-member shell densities (inslope, inslopeerr).

Pinning (tests/test_halo_profiles_oracle.py): a uniform sphere and a point-mass-dominated
halo for the circular velocity, a noise-free NFW profile built bin by bin (the fit must
return the generating rs), a noise-free power law for the inner slope, and the early returns.
"""
from __future__ import annotations

import warnings

import numpy as np

PROFILE_FIELDS = ["vmaxcir", "rmaxcir", "cNFW", "cNFWerr", "inslope", "inslopeerr"]


def _wrap(d, L):
    d = np.array(d, dtype=np.float64, copy=True)
    d[d > L / 2] -= L
    d[d < -L / 2] += L
    return d


def log_nfw(r, rs, logrho0):
    """The fitted model (:1864-1866)."""
    x = r / rs
    return logrho0 - np.log10(x) - 2.0 * np.log10(1.0 + x)


def extended_profiles(pos, mass, centre, rvir, L, gravconst):
    """compute_extended_profiles_1e :1819-1926 for an all-dark-matter halo -> dict of
    PROFILE_FIELDS (NaN where the reference leaves or sets NaN)."""
    from scipy.optimize import curve_fit

    out = dict.fromkeys(PROFILE_FIELDS, np.nan)
    drxs = _wrap(pos[:, 0] - centre[0], L)
    drys = _wrap(pos[:, 1] - centre[1], L)
    drzs = _wrap(pos[:, 2] - centre[2], L)
    dist = np.sqrt(drxs ** 2 + drys ** 2 + drzs ** 2)
    nonzero = dist > 0
    if np.sum(nonzero) > 0:                                        # :1832-1840
        order = np.argsort(dist[nonzero])
        sdist = dist[nonzero][order]
        smass = mass[nonzero][order]
        vcir = np.sqrt(gravconst * np.cumsum(smass) / sdist)
        imax = np.argmax(vcir)
        out["vmaxcir"] = vcir[imax]
        out["rmaxcir"] = sdist[imax]
    dmmask = nonzero                                               # every member is DM
    if np.sum(dmmask) < 8 or not (rvir > 0):                       # :1842-1846
        return out
    dmdist = dist[dmmask]
    dmmass = mass[dmmask]
    order = np.argsort(dmdist)
    dmdist = dmdist[order]
    dmmass = dmmass[order]
    inside = dmdist <= rvir
    if np.sum(inside) < 8:
        return out
    dmdist = dmdist[inside]
    dmmass = dmmass[inside]
    try:                                                           # :1861-1891
        nbins = max(8, min(80, int(np.sqrt(len(dmdist)))))
        bins = np.logspace(np.log10(dmdist[0]), np.log10(dmdist[-1]), nbins + 1)
        shell_mass, _ = np.histogram(dmdist, bins=bins, weights=dmmass)
        shell_num, _ = np.histogram(dmdist, bins=bins)
        shell_vol = 4.0 / 3.0 * np.pi * (bins[1:] ** 3 - bins[:-1] ** 3)
        radius = 0.5 * (bins[1:] + bins[:-1])
        valid = (shell_num > 0) & (shell_mass > 0) & (shell_vol > 0)
        if np.sum(valid) >= 4:
            rho = shell_mass[valid] / shell_vol[valid]
            radius = radius[valid]
            out["_R0"] = radius[0]                                 # for the tests' bars
            sigma = 1.0 / np.sqrt(shell_num[valid])
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                popt, pcov = curve_fit(log_nfw, radius, np.log10(rho),
                                       p0=[max(rvir / 10.0, radius[0]), np.log10(rho[0])],
                                       sigma=sigma, absolute_sigma=False,
                                       bounds=([0.0, -np.inf], [rvir, np.inf]))
            rs = popt[0]
            rs_err = np.sqrt(pcov[0, 0]) if pcov.size else np.nan
            out["cNFW"] = rvir / rs
            out["cNFWerr"] = out["cNFW"] * np.sqrt((rs_err / rs) ** 2) if rs > 0 else np.nan
    except Exception:
        out["cNFW"] = np.nan
        out["cNFWerr"] = np.nan
        out["_failed"] = True
    try:                                                           # :1893-1926
        cnfw = out["cNFW"]
        rs = rvir / cnfw if np.isfinite(cnfw) and cnfw > 0 else rvir
        rcut = min(rs, rvir)
        mask = dmdist < rcut
        if np.sum(mask) >= 4:
            r_in = dmdist[mask]
            m_in = dmmass[mask]
            xx = np.log10(r_in / rvir)
            shell_vol = 4.0 / 3.0 * np.pi * r_in ** 3
            shell_vol = np.diff(np.insert(shell_vol, 0, 0.0))
            rel_width = np.empty_like(r_in)
            rel_width[0] = 1.0
            rel_width[1:] = (r_in[1:] - r_in[:-1]) / np.maximum(r_in[1:], np.finfo(np.float64).tiny)
            valid = (shell_vol > 0) & (m_in > 0) & (rel_width > 1.0e-9)
            xx = xx[valid]
            yy = np.log10(m_in[valid] / shell_vol[valid])
            if len(xx) >= 4:
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore")
                    popt, pcov = curve_fit(lambda x, a, b: a * x + b, xx, yy)
                out["inslope"] = popt[0]
                out["inslopeerr"] = np.sqrt(pcov[0, 0]) if pcov.size else np.nan
    except Exception:
        out["inslope"] = np.nan
        out["inslopeerr"] = np.nan
    return out


def all_halos(offsets, pos, mass, centres, rvirs, L=1.0, gravconst=430.1, extra=False):
    """One row of PROFILE_FIELDS per halo; with ``extra`` also, per halo, the innermost fitted
    bin radius (NaN where no fit was attempted) and whether curve_fit raised."""
    rows, r0, failed = [], [], []
    for h in range(len(offsets) - 1):
        s, e = int(offsets[h]), int(offsets[h + 1])
        o = extended_profiles(pos[s:e], mass[s:e], centres[h], rvirs[h], L, gravconst)
        rows.append([o[k] for k in PROFILE_FIELDS])
        r0.append(o.get("_R0", np.nan))
        failed.append(o.get("_failed", False))
    if extra:
        return np.array(rows), np.array(r0), np.array(failed)
    return np.array(rows)
