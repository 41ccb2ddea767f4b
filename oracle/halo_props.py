"""NumPy restatement of the per-halo property reductions -- TEST INFRASTRUCTURE ONLY.

Follows reference src/compute_halo_props.py per function (cited below) for one halo
given its members (positions, velocities, masses, densities) in physical units, in the
configuration the hot path covers: periodic box (FlagPeriod=1), no stellar payload,
SC/DPMMC off (refine_centers_1b :1555 then leaves the centre unchanged), virial
potential profile 'shell' with self mode 'count' and norm policy
'profile_outer_for_large' (the defaults at :1035-1047).

Output order of :func:`halo_props` (one float64 row per halo) -- see FIELDS.

Pinning (tests/test_halo_oracle.py): closed forms (two-body angular momentum,
rigid rotation, the inertia tensor of a symmetric cube shell, a two-particle direct
potential energy), invariance of L and I under a periodic translation across the box
edge, eigen-decomposition vs scipy.linalg.eigh used as the reference does.
"""
from __future__ import annotations

import numpy as np

FIELDS = ["m", "px", "py", "pz", "vx", "vy", "vz", "sigma", "Lx", "Ly", "Lz", "r",
          "Ixx", "Ixy", "Ixz", "Iyy", "Iyz", "Izz", "sha", "shb", "shc", "ek", "ep", "et",
          "rvir", "mvir", "cvel", "tvir", "spin"]
F = {k: i for i, k in enumerate(FIELDS)}


def _wrap(d, L):
    """correct_for_periodicity_1d :888-897 (strict > L/2 and < -L/2)."""
    d = d.copy()
    d[d > L / 2] -= L
    d[d < -L / 2] += L
    return d


def _wrap_scalar(x, L):
    """correct_for_periodicity :856-885 on one coordinate."""
    if x > L / 2:
        x = x - L
    if x < -L / 2:
        x = x + L
    return x


def _shell_scale(nr):
    """_shell_scale :1061-1066."""
    s = np.arange(nr, dtype=np.float64) / np.float64(nr - 1)
    s[0] = 0.5 / np.float64(nr - 1)
    return s


def halo_props(pos, vel, mass, density=None, *, lbox=1.0, hub_pt=70.3, gravconst=430.1,
               vir_overdens=340.0, rho_mean=1.0, cdm=True, nshell=1000, direct_pe_nmax=1000,
               eig=None):
    """All properties of one halo.  ``eig`` optionally supplies (d, v) to use instead of
    scipy.linalg.eigh (to compare the virial scan on identical axes)."""
    from scipy.linalg import eigh

    out = np.full(len(FIELDS), np.nan)
    L = lbox
    n = mass.shape[0]
    # det_mass_17 :754-780
    m = np.sum(mass)
    out[F["m"]] = m
    # det_center_18 :901-966
    if cdm:
        p0 = pos[0]
        drx = _wrap(pos[:, 0] - p0[0], L)
        dry = _wrap(pos[:, 1] - p0[1], L)
        drz = _wrap(pos[:, 2] - p0[2], L)
        pc = [np.sum(mass * drx) / m + p0[0], np.sum(mass * dry) / m + p0[1],
              np.sum(mass * drz) / m + p0[2]]
        pc = [_wrap_scalar(x, L) for x in pc]
        drx = _wrap(pos[:, 0] - pc[0], L)
        dry = _wrap(pos[:, 1] - pc[1], L)
        drz = _wrap(pos[:, 2] - pc[2], L)
        d2 = drx ** 2 + dry ** 2 + drz ** 2
        imin = int(np.argmin(d2))
        cpos = pos[imin]
    else:
        cpos = pos[int(np.argmax(density))]
    px, py, pz = cpos
    out[F["px"]], out[F["py"]], out[F["pz"]] = px, py, pz
    vx = np.sum(mass * vel[:, 0]) / m
    vy = np.sum(mass * vel[:, 1]) / m
    vz = np.sum(mass * vel[:, 2]) / m
    out[F["vx"]], out[F["vy"]], out[F["vz"]] = vx, vy, vz
    dv2 = mass * ((vel[:, 0] - vx) ** 2 + (vel[:, 1] - vy) ** 2 + (vel[:, 2] - vz) ** 2)
    out[F["sigma"]] = np.sqrt(np.sum(dv2) / m)
    # compute_ang_mom_19 :783-824
    drx = _wrap(pos[:, 0] - px, L)
    dry = _wrap(pos[:, 1] - py, L)
    drz = _wrap(pos[:, 2] - pz, L)
    pvx = mass * (vel[:, 0] - vx)
    pvy = mass * (vel[:, 1] - vy)
    pvz = mass * (vel[:, 2] - vz)
    out[F["Lx"]] = np.sum(dry * pvz - drz * pvy)
    out[F["Ly"]] = np.sum(drz * pvx - drx * pvz)
    out[F["Lz"]] = np.sum(drx * pvy - dry * pvx)
    # r_halos_1a :828-851
    out[F["r"]] = np.sqrt(np.max(drx ** 2 + dry ** 2 + drz ** 2))
    # det_inertial_tensor_1b20 :1932-1951
    I = np.zeros((3, 3))
    I[0, 0] = np.sum(mass * drx * drx)
    I[0, 1] = I[1, 0] = np.sum(mass * drx * dry)
    I[0, 2] = I[2, 0] = np.sum(mass * drx * drz)
    I[1, 1] = np.sum(mass * dry * dry)
    I[1, 2] = I[2, 1] = np.sum(mass * dry * drz)
    I[2, 2] = np.sum(mass * drz * drz)
    out[F["Ixx"]], out[F["Ixy"]], out[F["Ixz"]] = I[0, 0], I[0, 1], I[0, 2]
    out[F["Iyy"]], out[F["Iyz"]], out[F["Izz"]] = I[1, 1], I[1, 2], I[2, 2]
    # det_main_axis_1b2 :1954-1967
    d, v = eigh(I) if eig is None else eig
    d = np.sqrt(d / m)
    sha, shb, shc = d
    out[F["sha"]], out[F["shb"]], out[F["shc"]] = sha, shb, shc
    # det_halo_energies_1b3 :1400-1444 (pair sum below direct_pe_nmax; the RF formula is
    # overwritten by the shell profile for larger halos under the default policy)
    ep = np.nan
    if n < direct_pe_nmax:
        i, j = np.triu_indices(n, 1)
        drs = pos[j] - pos[i]
        drs[drs > L / 2] -= L
        drs[drs < -L / 2] += L
        dists = np.sqrt(np.sum(drs ** 2, axis=1)) / L
        ep = np.sum(-(mass[i] * mass[j]) / dists) * gravconst / L
    vtx = vel[:, 0] - vx + drx * hub_pt
    vty = vel[:, 1] - vy + dry * hub_pt
    vtz = vel[:, 2] - vz + drz * hub_pt
    v2 = vtx ** 2 + vty ** 2 + vtz ** 2
    ek = 0.5 * np.sum(mass * v2)
    # tab_props_inside_1b40 :969-1032
    nr = nshell
    dra = drx * v[0, 0] + dry * v[1, 0] + drz * v[2, 0]
    drb = drx * v[0, 1] + dry * v[1, 1] + drz * v[2, 1]
    drc = drx * v[0, 2] + dry * v[1, 2] + drz * v[2, 2]
    r_ell = np.sqrt((dra / sha) ** 2 + (drb / shb) ** 2 + (drc / shc) ** 2)
    rmax = np.max(r_ell)
    eps = 1e-2
    amax = rmax * sha * (1.0 + eps)
    bmax = rmax * shb * (1.0 + eps)
    cmax = rmax * shc * (1.0 + eps)
    r_ell = np.sqrt((dra / amax) ** 2 + (drb / bmax) ** 2 + (drc / cmax) ** 2)
    i_ell = (r_ell * nr).astype(int)
    if np.any(i_ell >= nr):
        raise ValueError("tab_props_inside: a member falls outside the outer ellipsoid")
    tabm = np.zeros(nr)
    tabk = np.zeros(nr)
    np.add.at(tabm, i_ell, mass)
    np.add.at(tabk, i_ell, 0.5 * mass * v2)
    cnt = np.bincount(i_ell, minlength=nr).astype(np.float64)
    tabm = np.cumsum(tabm)
    tabk = np.cumsum(tabk)
    # potential_profile_1b40 :1069-1123 (shell / count / profile_outer_for_large)
    dm = np.diff(tabm, prepend=0.0)
    mass_before = tabm - dm
    self_factor = np.zeros(nr)
    occ = cnt > 0
    self_factor[occ] = 0.5 * (cnt[occ] - 1.0) / cnt[occ]
    r_eff = _shell_scale(nr) * (amax * bmax * cmax) ** (1.0 / 3.0)
    tabp = -gravconst * np.cumsum((mass_before + self_factor * dm) * dm / r_eff)
    outer = tabp[nr - 1]
    if n < direct_pe_nmax:
        if ep != outer and outer != 0.0:
            tabp = tabp / outer * ep
    else:
        ep = outer
    et = ek + ep
    out[F["ek"]], out[F["ep"]], out[F["et"]] = ek, ep, et
    # det_vir_props_1b4 :1127-1229
    ttab = nr
    mvir = tabm[ttab - 1]
    i1 = 1
    for i1 in range(ttab - 1, 0, -1):
        if tabm[i1 - 1] < tabm[i1]:
            break
    avir = i1 / (ttab - 1) * amax
    bvir = i1 / (ttab - 1) * bmax
    cvir = i1 / (ttab - 1) * cmax
    rvir = (avir * bvir * cvir) ** (1. / 3.)
    virth_old = 1.0
    virth = 100.0
    for i0 in range(ttab - 1, -1, -1):
        if (tabk[i0] + tabp[i0]) >= 0.0:
            continue
        virth = abs((2.0 * tabk[i0] + tabp[i0]) / (tabk[i0] + tabp[i0]))
        if virth < virth_old:
            mvir = tabm[i0]
            avir = min(avir, i0 / (ttab - 1) * amax)
            bvir = min(bvir, i0 / (ttab - 1) * bmax)
            cvir = min(cvir, i0 / (ttab - 1) * cmax)
            rvir = (avir * bvir * cvir) ** (1. / 3.)
            virth_old = virth
            if virth <= 0.20:
                break
    volmin = 4. / 3. * np.pi * (amax / (ttab - 1)) * (bmax / (ttab - 1)) * (cmax / (ttab - 1))
    if (virth > 0.20) or (mvir < vir_overdens * rho_mean * 4. / 3. * np.pi * avir * bvir * cvir):
        ii1 = 1
        for ii1 in range(ttab - 1, 0, -1):
            mv = vir_overdens * rho_mean * volmin * ii1 ** 3
            if (tabm[ii1] >= mv) and (tabm[ii1 - 1] < tabm[ttab - 1]):
                break
        mvir = tabm[ii1]
        avir = ii1 / (ttab - 1) * amax
        bvir = ii1 / (ttab - 1) * bmax
        cvir = ii1 / (ttab - 1) * cmax
        rvir = (avir * bvir * cvir) ** (1. / 3.)
    if mvir > 0.0 and rvir > 0.0:
        out[F["rvir"]] = min(rvir, out[F["r"]])
        out[F["mvir"]] = mvir
        out[F["cvel"]] = np.sqrt(gravconst * mvir / out[F["rvir"]])
        out[F["tvir"]] = 35.9 * gravconst * mvir / out[F["rvir"]]
    # compute_spin_parameter_1c :1599-1602
    hl = np.sqrt(out[F["Lx"]] ** 2 + out[F["Ly"]] ** 2 + out[F["Lz"]] ** 2)
    out[F["spin"]] = hl * np.sqrt(np.abs(et)) / m ** 2.5 / gravconst
    return out


def all_halos(offsets, pos, vel, mass, density=None, **kw):
    rows = []
    for h in range(len(offsets) - 1):
        s, e = int(offsets[h]), int(offsets[h + 1])
        rows.append(halo_props(pos[s:e], vel[s:e], mass[s:e],
                               None if density is None else density[s:e], **kw))
    return np.array(rows)
