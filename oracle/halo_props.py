"""NumPy restatement of the per-halo property reductions -- TEST INFRASTRUCTURE ONLY.

Follows reference src/compute_halo_props.py per function (cited below) for one halo
given its members (positions, velocities, masses, densities) in physical units: no stellar
payload, an AdaptaHOP selection method (not FOF, so refine_centers_1b :1555-1596 recentres
with the DPMMC and SC options when they are on), FlagPeriod 1 (periodic, the default) or 0,
and the virial profile switches of potential_profile_1b40 (:1036-1123): profile 'shell'
(default) or 'baseline', self mode 'count' (default) or 'half_shell', norm policy
'profile_outer_for_large' (default) or 'legacy'.

Each reference function has its own function here so that each can be pinned on its
own; :func:`halo_props` composes them in the reference's order and returns one float64
row per halo in the order of FIELDS.

Pinning (tests/test_halo_oracle.py): closed forms (two-body and rigid-rotation angular
momentum, the inertia tensor of a cube shell, direct potential energies of regular
2-, 3- and 4-member systems, Hubble-flow kinetic energy, the shell profile of a single
shell, the overdensity branch of a static halo), invariance under a periodic translation
across the box edge and rotation covariance of L and I, and eigh ordering; for the
recentring, the cases where it reduces to the centre of det_center_18, a core found
through a decoy, the grid's cell choice, and the reference's indexing of the densest
member (tests/test_halo_oracle.py); for the per-shell angular momentum of ANG_MOM_OF_R,
members placed by hand on the principal axes of a rotated frame and the shells summing to the
halo's mass and angular momentum.
"""
from __future__ import annotations

import numpy as np

FIELDS = ["m", "px", "py", "pz", "vx", "vy", "vz", "sigma", "Lx", "Ly", "Lz", "r",
          "Ixx", "Ixy", "Ixz", "Iyy", "Iyz", "Izz", "sha", "shb", "shc", "ek", "ep", "et",
          "rvir", "mvir", "cvel", "tvir", "spin", "rho_0", "r_c"]
F = {k: i for i, k in enumerate(FIELDS)}


def wrap(d, L):
    """correct_for_periodicity_1d :888-897 (strict > L/2 and < -L/2)."""
    d = np.array(d, dtype=np.float64, copy=True)
    d[d > L / 2] -= L
    d[d < -L / 2] += L
    return d


def wrap_scalar(x, L):
    """correct_for_periodicity :856-885 on one coordinate."""
    if x > L / 2:
        x = x - L
    if x < -L / 2:
        x = x + L
    return x


def det_center(pos, vel, mass, density, L, cdm):
    """det_center_18 :901-966 -> (centre position (3,), mean velocity (3,), sigma)."""
    m = np.sum(mass)
    if cdm:
        p0 = pos[0]
        pc = [wrap_scalar(np.sum(mass * wrap(pos[:, d] - p0[d], L)) / m + p0[d], L)
              for d in range(3)]
        d2 = sum(wrap(pos[:, d] - pc[d], L) ** 2 for d in range(3))
        cpos = pos[int(np.argmin(d2))]
    else:
        cpos = pos[int(np.argmax(density))]
    vbar = np.array([np.sum(mass * vel[:, d]) / m for d in range(3)])
    dv2 = mass * ((vel[:, 0] - vbar[0]) ** 2 + (vel[:, 1] - vbar[1]) ** 2 +
                  (vel[:, 2] - vbar[2]) ** 2)
    return np.array(cpos, dtype=np.float64), vbar, np.sqrt(np.sum(dv2) / m)


def radius_from_center(pos, c, L):
    """_radius_from_center :1543-1548: distance of the farthest member, wrapped."""
    if len(pos) == 0:
        return np.float64(np.nan)
    d = rel(pos, c, L)
    return np.sqrt(np.max(d[0] ** 2 + d[1] ** 2 + d[2] ** 2))


def center_multiscale(pos, mass, density, c, r0, nx, dcell_min, max_levels=256):
    """det_halo_center_multiscale_1b0 :1447-1495, the recursion unrolled.

    Each level assigns member masses to an nx^3 NGP grid over the cube [c - r0, c + r0]
    (positions not wrapped; the index is the float quotient truncated toward zero, so
    quotients in (-1, 0) land in cell 0, :1460-1465), takes the first cell of maximum mass
    in (i, j, k) order (:1475-1477) and, while the cell is wider than 3 dcell_min, recurses
    on a 3^3 grid over the cube of half-width one cell around that cell's centre.  The last
    level returns the position of member number a among the members inside its grid, in
    member order, where a is the index of the densest member of the halo (first on ties):
    :1486-1490 index the in-grid members with the argmax over all members.  IndexError where
    that index is out of range, as in the reference."""
    x0 = np.array(c, dtype=np.float64)
    r0 = np.float64(r0)
    for _ in range(max_levels):
        xmin = x0 - r0
        deltax = 2 * r0 / nx
        ijk = ((pos - xmin) / deltax).astype(np.int64)
        valid = np.all((ijk >= 0) & (ijk < nx), axis=1)
        grid = np.zeros((nx, nx, nx))
        np.add.at(grid, (ijk[valid, 0], ijk[valid, 1], ijk[valid, 2]), mass[valid])
        cell = np.argwhere(grid == np.max(grid))[0]
        if deltax > 3 * dcell_min:
            x0 = xmin + (cell + 0.5) * deltax
            r0 = deltax
            nx = 3
            continue
        a = int(np.argmax(density))
        inside = np.nonzero(valid)[0]
        if a >= len(inside):
            raise IndexError(f"densest member {a} beyond the {len(inside)} members of the cell grid")
        return np.array(pos[inside[a]], dtype=np.float64)
    raise RecursionError("multiscale centre: too many levels")


def center_sphere(pos, mass, c, r0, L, dcell_min, eps_sc, max_iter=1000):
    """det_halo_center_sphere_1b1 :1498-1540, the recursion unrolled: the centre of mass of
    the members within r of the current centre (wrapped offsets, dr^2 <= r^2) becomes the
    centre and r shrinks by (1 - eps_SC) while r > dcell_min and the sphere holds mass; then
    the member nearest that centre of mass (first on ties).  ValueError where the reference
    raises (squared distance above 1)."""
    x0 = np.array(c, dtype=np.float64)
    r = np.float64(r0)
    for _ in range(max_iter):
        drx, dry, drz = rel(pos, x0, L)
        inside = drx ** 2 + dry ** 2 + drz ** 2 <= r * r
        m = mass[inside]
        mtot = np.sum(m)
        if mtot > 0:
            xc = np.array([np.sum(m * drx[inside]) / mtot + x0[0],
                           np.sum(m * dry[inside]) / mtot + x0[1],
                           np.sum(m * drz[inside]) / mtot + x0[2]])
        else:
            xc = x0
        if r > dcell_min and mtot > 0:
            x0 = xc
            r = (1 - eps_sc) * r
            continue
        drx, dry, drz = rel(pos, xc, L)
        d2 = drx ** 2 + dry ** 2 + drz ** 2
        i = int(np.argmin(d2))
        if d2[i] > 1.0:
            raise ValueError(f"distmin(={d2[i]}) > 1")
        return np.array(pos[i], dtype=np.float64)
    raise RecursionError("sphere centre: too many iterations")


def refine_centers(pos, mass, density, c, L, sc, dpmmc, dcell_min, eps_sc):
    """refine_centers_1b :1555-1567 (total centre; the stellar centre needs the stellar
    payload): DPMMC then SC, each from the radius of the farthest member."""
    r0 = radius_from_center(pos, c, L)
    if np.isfinite(r0) and r0 > 0:
        if dpmmc:
            c = center_multiscale(pos, mass, density, c, r0, 9, dcell_min)
            r0 = radius_from_center(pos, c, L)
        if sc:
            c = center_sphere(pos, mass, c, r0, L, dcell_min, eps_sc)
    return c


def rel(pos, c, L):
    """Member offsets from the centre, wrapped (used by :783-1032, :1400-1444, :1932-1951)."""
    return [wrap(pos[:, d] - c[d], L) for d in range(3)]


def ang_mom(pos, vel, mass, c, vbar, L):
    """compute_ang_mom_19 :783-824: L = sum dr x m (v - vbar)."""
    drx, dry, drz = rel(pos, c, L)
    pvx, pvy, pvz = (mass * (vel[:, d] - vbar[d]) for d in range(3))
    return np.array([np.sum(dry * pvz - drz * pvy), np.sum(drz * pvx - drx * pvz),
                     np.sum(drx * pvy - dry * pvx)])


def r_halo(pos, c, L):
    """r_halos_1a :828-851: distance of the farthest member from the centre."""
    drx, dry, drz = rel(pos, c, L)
    return np.sqrt(np.max(drx ** 2 + dry ** 2 + drz ** 2))


def inertia(pos, mass, c, L):
    """det_inertial_tensor_1b20 :1932-1951: I = sum m dr (x) dr."""
    dr = rel(pos, c, L)
    I = np.zeros((3, 3))
    for a in range(3):
        for b in range(a, 3):
            I[a, b] = I[b, a] = np.sum(mass * dr[a] * dr[b])
    return I


def main_axis(I, m):
    """det_main_axis_1b2 :1954-1967: eigh (ascending) -> axes sqrt(lambda/m), vectors."""
    from scipy.linalg import eigh
    d, v = eigh(I)
    return np.sqrt(d / m), v


def direct_pe(pos, mass, L, gravconst, periodic=True):
    """det_halo_energies_1b3 :1417-1427: pair sum over i<j of -m_i m_j / (d_ij/L) * G/L
    (the wrap only with FlagPeriod != 0)."""
    n = mass.shape[0]
    i, j = np.triu_indices(n, 1)
    drs = pos[j] - pos[i]
    if periodic:
        drs[drs > L / 2] -= L
        drs[drs < -L / 2] += L
    dists = np.sqrt(np.sum(drs ** 2, axis=1)) / L
    return np.sum(-(mass[i] * mass[j]) / dists) * gravconst / L


def rf_pe(m, axes, gravconst):
    """det_halo_energies_1b3 :1407-1412: -3/5 G m^2 RF(a^2, b^2, c^2) (scipy's elliprf, as
    num_rec.rf)."""
    from scipy.special import elliprf
    return (-3 / 5) * gravconst * m ** 2 * elliprf(axes[0] ** 2, axes[1] ** 2, axes[2] ** 2)


def kinetic(pos, vel, mass, c, vbar, L, hub_pt):
    """det_halo_energies_1b3 :1430-1441 -> (ek, per-member |v_t|^2 with the Hubble flow)."""
    drx, dry, drz = rel(pos, c, L)
    v2 = ((vel[:, 0] - vbar[0] + drx * hub_pt) ** 2 + (vel[:, 1] - vbar[1] + dry * hub_pt) ** 2 +
          (vel[:, 2] - vbar[2] + drz * hub_pt) ** 2)
    return 0.5 * np.sum(mass * v2), v2


def shell_scale(nr):
    """_shell_scale :1061-1066."""
    s = np.arange(nr, dtype=np.float64) / np.float64(nr - 1)
    s[0] = 0.5 / np.float64(nr - 1)
    return s


# the smallest distance of r_ell * nr from a shell boundary seen by the last shell_tables call
# (how far the member nearest a boundary is from changing shell: i_ell = int(r_ell * nr))
LAST_SHELL_MARGIN = [np.inf]


def shell_tables(pos, mass, v2, c, axes, vecs, L, nr):
    """tab_props_inside_1b40 :969-1032 -> (cumulative mass, cumulative kinetic energy,
    shell counts, amax, bmax, cmax)."""
    drx, dry, drz = rel(pos, c, L)
    dra = drx * vecs[0, 0] + dry * vecs[1, 0] + drz * vecs[2, 0]
    drb = drx * vecs[0, 1] + dry * vecs[1, 1] + drz * vecs[2, 1]
    drc = drx * vecs[0, 2] + dry * vecs[1, 2] + drz * vecs[2, 2]
    sha, shb, shc = axes
    rmax = np.max(np.sqrt((dra / sha) ** 2 + (drb / shb) ** 2 + (drc / shc) ** 2))
    eps = 1e-2
    amax, bmax, cmax = rmax * sha * (1.0 + eps), rmax * shb * (1.0 + eps), rmax * shc * (1.0 + eps)
    r_ell = np.sqrt((dra / amax) ** 2 + (drb / bmax) ** 2 + (drc / cmax) ** 2)
    i_ell = (r_ell * nr).astype(int)
    # (members at the centre, r_ell = 0, are in shell 0 in any arithmetic)
    frac = (r_ell * nr - i_ell)[r_ell > 0]
    LAST_SHELL_MARGIN[0] = float(np.min(np.minimum(frac, 1.0 - frac))) if frac.size else np.inf
    if np.any(i_ell >= nr):
        raise ValueError("tab_props_inside: a member falls outside the outer ellipsoid")
    tabm = np.zeros(nr)
    tabk = np.zeros(nr)
    np.add.at(tabm, i_ell, mass)
    np.add.at(tabk, i_ell, 0.5 * mass * v2)
    cnt = np.bincount(i_ell, minlength=nr).astype(np.float64)
    return np.cumsum(tabm), np.cumsum(tabk), cnt, amax, bmax, cmax


def potential_profile(tabm, cnt, amax, bmax, cmax, gravconst, mode="shell", self_mode="count",
                      axes=None):
    """potential_profile_1b40 :1069-1107 (before the outer-value normalisation): mode 'shell'
    with self mode 'count' or 'half_shell', or mode 'baseline'."""
    nr = tabm.shape[0]
    if mode == "baseline":
        from scipy.special import elliprf
        return -3 / 5 * gravconst * tabm ** 2 * elliprf(axes[0] ** 2, axes[1] ** 2, axes[2] ** 2)
    dm = np.diff(tabm, prepend=0.0)
    mass_before = tabm - dm
    if self_mode == "half_shell":
        self_factor = 0.5
    else:
        self_factor = np.zeros(nr)
        occ = cnt > 0
        self_factor[occ] = 0.5 * (cnt[occ] - 1.0) / cnt[occ]
    r_eff = shell_scale(nr) * (amax * bmax * cmax) ** (1.0 / 3.0)
    return -gravconst * np.cumsum((mass_before + self_factor * dm) * dm / r_eff)


def vir_props(tabm, tabk, tabp, amax, bmax, cmax, vir_overdens, rho_mean):
    """det_vir_props_1b4 :1127-1229 -> (mvir, rvir) before the min with r."""
    ttab = tabm.shape[0]
    mvir = tabm[ttab - 1]
    i1 = 1
    for i1 in range(ttab - 1, 0, -1):          # frange(ttab-1, 1, -1): down to 1 inclusive
        if tabm[i1 - 1] < tabm[i1]:
            break
    avir = i1 / (ttab - 1) * amax
    bvir = i1 / (ttab - 1) * bmax
    cvir = i1 / (ttab - 1) * cmax
    rvir = (avir * bvir * cvir) ** (1. / 3.)
    virth_old = 1.0
    virth = 100.0
    for i0 in range(ttab - 1, -1, -1):          # frange(ttab-1, 0, -1)
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
    return mvir, rvir


def ang_momentum_per_shell(pos, vel, mass, c, vbar, vecs, amax, bmax, cmax, L, nshells=100):
    """det_ang_momentum_per_shell_1b5 :1970-2013 -> (m, Lx, Ly, Lz) per shell, each (nshells,).
    Members in the reference's order, one at a time: dr wrapped as correct_for_periodicity
    (:854-887), projected on the principal axes (dra = dr.x v[0,0] + dr.y v[1,0] + dr.z v[2,0],
    ...), i_ell = int(r_ell * nshells) with r_ell = sqrt((dra/amax)**2 + (drb/bmax)**2 +
    (drc/cmax)**2), dv = v - the halo's mean velocity, and the shell gets m and
    m * (dr x dv).  i_ell > nshells is the reference's ValueError, i_ell = nshells its
    IndexError."""
    m = np.zeros(nshells)
    lx = np.zeros(nshells)
    ly = np.zeros(nshells)
    lz = np.zeros(nshells)
    for i in range(mass.shape[0]):
        dx = wrap_scalar(pos[i, 0] - c[0], L)
        dy = wrap_scalar(pos[i, 1] - c[1], L)
        dz = wrap_scalar(pos[i, 2] - c[2], L)
        dra = dx * vecs[0, 0] + dy * vecs[1, 0] + dz * vecs[2, 0]
        drb = dx * vecs[0, 1] + dy * vecs[1, 1] + dz * vecs[2, 1]
        drc = dx * vecs[0, 2] + dy * vecs[1, 2] + dz * vecs[2, 2]
        r_ell = np.sqrt((dra / amax) ** 2 + (drb / bmax) ** 2 + (drc / cmax) ** 2)
        i_ell = int(r_ell * nshells)
        if i_ell >= nshells:
            raise ValueError(f"{i_ell},{nshells}")
        vx, vy, vz = vel[i, 0] - vbar[0], vel[i, 1] - vbar[1], vel[i, 2] - vbar[2]
        m[i_ell] += mass[i]
        lx[i_ell] += mass[i] * (dy * vz - dz * vy)
        ly[i_ell] += mass[i] * (dz * vx - dx * vz)
        lz[i_ell] += mass[i] * (dx * vy - dy * vx)
    return m, lx, ly, lz


def halo_props(pos, vel, mass, density=None, *, lbox=1.0, hub_pt=70.3, gravconst=430.1,
               vir_overdens=340.0, rho_mean=1.0, cdm=True, nshell=1000, direct_pe_nmax=1000,
               sc=False, dpmmc=False, dcell_min=0.0, eps_sc=0.0, flag_period=1,
               virial_pe_profile="shell", virial_pe_norm_policy="profile_outer_for_large",
               virial_shell_self_mode="count", ang_mom_nshells=0, extras=None):
    """All properties of one halo, composed as det_vir_1b :1241-1259 and the property loop
    (det_mass_17, det_center_18, refine_centers_1b, compute_ang_mom_19, r_halos_1a,
    det_main_axis_1b2, det_halo_energies_1b3, det_vir_props_1b4, compute_spin_parameter_1c
    :1599-1602; the order of _compute_halo_props :270-290).  With ``ang_mom_nshells`` > 0 and
    a dict ``extras``, the per-shell angular momentum of det_ang_momentum_per_shell_1b5
    (:1257-1258, for a finite rvir) goes into extras["shells"] (4, nshells) and the ellipsoid
    into extras["axes"] (amax, bmax, cmax, v row-major); NaN otherwise."""
    out = np.full(len(FIELDS), np.nan)
    # FlagPeriod = 0: correct_for_periodicity returns its input (:862, :889), which the strict
    # comparisons with L/2 reproduce for an infinite L; the pair sum drops its wrap
    L = lbox if flag_period else np.inf
    n = mass.shape[0]
    m = np.sum(mass)                                          # det_mass_17 :754-780
    out[F["m"]] = m
    c, vbar, sigma = det_center(pos, vel, mass, density, L, cdm)
    if sc or dpmmc:
        c = refine_centers(pos, mass, density, c, L, sc, dpmmc, dcell_min, eps_sc)
    out[F["px"]], out[F["py"]], out[F["pz"]] = c
    out[F["vx"]], out[F["vy"]], out[F["vz"]] = vbar
    out[F["sigma"]] = sigma
    Lv = ang_mom(pos, vel, mass, c, vbar, L)
    out[F["Lx"]], out[F["Ly"]], out[F["Lz"]] = Lv
    out[F["r"]] = r_halo(pos, c, L)
    I = inertia(pos, mass, c, L)
    out[F["Ixx"]], out[F["Ixy"]], out[F["Ixz"]] = I[0, 0], I[0, 1], I[0, 2]
    out[F["Iyy"]], out[F["Iyz"]], out[F["Izz"]] = I[1, 1], I[1, 2], I[2, 2]
    axes, vecs = main_axis(I, m)
    out[F["sha"]], out[F["shb"]], out[F["shc"]] = axes
    # det_halo_energies_1b3: pair sum below direct_pe_nmax; the RF estimate of larger halos
    # is overwritten by the profile's outer value under the default norm policy
    ep = (direct_pe(pos, mass, lbox, gravconst, bool(flag_period)) if n < direct_pe_nmax
          else rf_pe(m, axes, gravconst))
    ek, v2 = kinetic(pos, vel, mass, c, vbar, L, hub_pt)
    tabm, tabk, cnt, amax, bmax, cmax = shell_tables(pos, mass, v2, c, axes, vecs, L, nshell)
    tabp = potential_profile(tabm, cnt, amax, bmax, cmax, gravconst, virial_pe_profile,
                             virial_shell_self_mode, axes)
    outer = tabp[-1]
    if virial_pe_norm_policy == "legacy":
        if ep != outer and outer != 0.0:
            tabp = tabp / outer * ep
    elif n < direct_pe_nmax:
        if ep != outer and outer != 0.0:
            tabp = tabp / outer * ep
    else:
        ep = outer
    et = ek + ep
    out[F["ek"]], out[F["ep"]], out[F["et"]] = ek, ep, et
    mvir, rvir = vir_props(tabm, tabk, tabp, amax, bmax, cmax, vir_overdens, rho_mean)
    if mvir > 0.0 and rvir > 0.0:
        out[F["rvir"]] = min(rvir, out[F["r"]])
        out[F["mvir"]] = mvir
        out[F["cvel"]] = np.sqrt(gravconst * mvir / out[F["rvir"]])
        out[F["tvir"]] = 35.9 * gravconst * mvir / out[F["rvir"]]
        # compute_halo_profile_1b41 :1232-1238, profile 'TSIS'
        out[F["rho_0"]] = out[F["mvir"]] / (4.0 * np.pi * out[F["rvir"]] ** 3)
        out[F["r_c"]] = out[F["rvir"]]
    hl = np.sqrt(np.sum(Lv ** 2))
    out[F["spin"]] = hl * np.sqrt(np.abs(et)) / m ** 2.5 / gravconst
    if ang_mom_nshells > 0 and extras is not None:
        extras["shells"] = np.full((4, ang_mom_nshells), np.nan)
        extras["axes"] = np.full(12, np.nan)
        if np.isfinite(out[F["rvir"]]):
            extras["shells"] = np.array(ang_momentum_per_shell(pos, vel, mass, c, vbar, vecs,
                                                               amax, bmax, cmax, L,
                                                               ang_mom_nshells))
            extras["axes"] = np.concatenate([[amax, bmax, cmax], vecs.ravel()])
    return out


def all_halos(offsets, pos, vel, mass, density=None, **kw):
    rows = []
    for h in range(len(offsets) - 1):
        s, e = int(offsets[h]), int(offsets[h + 1])
        rows.append(halo_props(pos[s:e], vel[s:e], mass[s:e],
                               None if density is None else density[s:e], **kw))
    return np.array(rows)
