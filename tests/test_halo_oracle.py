"""Pins of the per-halo property oracle (oracle/halo_props.py) by closed forms and
invariants -- the reference's Python reductions are not imported, so these checks are
what fixes the restatement (src/compute_halo_props.py:754-1229)."""
import numpy as np
import pytest

from oracle import halo_props as HP

F = HP.F
G = 430.1
L1 = 1.0


def props(pos, vel, mass, **kw):
    kw.setdefault("hub_pt", 0.0)
    return HP.halo_props(np.asarray(pos, float), np.asarray(vel, float), np.asarray(mass, float),
                         **kw)


def _static_symmetric_cloud(seed, npairs, scale):
    # a member at the origin plus point-symmetric pairs: the COM is the origin member
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(npairs, 3)) * scale
    return np.vstack([[0, 0, 0], p, -p])


def _cloud(seed, n):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * 0.01
    vel = rng.normal(size=(n, 3)) * 50.0
    mass = np.exp(rng.normal(size=n) * 0.3) / n
    return pos, vel, mass


def test_two_body_angular_momentum_and_centre():
    # a pair at +-d on x moving +-v on y inside a static point-symmetric cloud: the centre
    # is the origin member, mean velocity 0, and only the pair carries angular momentum,
    # L = 2 * (d x) x (m v y) = 2 d m v z
    d, v, m, mc = 0.01, 3.0, 0.5, 0.1
    cloud = _static_symmetric_cloud(0, 20, 0.004)
    pos = np.vstack([cloud, [[d, 0, 0], [-d, 0, 0]]])
    vel = np.zeros_like(pos)
    vel[-2:] = [[0, v, 0], [0, -v, 0]]
    mass = np.r_[np.full(len(cloud), mc), m, m]
    o = props(pos, vel, mass)
    assert (o[F["px"]], o[F["py"]], o[F["pz"]]) == (0.0, 0.0, 0.0)
    assert (o[F["vx"]], o[F["vy"]], o[F["vz"]]) == (0.0, 0.0, 0.0)
    assert o[F["Lx"]] == 0 and o[F["Ly"]] == 0
    assert o[F["Lz"]] == pytest.approx(2 * d * m * v, rel=1e-14)
    assert o[F["sigma"]] == pytest.approx(v * np.sqrt(2 * m / mass.sum()), rel=1e-14)
    assert o[F["m"]] == pytest.approx(mass.sum(), rel=1e-15)


def test_centre_is_the_member_nearest_the_periodic_com():
    # COM about the first member, wrapped: two members either side of the x edge
    pos = np.array([[0.499, 0, 0], [-0.499, 0, 0], [0.4995, 0.001, 0]])
    c, vbar, _ = HP.det_center(pos, np.zeros((3, 3)), np.ones(3), None, L1, True)
    # COM in the unwrapped frame of member 0 is x = 0.499 + (0 + 0.002 + 0.0005)/3
    assert np.array_equal(c, pos[2]) and np.all(vbar == 0)
    # ties go to the first member
    pos = np.array([[-0.01, 0, 0], [0.01, 0, 0]])
    c, _, _ = HP.det_center(pos, np.zeros((2, 3)), np.ones(2), None, L1, True)
    assert np.array_equal(c, pos[0])


def test_rigid_rotation_angular_momentum():
    # ring of radius R around the centre, v = omega x r: L = omega * sum m R^2 along z
    R, om, nring, m = 0.02, 5.0, 12, 1e-3
    phi = 2 * np.pi * np.arange(nring) / nring
    pos = np.c_[R * np.cos(phi), R * np.sin(phi), np.zeros(nring)]
    vel = np.c_[-om * R * np.sin(phi), om * R * np.cos(phi), np.zeros(nring)]
    Lv = HP.ang_mom(pos, vel, np.full(nring, m), np.zeros(3), np.zeros(3), L1)
    assert Lv[2] == pytest.approx(om * nring * m * R * R, rel=1e-12)
    assert abs(Lv[0]) < 1e-18 and abs(Lv[1]) < 1e-18


def test_inertia_tensor_and_axes_of_a_cube_shell():
    # 8 cube corners (+-a)^3 and a centre member: I = 8 m a^2 * identity, all three axes
    # sqrt(8 m a^2 / (9 m)) (det_inertial_tensor_1b20, det_main_axis_1b2)
    a, m = 0.003, 2.0
    corners = np.array([[sx, sy, sz] for sx in (-a, a) for sy in (-a, a) for sz in (-a, a)])
    pos = np.vstack([[0, 0, 0], corners])
    o = props(pos, np.zeros((9, 3)), np.full(9, m))
    for k in ("Ixx", "Iyy", "Izz"):
        assert o[F[k]] == pytest.approx(8 * m * a * a, rel=1e-14)
    for k in ("Ixy", "Ixz", "Iyz"):
        assert abs(o[F[k]]) < 1e-20
    for k in ("sha", "shb", "shc"):
        assert o[F[k]] == pytest.approx(a * np.sqrt(8 / 9), rel=1e-12)
    assert o[F["r"]] == pytest.approx(a * np.sqrt(3), rel=1e-15)


def test_axes_are_ascending_and_follow_an_elongation():
    # an elongated cloud along z: sha <= shb <= shc, the largest axis being z's spread
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(400, 3)) * [0.001, 0.002, 0.004]
    I = HP.inertia(pos, np.ones(400), np.zeros(3), L1)
    axes, vecs = HP.main_axis(I, 400.0)
    assert axes[0] <= axes[1] <= axes[2]
    assert axes[2] == pytest.approx(0.004, rel=0.1) and abs(vecs[2, 2]) > 0.99


@pytest.mark.parametrize("shape,npairs", [("pair", 1), ("triangle", 3), ("tetrahedron", 6)])
def test_direct_potential_energy_of_regular_systems(shape, npairs):
    # every pair at the same separation s: ep = -npairs * G m^2 / s (det_halo_energies_1b3)
    s, m = 0.004, 0.25
    if shape == "pair":
        pos = [[0, 0, 0], [s, 0, 0]]
    elif shape == "triangle":
        pos = [[0, 0, 0], [s, 0, 0], [s / 2, s * np.sqrt(3) / 2, 0]]
    else:
        pos = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) * s / np.sqrt(8)
    pos = np.asarray(pos, float)
    n = len(pos)
    ep = HP.direct_pe(pos, np.full(n, m), L1, G)
    assert ep == pytest.approx(-npairs * G * m * m / s, rel=1e-12)
    # the same pair across the periodic edge
    if shape == "pair":
        ep2 = HP.direct_pe(np.array([[0.4995, 0, 0], [-0.4995 + s - 0.001, 0, 0]]),
                           np.full(2, m), L1, G)
        assert ep2 == pytest.approx(-G * m * m / s, rel=1e-9)


def test_kinetic_energy_includes_the_hubble_flow():
    # members at rest, H = 70: ek = 1/2 sum m (H dr)^2 about the centre
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0, -0.02, 0]])
    ek, _ = HP.kinetic(pos, np.zeros((3, 3)), np.array([1.0, 2.0, 3.0]), np.zeros(3),
                       np.zeros(3), L1, 70.0)
    assert ek == pytest.approx(0.5 * (2 * 0.7 ** 2 + 3 * 1.4 ** 2), rel=1e-13)


def test_shell_profile_of_single_occupied_shells():
    # one member per occupied shell: no self pairs, so shell b adds -G m_before m_b / r_b;
    # two members in one shell add the self term -G (1/4) m^2 / r_b (0.5 (n-1)/n of dm^2)
    nr = 10
    tabm = np.zeros(nr)
    cnt = np.zeros(nr)
    tabm[3:] = 1.0
    cnt[3] = 1
    tabm[6:] = 3.0
    cnt[6] = 2
    p = HP.potential_profile(tabm, cnt, 0.9, 0.9, 0.9, G)
    r6 = 6 / 9 * 0.9
    assert np.all(p[:6] == 0.0)
    assert p[6] == pytest.approx(-G * (1.0 * 2.0 + 0.25 * 2.0 * 2.0) / r6, rel=1e-14)
    assert np.all(p[6:] == p[6])


def test_translation_across_the_periodic_edge():
    # moving a halo so that it straddles the box edge changes only its centre
    pos, vel, mass = _cloud(1, 300)
    a = props(pos, vel, mass, hub_pt=70.3)
    shift = np.array([0.5 - 0.002, -0.5 + 0.001, 0.25])
    p2 = np.mod(pos + shift + 0.5, 1.0) - 0.5
    assert np.any(p2[:, 0] < 0) and np.any(p2[:, 0] > 0.4)
    b = props(p2, vel, mass, hub_pt=70.3)
    for k in HP.FIELDS:
        if k in ("px", "py", "pz"):
            continue
        assert b[F[k]] == pytest.approx(a[F[k]], rel=1e-9, abs=1e-12 * abs(a[F["m"]])), k
    c = np.array([b[F["px"]], b[F["py"]], b[F["pz"]]])
    c0 = np.mod(np.array([a[F["px"]], a[F["py"]], a[F["pz"]]]) + shift + 0.5, 1.0) - 0.5
    assert np.allclose(c, c0, atol=1e-15)


def test_rotation_covariance_of_L_and_I():
    # rotating positions and velocities rotates L as a vector and I as a tensor
    pos, vel, mass = _cloud(2, 500)
    th = 0.7
    Rm = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    a = props(pos, vel, mass)
    b = props(pos @ Rm.T, vel @ Rm.T, mass)
    La = np.array([a[F["Lx"]], a[F["Ly"]], a[F["Lz"]]])
    Lb = np.array([b[F["Lx"]], b[F["Ly"]], b[F["Lz"]]])
    assert np.allclose(Rm @ La, Lb, rtol=1e-10, atol=1e-12 * np.linalg.norm(La))

    def tens(o):
        return np.array([[o[F["Ixx"]], o[F["Ixy"]], o[F["Ixz"]]],
                         [o[F["Ixy"]], o[F["Iyy"]], o[F["Iyz"]]],
                         [o[F["Ixz"]], o[F["Iyz"]], o[F["Izz"]]]])
    assert np.allclose(Rm @ tens(a) @ Rm.T, tens(b), rtol=1e-10, atol=1e-14)
    for k in ("sha", "shb", "shc", "ek", "ep", "spin"):
        assert b[F[k]] == pytest.approx(a[F[k]], rel=1e-9), k


def test_virial_overdensity_fallback_on_a_static_halo():
    # members at rest with no Hubble flow: K = 0, so |2K+W|/|K+W| = 1 never improves on
    # the initial 1 and the overdensity branch decides.  With a negligible overdensity the
    # first shell (from outside) whose inner neighbour holds less than the total is the
    # outermost occupied one: r_ell max = 1/(1+eps) -> bin int(nr/1.01) = 990 of 1000, so
    # mvir = m and rvir = 990/999 * (amax bmax cmax)^(1/3), amax = 1.01 rmax_ell sha, ...
    pos, _, mass = _cloud(3, 800)
    o = props(pos, np.zeros((800, 3)), mass, vir_overdens=1e-30)
    assert o[F["mvir"]] == pytest.approx(o[F["m"]], rel=1e-14)
    p0 = np.array([o[F["px"]], o[F["py"]], o[F["pz"]]])
    I = HP.inertia(pos, mass, p0, L1)
    lam, V = np.linalg.eigh(I)
    ax = np.sqrt(lam / o[F["m"]])
    q = ((pos - p0) @ V) / ax
    rmax = np.sqrt((q ** 2).sum(1)).max()
    expect = min(990 / 999 * np.prod(1.01 * rmax * ax) ** (1 / 3), o[F["r"]])
    assert o[F["rvir"]] == pytest.approx(expect, rel=1e-10)
    assert o[F["cvel"]] == pytest.approx(np.sqrt(G * o[F["mvir"]] / o[F["rvir"]]), rel=1e-14)


def test_virial_scan_takes_the_first_well_virialised_shell():
    # tables where shell 7 (from outside) is bound with |2K+W|/|K+W| = 0.1: the scan stops
    # there, mvir = tabm[7], axes scaled by 7/(ttab-1)
    nr = 10
    tabm = np.arange(1, nr + 1, dtype=float)
    tabk = np.full(nr, 5.0)
    tabp = np.full(nr, 1.0)          # unbound everywhere but shell 7
    tabk[7], tabp[7] = 9.0, -19.0    # K+W = -10, 2K+W = -1 -> 0.1
    mvir, rvir = HP.vir_props(tabm, tabk, tabp, 1.0, 2.0, 4.0, 1e-30, 1.0)
    assert mvir == 8.0
    assert rvir == pytest.approx(7 / 9 * 2.0, rel=1e-14)


def test_large_halo_takes_the_shell_profile_energy():
    # at or above the pair-sum limit ep is the outer value of the shell profile: for a
    # bound cloud it is negative and of the order of the direct pair sum
    pos, vel, mass = _cloud(4, 1200)
    big = props(pos, vel * 0.01, mass, direct_pe_nmax=1000)
    direct = props(pos, vel * 0.01, mass, direct_pe_nmax=5000)
    assert big[F["ep"]] < 0
    assert big[F["ep"]] == pytest.approx(direct[F["ep"]], rel=0.5)
    assert big[F["et"]] == pytest.approx(big[F["ek"]] + big[F["ep"]], rel=1e-15)
    assert direct[F["ep"]] == pytest.approx(HP.direct_pe(pos, mass, L1, G), rel=1e-15)


def test_densest_member_centre_takes_the_first_of_equal_maxima():
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0.02, 0, 0], [0.03, 0, 0]])
    dens = np.array([1.0, 5.0, 5.0, 2.0])
    c, _, _ = HP.det_center(pos, np.zeros((4, 3)), np.ones(4), dens, L1, False)
    assert np.array_equal(c, pos[1])


def test_spin_parameter_closed_form():
    pos, vel, mass = _cloud(5, 200)
    o = props(pos, vel, mass, hub_pt=70.3)
    hl = np.sqrt(o[F["Lx"]] ** 2 + o[F["Ly"]] ** 2 + o[F["Lz"]] ** 2)
    assert o[F["spin"]] == pytest.approx(hl * np.sqrt(abs(o[F["et"]])) / o[F["m"]] ** 2.5 / G,
                                         rel=1e-15)


# ---- recentring (refine_centers_1b :1555-1596)

def _line(xs, L=10.0):
    pos = np.zeros((len(xs), 3))
    pos[:, 0] = xs
    return pos


def test_shrinking_sphere_hand_worked():
    """det_halo_center_sphere_1b1 :1498-1540 on seven members on a line, eps_SC=0.75,
    dcell_min=0.5, starting on member 0 (r0 = 2, the farthest member at x=-2).
    Sphere 1 (r=2, dr^2 <= r^2 keeps x=-2): masses 17, sum m x = 2.1, centre 0.123529...;
    r becomes 0.5, not above dcell_min, so the last sphere: members 0, 5, 6 (masses 2, 2, 3)
    lie within 0.5 of it, their centre of mass is -0.00714..., nearest member 0.
    A strict < (x=-2 dropped), a shrink by eps instead of 1-eps, the nearest member to
    the previous centre, or >= dcell_min give members 3, 5, 5 and 6 instead."""
    xs = [0.0, 1.69, -2.0, 0.69, 0.85, 0.17, -0.13]
    m = np.array([2.0, 3.0, 3.0, 2.0, 2.0, 2.0, 3.0])
    pos = _line(xs)
    c = HP.refine_centers(pos, m, None, pos[0], 10.0, True, False, 0.5, 0.75)
    assert np.array_equal(c, pos[0])
    # the same members moved across the periodic edge: the same member, wrapped
    shifted = pos.copy()
    shifted[:, 0] += 4.0
    shifted[shifted[:, 0] >= 5.0, 0] -= 10.0
    c = HP.refine_centers(shifted, m, None, shifted[0], 10.0, True, False, 0.5, 0.75)
    assert np.array_equal(c, shifted[0])


def test_shrinking_sphere_finds_a_core_through_a_decoy():
    """A compact core (a member at P plus point-symmetric pairs within 1e-4) inside a broad
    cloud centred 0.02 away, across the periodic edge: starting from the cloud's
    centre-of-mass member, the shrinking sphere ends on the member at P."""
    rng = np.random.default_rng(3)
    P = np.array([0.4995, -0.2, 0.1])
    q = rng.normal(size=(40, 3)) * 1e-4
    core = np.vstack([P, P + q, P - q])
    cloud = rng.normal(size=(300, 3)) * 0.01 + P + np.array([0.02, 0.0, 0.0])
    cloud = cloud[np.linalg.norm(HP.wrap(cloud - P, 1.0), axis=1) > 3e-3]
    pos = np.vstack([cloud, core])
    pos = (pos + 0.5) % 1.0 - 0.5
    m = np.full(len(pos), 1.0)
    c0, _, _ = HP.det_center(pos, np.zeros_like(pos), m, None, 1.0, True)
    assert np.linalg.norm(HP.wrap(c0 - P, 1.0)) > 5e-3        # the plain centre is elsewhere
    c = HP.refine_centers(pos, m, None, c0, 1.0, True, False, 1e-3, 0.3)
    assert np.array_equal(c, pos[len(cloud)])


def test_shrinking_sphere_above_the_cell_size_is_the_com_centre():
    """dcell_min above the starting radius: one sphere holding every member, then the member
    nearest its centre of mass -- the centre det_center_18 chose."""
    pos, vel, mass = _cloud(5, 200)
    c0, _, _ = HP.det_center(pos, vel, mass, None, 1.0, True)
    c = HP.refine_centers(pos, mass, None, pos[17], 1.0, True, False, 10.0, 0.3)
    assert np.array_equal(c, c0)


def _grid_halo():
    """Members around the origin for a 9^3 grid of half-width 1 (cells 2/9 wide): far
    members 1-3 and a clump of six (members 4-9) in cell (6, 6, 6)."""
    far = [[0.9, -0.8, 0.1], [-0.7, 0.8, -0.6], [0.2, -0.9, 0.85]]
    d = np.array([[0.0, 0.01, -0.02], [0.02, -0.01, 0.0], [-0.02, 0.0, 0.01],
                  [0.01, 0.02, 0.02], [-0.01, -0.02, -0.01], [0.03, 0.03, -0.03]])
    pos = np.vstack([[0.0, 0.0, 0.0], far, 0.45 + d])
    return pos, np.ones(len(pos))


def test_multiscale_centre_indexes_the_cell_members_with_the_densest_member():
    """det_halo_center_multiscale_1b0 :1447-1495 with dcell_min=0.06: level 0 (cells 2/9 >
    0.18) picks the clump's cell (mass 6) and recurses on 3^3 cells over [0.222, 0.667]^3,
    where only the clump lies; cells 0.148 wide end the recursion.  The densest member of
    the halo is member 1, so the centre is the clump's member number 1 (member 5): the
    reference indexes the in-grid members with the argmax over all members (:1486-1490).
    Not the densest member (1), nor the densest in the cell (member 8)."""
    pos, m = _grid_halo()
    dens = np.array([1.0, 100.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 50.0, 8.0])
    c = HP.center_multiscale(pos, m, dens, pos[0], 1.0, 9, 0.06)
    assert np.array_equal(c, pos[5])
    # one level (cells not above 3 dcell_min): every member but the far corner ones lies in
    # the grid, in member order, so the densest member's own position
    c = HP.center_multiscale(pos, m, dens, pos[0], 1.0, 9, 1.0)
    assert np.array_equal(c, pos[1])
    # densest member 9 beyond the six members of the last grid: the reference's IndexError
    dens[9] = 1e3
    with pytest.raises(IndexError):
        HP.center_multiscale(pos, m, dens, pos[0], 1.0, 9, 0.06)


def test_multiscale_takes_the_first_cell_of_equal_maximum_mass_in_ijk_order():
    """Two clumps of three in cells (6, 2, 6) (members 1-3) and (2, 6, 6) (members 4-6):
    np.where's first maximum is (2, 6, 6) (C order; Fortran order would pick (6, 2, 6)).
    The densest member is member 0, so the centre is the first member of the chosen clump."""
    a = np.array([[0.45, -0.45, 0.45]]) + np.array([[0, 0.01, 0], [0.01, 0, 0], [0, 0, 0.01]])
    b = np.array([[-0.45, 0.45, 0.45]]) + np.array([[0, 0.01, 0], [0.01, 0, 0], [0, 0, 0.01]])
    pos = np.vstack([[0.0, 0.0, 0.0], a, b])
    m = np.ones(len(pos))
    dens = np.array([9.0, 1, 1, 1, 1, 1, 1])
    c = HP.center_multiscale(pos, m, dens, pos[0], 1.0, 9, 0.06)
    assert np.array_equal(c, pos[4])


def test_refine_centers_runs_dpmmc_then_sc():
    """refine_centers_1b :1555-1567: DPMMC, the radius from its centre, then SC from there."""
    pos, m = _grid_halo()
    dens = np.array([1.0, 100.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 50.0, 8.0])
    c1 = HP.center_multiscale(pos, m, dens, pos[0], HP.radius_from_center(pos, pos[0], 10.0),
                              9, 0.06)
    r1 = HP.radius_from_center(pos, c1, 10.0)
    c2 = HP.center_sphere(pos, m, c1, r1, 10.0, 0.06, 0.3)
    c = HP.refine_centers(pos, m, dens, pos[0], 10.0, True, True, 0.06, 0.3)
    assert np.array_equal(c, c2)
    # off: the centre is left alone
    assert np.array_equal(HP.refine_centers(pos, m, dens, pos[3], 10.0, False, False, 0.06, 0.3),
                          pos[3])


def test_rf_potential_of_a_sphere():
    """det_halo_energies_1b3's estimate -3/5 G m^2 RF(a^2, b^2, c^2) (:1407-1412): for equal
    axes RF(a^2, a^2, a^2) = 1/a, the potential energy -3/5 G m^2 / a of a uniform sphere."""
    from oracle.halo_props import rf_pe
    assert np.isclose(rf_pe(2.0, (0.5, 0.5, 0.5), 430.1), -0.6 * 430.1 * 4.0 / 0.5, rtol=1e-15)
    # RF decreases in each argument, so axes (1, 0.5, 0.5) lie between the spheres of radius
    # 0.5 and 1
    e = rf_pe(1.0, (1.0, 0.5, 0.5), 1.0)
    assert -0.6 / 0.5 < e < -0.6 / 1.0


def test_virial_profile_switches():
    """potential_profile_1b40's switches on a two-shell table (:1085-1105): 'half_shell' takes
    half of each shell's own mass against itself where 'count' takes (n-1)/(2n); 'baseline'
    is -3/5 G M(<r)^2 RF."""
    from oracle.halo_props import potential_profile, shell_scale
    from scipy.special import elliprf
    tabm = np.array([2.0, 5.0])
    cnt = np.array([2.0, 3.0])
    a = b = c = 1.0
    r = shell_scale(2) * 1.0                           # 0.5, 1.0
    G = 2.0
    count = potential_profile(tabm, cnt, a, b, c, G)
    # shell 0: (0 + 1/4 * 2) * 2 / 0.5; shell 1: (2 + 1/3 * 3) * 3 / 1
    assert np.allclose(count, -G * np.cumsum([0.25 * 2 * 2 / r[0], (2 + 3 / 3) * 3 / r[1]]))
    half = potential_profile(tabm, cnt, a, b, c, G, self_mode="half_shell")
    assert np.allclose(half, -G * np.cumsum([0.5 * 2 * 2 / r[0], (2 + 0.5 * 3) * 3 / r[1]]))
    axes = (1.0, 0.8, 0.5)
    base = potential_profile(tabm, cnt, a, b, c, G, mode="baseline", axes=axes)
    assert np.allclose(base, -0.6 * G * tabm ** 2 * elliprf(1.0, 0.64, 0.25))


def test_halo_switches_compose():
    """In halo_props: 'legacy' rescales the profile of a large halo to the RF estimate, which
    stays its ep; the default policy sets ep to the profile's outer value.  FlagPeriod 0 on a
    halo across the box edge measures it unwrapped."""
    from oracle.halo_props import F, halo_props, rf_pe
    rng = np.random.default_rng(2)
    n = 1200
    pos = rng.normal(0, 0.01, (n, 3))
    vel = rng.normal(0, 50, (n, 3))
    mass = np.full(n, 1e-3)
    d = halo_props(pos, vel, mass)
    leg = halo_props(pos, vel, mass, virial_pe_norm_policy="legacy")
    assert np.isclose(leg[F["ep"]], rf_pe(d[F["m"]], (d[F["sha"]], d[F["shb"]], d[F["shc"]]),
                                          430.1), rtol=1e-14)
    assert leg[F["ep"]] != d[F["ep"]]
    for k in ("m", "px", "Lx", "r", "sha", "ek"):
        assert leg[F[k]] == d[F[k]]
    # across the edge: unwrapped, the members reach across the box
    edge = pos + np.array([0.5, 0.0, 0.0])
    edge[:, 0] = np.where(edge[:, 0] >= 0.5, edge[:, 0] - 1.0, edge[:, 0])
    wrapped = halo_props(edge, vel, mass)
    flat = halo_props(edge, vel, mass, flag_period=0)
    assert np.isclose(wrapped[F["r"]], d[F["r"]], rtol=1e-9)
    assert flat[F["r"]] > 0.4


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_angular_momentum_shells_by_hand():
    """det_ang_momentum_per_shell_1b5 on members placed by hand along the principal axes of a
    rotated frame (the axes are the columns of v, dra = dr . v[:, 0]): each lands in the
    shell int(r_ell * nshells) of its fraction of amax / bmax / cmax, and the shell holds m
    and m (dr x dv) with dv against the mean velocity.  A transposed v (rows as axes) puts
    them elsewhere.  A member at r_ell >= 1 is the reference's error."""
    from oracle.halo_props import ang_momentum_per_shell
    v = _rotation(7)
    amax, bmax, cmax = 0.3, 0.2, 0.1
    c = np.array([0.49, -0.2, 0.1])
    vbar = np.array([10.0, -5.0, 2.0])
    # (axis, fraction of that semi-axis, velocity offset)
    members = [(0, 0.055, (1.0, 2.0, 3.0)), (1, 0.505, (0.0, -4.0, 1.0)),
               (2, 0.9995, (2.0, 0.0, -1.0)), (0, -0.3055, (0.5, 0.5, 0.5))]
    semi = (amax, bmax, cmax)
    raw = np.array([c + f * semi[a] * v[:, a] for a, f, _ in members])
    assert (raw >= 0.5).any()                        # some member wraps across the edge
    pos = np.where(raw >= 0.5, raw - 1.0, raw)
    vel = np.array([vbar + np.array(dv) for _, _, dv in members])
    mass = np.array([1.0, 2.0, 3.0, 4.0])
    m, lx, ly, lz = ang_momentum_per_shell(pos, vel, mass, c, vbar, v, amax, bmax, cmax, 1.0)
    want = {5: [0], 50: [1], 99: [2], 30: [3]}
    assert set(np.nonzero(m)[0]) == set(want)
    for shell, idx in want.items():
        i = idx[0]
        a, f, dv = members[i]
        dr = f * semi[a] * v[:, a]
        Lw = mass[i] * np.cross(dr, np.array(dv))
        assert m[shell] == mass[i]
        assert np.allclose([lx[shell], ly[shell], lz[shell]], Lw, rtol=1e-12, atol=1e-15)
    mt, *_ = ang_momentum_per_shell(pos, vel, mass, c, vbar, v.T.copy(), amax, bmax, cmax, 1.0)
    assert set(np.nonzero(mt)[0]) != set(want)
    out = np.vstack([pos, c + 1.01 * amax * v[:, 0]])
    with pytest.raises(ValueError):
        ang_momentum_per_shell(out, np.vstack([vel, vbar]), np.append(mass, 1.0), c, vbar, v,
                               amax, bmax, cmax, 1.0)


def test_angular_momentum_shells_sum_to_the_halo():
    """Summed over the shells, the per-shell masses give the halo's mass and the per-shell
    angular momenta its L (compute_ang_mom_19, pinned by its own closed forms); the shells
    of halo_props are those of the all-member ellipsoid, so the outermost occupied shell is
    below nshells, and a halo without a finite rvir has NaN shells."""
    from oracle.halo_props import F, halo_props
    rng = np.random.default_rng(3)
    n = 800
    pos = rng.normal(0, 0.01, (n, 3)) * np.array([1.0, 0.6, 0.3])
    pos[:, 0] += 0.495                                 # across the periodic edge
    pos = np.where(pos >= 0.5, pos - 1.0, pos)
    vel = rng.normal(0, 60, (n, 3)) + np.cross(pos, [0.0, 0.0, 3000.0])
    mass = rng.uniform(0.5, 1.5, n) * 1e-3
    ex = {}
    out = halo_props(pos, vel, mass, ang_mom_nshells=100, extras=ex)
    assert np.isfinite(out[F["rvir"]])
    m, lx, ly, lz = ex["shells"]
    assert np.isclose(m.sum(), out[F["m"]], rtol=1e-13)
    Lh = out[[F["Lx"], F["Ly"], F["Lz"]]]
    assert np.allclose([lx.sum(), ly.sum(), lz.sum()], Lh, rtol=1e-11,
                       atol=1e-11 * np.linalg.norm(Lh))
    assert m[-1] > 0 or m[-2] > 0                      # the ellipsoid is fitted to the members
    ax = ex["axes"]
    assert np.allclose(ax[:3] / ax[0], [1.0, out[F["shb"]] / out[F["sha"]],
                                        out[F["shc"]] / out[F["sha"]]], rtol=1e-12)


def test_ang_mom_of_r_file_layout(tmp_path):
    """write_ang_mom_of_r: Fortran sequential records as FortranFile writes them, the
    header (nb_of_halos, nb_of_subhalos), (nshells), then one record per halo with a finite
    rvir: id, the 11 halo fields, amax, bmax, cmax, v, and the four per-shell arrays."""
    from scipy.io import FortranFile
    from amdahop import HALO_FIELDS, AngMomShells, write_ang_mom_of_r
    rng = np.random.default_rng(5)
    nh, nsh = 4, 7
    out = rng.normal(size=(nh, len(HALO_FIELDS)))
    out[2, HALO_FIELDS.index("rvir")] = np.nan
    shells = rng.normal(size=(nh, 4, nsh))
    axes = rng.normal(size=(nh, 12))
    ids = np.array([3, 5, 8, 9])
    path = tmp_path / "ang_mom_of_r.007"
    assert write_ang_mom_of_r(str(path), ids, out, AngMomShells(shells, axes), 3, 1) == 3
    f = FortranFile(str(path), "r")
    assert list(f.read_record(np.int64)) == [3, 1]
    assert list(f.read_record(np.int64)) == [nsh]
    cols = [HALO_FIELDS.index(k) for k in ("px", "py", "pz", "vx", "vy", "vz", "rvir", "mvir",
                                            "m", "r", "spin")]
    for h in (0, 1, 3):
        rec = f.read_record(np.dtype([("id", "<i4"), ("f", "<f8", 11), ("ax", "<f8", 12),
                                      ("sh", "<f8", (4, nsh))]))[0]
        assert rec["id"] == ids[h]
        assert np.array_equal(rec["f"], out[h, cols])
        assert np.array_equal(rec["ax"], axes[h])
        assert np.array_equal(rec["sh"], shells[h])
    with pytest.raises(Exception):
        f.read_record(np.int64)
    f.close()
