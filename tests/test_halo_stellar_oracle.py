"""Pins of the stellar-property oracle (oracle/halo_stellar.py, restating the stellar parts
of src/compute_halo_props.py) by constructions with answers known in closed form."""
import numpy as np
import pytest

from oracle import halo_stellar as HS

F = {k: i for i, k in enumerate(HS.STELLAR_FIELDS)}


def props(pos, vel, mass, dens, star, age=None, metal=None, vbar=(0, 0, 0), **kw):
    n = len(mass)
    age = np.zeros(n) if age is None else np.asarray(age, float)
    metal = np.zeros(n) if metal is None else np.asarray(metal, float)
    return HS.stellar_props(np.asarray(pos, float), np.asarray(vel, float),
                            np.asarray(mass, float), np.asarray(dens, float),
                            np.asarray(star, bool), age, metal, np.asarray(vbar, float), **kw)


def test_half_and_ninety_percent_radii_by_counting():
    """Eleven unit-mass stars at 0, 1, ..., 10 mm from the densest (at 0): the cumulative
    mass first reaches half of 11 at the sixth star (5 mm) and 90% (9.9) at the tenth (9 mm)."""
    r = np.arange(11) * 1e-3
    pos = np.zeros((11, 3))
    pos[:, 0] = r
    dens = np.ones(11)
    dens[0] = 10.0
    o = props(pos, np.zeros((11, 3)), np.ones(11), dens, np.ones(11))
    assert (o["px*"], o["py*"], o["pz*"]) == (0.0, 0.0, 0.0)
    assert o["r50"] == r[5] and o["r90"] == r[9] and o["r*"] == r[10]
    assert o["nstar"] == 11 and o["ndm"] == 0 and o["m*"] == 11.0


def test_ages_metals_and_star_formation_rates():
    pos = np.array([[0, 0, 0], [0.001, 0, 0], [0.002, 0, 0], [0.003, 0, 0]], float)
    m = np.array([1.0, 2.0, 1.0, 4.0])
    age = np.array([0.005, 0.05, 0.5, 5.0])        # Gyr: 10 Myr and 100 Myr cuts
    metal = np.array([0.02, 0.01, 0.0, 0.03])
    o = props(pos, np.zeros((4, 3)), m, [9, 1, 1, 1], np.ones(4), age, metal)
    assert o["age"] == pytest.approx((0.005 + 0.1 + 0.5 + 20.0) / 8, rel=1e-15)
    assert o["metal"] == pytest.approx((0.02 + 0.02 + 0.12) / 8, rel=1e-15)
    assert o["SFR"] == 3.0e3 and o["SFR10"] == 1.0e4
    # cumulative 1, 3, 4, 8: half (4) at the third star (2 mm), 90% (7.2) at the fourth (3 mm);
    # inside r50 is strict: the first two stars
    assert o["r50"] == 0.002 and o["r90"] == 0.003
    assert o["SFR_r50"] == 3.0e3 and o["SFR10_r50"] == 1.0e4 and o["SFR_r90"] == 3.0e3


def test_cold_rotating_ring():
    """Unit-mass stars on a ring of radius 0.01 with tangential speed 200 and a light central
    star (the densest, at rest): L* along +z, vrot = 200 and sigcyl = 0 over the ring stars
    (the centre has no cylindrical radius), sig3d^2 the mass-weighted mean of v^2."""
    n = 16
    t = 2 * np.pi * np.arange(n) / n
    ring = np.stack([0.01 * np.cos(t), 0.01 * np.sin(t), np.zeros(n)], axis=1)
    vring = np.stack([-200 * np.sin(t), 200 * np.cos(t), np.zeros(n)], axis=1)
    pos = np.vstack([[0, 0, 0], ring])
    vel = np.vstack([[0, 0, 0], vring])
    m = np.concatenate([[0.5], np.ones(n)])
    dens = np.concatenate([[100.0], np.ones(n)])
    o = props(pos, vel, m, dens, np.ones(n + 1))
    assert o["Lz*"] == pytest.approx(n * 0.01 * 200, rel=1e-12)
    assert abs(o["Lx*"]) < 1e-12 and abs(o["Ly*"]) < 1e-12
    assert o["vrot"] == pytest.approx(200.0, rel=1e-12)
    assert o["sigcyl"] < 1e-9
    assert o["sig3d"] == pytest.approx(200.0 * np.sqrt(n / (n + 0.5)), rel=1e-12)
    assert o["r50"] == pytest.approx(0.01, rel=1e-12)


def test_no_rotation_axis_leaves_the_cylindrical_terms_nan():
    """Radial velocities only: L* = 0, so no rotation axis; sig3d is still defined."""
    p = np.array([[0.01, 0, 0], [0, 0.01, 0], [0, 0, 0.01]])
    pos = np.vstack([[0, 0, 0], p, -p])
    vel = np.vstack([[0, 0, 0], 50 * p / 0.01, -50 * p / 0.01])
    o = props(pos, vel, np.ones(7), [9, 1, 1, 1, 1, 1, 1], np.ones(7))
    assert (o["Lx*"], o["Ly*"], o["Lz*"]) == (0.0, 0.0, 0.0)
    assert np.isnan(o["vrot"]) and np.isnan(o["sigcyl"])
    assert o["sig3d"] == pytest.approx(50.0 * np.sqrt(6 / 7), rel=1e-12)


def test_dark_matter_only_defaults_contamination_and_dispersions():
    pos = np.zeros((3, 3))
    vel = np.array([[10.0, 0, 0], [-10.0, 0, 0], [0, 0, 0]])
    m = np.array([1.0, 1.0, 2.5])
    o = props(pos, vel, m, [1, 1, 1], np.zeros(3), vbar=(0, 0, 0), massp=1.0)
    assert o["mcontam"] == 2.5 and o["mdm"] == 4.5 and o["m*"] == 0.0
    assert o["sigma_dm"] == pytest.approx(np.sqrt(200 / 4.5), rel=1e-15) and o["sigma*"] == 0
    assert o["SFR"] == 0.0 and o["nstar"] == 0
    for k in ("px*", "Lx*", "r*", "r50", "age", "vrot", "sig3d"):
        assert np.isnan(o[k]), k
    # one star among them: sigma* about the halo's mean velocity
    o = props(pos, vel, m, [1, 1, 1], [False, True, False], vbar=(1.0, 0, 0))
    assert o["sigma*"] == pytest.approx(11.0, rel=1e-15) and o["r50"] == 0.0
    assert (o["Lx*"], o["Ly*"], o["Lz*"]) == (0.0, 0.0, 0.0) and np.isnan(o["sig3d"])


def test_stellar_centre_follows_sc_on_the_stars():
    """The stellar centre is the densest star, then the shrinking sphere over the stars only:
    dark matter elsewhere does not move it."""
    rng = np.random.default_rng(2)
    stars = rng.normal(size=(300, 3)) * 0.002 + 0.1
    dm = rng.normal(size=(300, 3)) * 0.01
    pos = np.vstack([stars, dm])
    star = np.r_[np.ones(300, bool), np.zeros(300, bool)]
    dens = rng.uniform(1, 2, 600)
    kw = dict(sc=True, dcell_min=1e-4, eps_sc=0.3)
    o = props(pos, np.zeros((600, 3)), np.ones(600), dens, star, **kw)
    o2 = props(stars, np.zeros((300, 3)), np.ones(300), dens[:300], np.ones(300), **kw)
    assert (o["px*"], o["py*"], o["pz*"]) == (o2["px*"], o2["py*"], o2["pz*"])
    assert not np.array_equal([o["px*"], o["py*"], o["pz*"]], stars[int(np.argmax(dens[:300]))])


def test_chemistry_means_skip_invalid_values():
    pos = np.array([[0, 0, 0], [0.001, 0, 0], [0.002, 0, 0], [0.003, 0, 0]], float)
    m = np.array([1.0, 3.0, 0.0, 2.0])                 # a massless star is skipped
    chem = np.array([[0.7, 1e-3], [0.8, np.nan], [5.0, 5.0], [0.6, 4e-3]])
    o = props(pos, np.zeros((4, 3)), m, [9, 1, 1, 1], np.ones(4), chem=chem)
    assert o["chem"][0] == pytest.approx((0.7 + 2.4 + 1.2) / 6, rel=1e-15)
    assert o["chem"][1] == pytest.approx((1e-3 + 8e-3) / 3, rel=1e-15)
    o = props(pos, np.zeros((4, 3)), m, [9, 1, 1, 1], np.zeros(4), chem=chem)
    assert np.isnan(o["chem"]).all()                   # no stars
