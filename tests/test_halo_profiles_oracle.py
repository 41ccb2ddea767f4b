"""Pins of the extended-profile oracle (oracle/halo_profiles.py, restating
compute_extended_profiles_1e, src/compute_halo_props.py:1819-1926) by constructions whose
answers are known in closed form."""
import numpy as np
import pytest

from oracle import halo_profiles as HPF

G = 430.1


def on_axis(radii):
    """Members on the x axis at the given radii from the origin: sqrt(r*r) == r exactly."""
    pos = np.zeros((len(radii), 3))
    pos[:, 0] = radii
    return pos


def round_trip(x):
    """A value near x that np.logspace reproduces as an end point (10**log10(x) == x), so that
    the members at the first and last distance stay inside the histogram."""
    for k in range(1000):
        y = x * (1 + k * 1e-6)
        if np.power(10.0, np.log10(y)) == y:
            return y
    raise AssertionError("no round-tripping value")


def test_circular_velocity_of_a_uniform_sphere_peaks_at_its_edge():
    # cumulative count i at r_i = R (i/n)^(1/3): vcir ~ i^(1/3), largest at the last member
    n, R = 500, 0.05
    r = R * (np.arange(1, n + 1) / n) ** (1.0 / 3.0)
    m = np.full(n, 2e-3)
    o = HPF.extended_profiles(on_axis(r), m, np.zeros(3), np.nan, 1.0, G)
    assert o["rmaxcir"] == r[-1]
    assert o["vmaxcir"] == pytest.approx(np.sqrt(G * m.sum() / R), rel=1e-14)
    assert np.isnan(o["cNFW"]) and np.isnan(o["inslope"])          # rvir NaN: no fits


def test_circular_velocity_of_a_point_mass_peaks_at_it_and_skips_the_centre():
    r = np.concatenate([[0.0, 1e-4], np.linspace(0.01, 0.05, 40)])
    m = np.concatenate([[1e3, 1.0], np.full(40, 1e-6)])            # member at the centre: skipped
    o = HPF.extended_profiles(on_axis(r), m, np.zeros(3), np.nan, 1.0, G)
    assert o["rmaxcir"] == 1e-4
    assert o["vmaxcir"] == np.sqrt(G * 1.0 / 1e-4)


def test_nfw_fit_returns_the_generating_scale_radius():
    """Ten log bins of ten members whose bin masses make the binned density exactly
    NFW(rs=0.02, log10 rho0=3) at the bin centres: the fit returns rs, cNFW = 5, and a
    vanishing error."""
    rvir, rs, a = 0.1, 0.02, 3.0
    d0, d1 = round_trip(1e-3), round_trip(0.09)
    edges = np.logspace(np.log10(d0), np.log10(d1), 11)
    assert edges[0] == d0 and edges[-1] == d1
    radii, masses = [], []
    for j in range(10):
        lo, hi = edges[j], edges[j + 1]
        rr = lo * (hi / lo) ** ((np.arange(10) + 0.5) / 10)
        if j == 0:
            rr[0] = d0
        if j == 9:
            rr[-1] = d1
        vol = 4.0 / 3.0 * np.pi * (hi ** 3 - lo ** 3)
        rho = 10.0 ** HPF.log_nfw(0.5 * (hi + lo), rs, a)
        radii.append(rr)
        masses.append(np.full(10, rho * vol / 10))
    o = HPF.extended_profiles(on_axis(np.concatenate(radii)), np.concatenate(masses),
                              np.zeros(3), rvir, 1.0, G)
    assert o["cNFW"] == pytest.approx(rvir / rs, rel=1e-6)
    assert o["cNFWerr"] < 1e-6 * o["cNFW"]


def test_inner_slope_of_a_power_law():
    """Members whose per-member shell densities follow r^-1.5 exactly; all but two within 2%
    of the innermost, so fewer than four log bins are occupied, the NFW fit is skipped and the
    slope is fitted inside rvir: inslope = -1.5, error ~0."""
    d0 = 1e-3
    r = np.concatenate([d0 * (1 + 0.02 * np.arange(30) / 30), [0.09, 0.1]])
    vol = np.diff(np.insert(4.0 / 3.0 * np.pi * r ** 3, 0, 0.0))
    m = 7.0 * r ** -1.5 * vol
    o = HPF.extended_profiles(on_axis(r), m, np.zeros(3), 0.2, 1.0, G)
    assert np.isnan(o["cNFW"])
    assert o["inslope"] == pytest.approx(-1.5, rel=1e-9)
    assert o["inslopeerr"] < 1e-9


def test_early_returns():
    r = np.linspace(0.01, 0.05, 7)                                 # 7 members: below 8
    o = HPF.extended_profiles(on_axis(r), np.ones(7), np.zeros(3), 0.1, 1.0, G)
    assert np.isfinite(o["vmaxcir"]) and np.isnan(o["cNFW"]) and np.isnan(o["inslope"])
    r = np.linspace(0.01, 0.05, 20)                                # 7 inside rvir
    o = HPF.extended_profiles(on_axis(r), np.ones(20), np.zeros(3), 0.0225, 1.0, G)
    assert np.isnan(o["cNFW"]) and np.isnan(o["inslope"])


def test_vectorised_bin_edges_equal_np_logspace_per_halo():
    """The product's edges (amdahop.profile_bin_edges) are the reference's np.logspace call
    per halo, bit for bit, including halos whose first and last distances are equal."""
    from amdahop import profile_bin_edges
    rng = np.random.default_rng(0)
    n = 500
    first = 10 ** rng.uniform(-5, -2, n)
    last = first * 10 ** rng.uniform(0, 3, n)
    last[:5] = first[:5]
    nin = rng.integers(8, 100000, n)
    edges, num = profile_bin_edges(first, last, nin)
    off = np.concatenate([[0], np.cumsum(num)])
    for h in range(n):
        nb = max(8, min(80, int(np.sqrt(nin[h]))))
        assert num[h] == nb + 1
        ref = np.logspace(np.log10(first[h]), np.log10(last[h]), nb + 1)
        assert np.array_equal(ref, edges[off[h]:off[h + 1]])
