"""Stellar properties on the GPU (ahop_halo_stellar_device) vs oracle/halo_stellar.py.

Bars: counts, the stellar centre (a star's position), r*, r50 and r90 exact (r50 / r90 come
from the same left-to-right cumulative sum in distance order as np.cumsum); masses, age,
metallicity and star formation rates within 1e-12; L* within 1e-10 of |L*|; the kinematics
within 1e-10 of the halo's sig3d (the reference sums pairwise in member order)."""
import numpy as np
import pytest

from amdahop import generate as G

pytestmark = pytest.mark.gpu


def with_stars(hs, seed, frac=0.2, starless=(3, 7)):
    """Members of each halo flagged as stars: the innermost fraction by distance from the
    halo's first member; some halos without stars.  Ages straddle the 10 and 100 Myr cuts."""
    rng = np.random.default_rng(seed)
    n = len(hs.mass)
    star = np.zeros(n, dtype=bool)
    for h in range(hs.nhalos):
        s, e = hs.offsets[h], hs.offsets[h + 1]
        if h in starless:
            continue
        d = np.linalg.norm((hs.pos[s:e] - hs.pos[s] + 0.5) % 1.0 - 0.5, axis=1)
        k = max(1, int(frac * (e - s)))
        star[s + np.argsort(d)[:k]] = True
    age = 10 ** rng.uniform(-3, 1, n)
    metal = rng.uniform(0, 0.04, n)
    dens = rng.uniform(1, 100, n)
    return star, age, metal, dens


def run(hs, gpu, star, age, metal, dens, chem=None, **kw):
    import torch
    from amdahop import halo_props, halo_stellar
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    props, _ = halo_props(hs.offsets, t(hs.pos), t(hs.vel), t(hs.mass), hub_pt=70.3)
    out, ms = halo_stellar(hs.offsets, t(hs.pos), t(hs.vel), t(hs.mass), t(dens), t(star),
                           t(age), t(metal), props, chem=None if chem is None else t(chem),
                           massp=float(np.median(hs.mass)), **kw)
    return props.cpu().numpy(), out.cpu().numpy(), ms


def compare(hs, props, out, star, age, metal, dens, **kw):
    from oracle.halo_stellar import STELLAR_FIELDS, all_halos
    ref = all_halos(hs.offsets, hs.pos, hs.vel, hs.mass, dens, star, age, metal, props[:, 4:7],
                    massp=float(np.median(hs.mass)), **kw)
    F = {k: i for i, k in enumerate(STELLAR_FIELDS)}
    assert np.array_equal(np.isnan(out), np.isnan(ref)), \
        [STELLAR_FIELDS[j] for j in np.nonzero(np.any(np.isnan(out) != np.isnan(ref), axis=0))[0]]
    for k in ("ndm", "nstar", "px*", "py*", "pz*", "r*", "r50", "r90"):
        assert np.array_equal(out[:, F[k]], ref[:, F[k]], equal_nan=True), k
    ok = ~np.isnan(ref)
    for k in ("mdm", "m*", "mcontam", "sigma_dm", "sigma*", "age", "metal", "SFR", "SFR_r50",
              "SFR_r90", "SFR10", "SFR10_r50", "SFR10_r90"):
        j = F[k]
        d = np.abs(out[ok[:, j], j] - ref[ok[:, j], j])
        assert np.all(d <= 1e-12 * np.maximum(np.abs(ref[ok[:, j], j]), 1e-300)), (k, d.max())
    Ln = np.sqrt(np.nansum(ref[:, [F["Lx*"], F["Ly*"], F["Lz*"]]] ** 2, axis=1))
    for k in ("Lx*", "Ly*", "Lz*"):
        j = F[k]
        d = np.abs(out[ok[:, j], j] - ref[ok[:, j], j])
        assert np.all(d <= 1e-10 * Ln[ok[:, j]]), (k, d.max())
    for sfx in ("", "_r50", "_r90"):
        scale = ref[:, F["sig3d" + sfx]]
        for k in ("vrot", "sig3d", "sigcyl"):
            j = F[k + sfx]
            d = np.abs(out[ok[:, j], j] - ref[ok[:, j], j])
            assert np.all(d <= 1e-10 * scale[ok[:, j]]), (k + sfx, d.max())
    return ref


def test_stellar_props_match_the_oracle(gpu):
    hs = G.nfw_halos(120, 13, alpha=0.9, nmin=30, nmax=20000, mass_sigma_dex=0.3)
    star, age, metal, dens = with_stars(hs, 1)
    print(f"{hs.nhalos} halos, {int(star.sum())} stars of {len(star)} members", flush=True)
    props, out, ms = run(hs, gpu, star, age, metal, dens)
    print(f"stellar: {ms:.3f} ms", flush=True)
    ref = compare(hs, props, out, star, age, metal, dens)
    assert np.isnan(ref[3, 7]) and ref[3, 18] == 0.0      # a halo without stars
    assert np.nansum(ref[:, 18]) > 0 and np.nansum(ref[:, 21]) > 0   # some recent star formation


def test_stellar_centre_with_sc_and_dpmmc(gpu):
    hs = G.nfw_halos(60, 14, alpha=0.8, nmin=100, nmax=10000, mass_sigma_dex=0.3)
    star, age, metal, dens = with_stars(hs, 2, frac=0.4)
    # DPMMC indexes the last grid's stars with the densest star's index: keep it among the first
    for h in range(hs.nhalos):
        s, e = hs.offsets[h], hs.offsets[h + 1]
        idx = s + np.nonzero(star[s:e])[0]
        if len(idx):
            dens[idx[0]] = 1e3
    for kw in (dict(sc=True, dcell_min=1e-4, eps_sc=0.3),
               dict(dpmmc=True, sc=True, dcell_min=5e-4, eps_sc=0.3)):
        print(kw, flush=True)
        props, out, _ = run(hs, gpu, star, age, metal, dens, **kw)
        compare(hs, props, out, star, age, metal, dens, **kw)


def test_all_dark_matter(gpu):
    hs = G.nfw_halos(20, 15, alpha=1.0, nmin=30, nmax=3000)
    n = len(hs.mass)
    z = np.zeros(n)
    props, out, _ = run(hs, gpu, np.zeros(n, dtype=bool), z, z, np.ones(n))
    compare(hs, props, out, np.zeros(n, dtype=bool), z, z, np.ones(n))


def test_stellar_chemistry(gpu):
    """Mass-weighted abundances over each halo's stars, NaN values and massless stars skipped
    (:1671-1679); within 1e-12 of the oracle, NaN where no star has a valid value."""
    from oracle.halo_stellar import STELLAR_FIELDS, all_halos
    hs = G.nfw_halos(40, 16, alpha=0.9, nmin=30, nmax=5000, mass_sigma_dex=0.3)
    star, age, metal, dens = with_stars(hs, 3)
    rng = np.random.default_rng(4)
    chem = rng.uniform(0, 1, (len(hs.mass), 9))
    chem[rng.random(chem.shape) < 0.05] = np.nan
    chem[:, 8] = np.nan                                 # an element with no valid value
    props, out, _ = run(hs, gpu, star, age, metal, dens, chem=chem)
    ref = all_halos(hs.offsets, hs.pos, hs.vel, hs.mass, dens, star, age, metal, props[:, 4:7],
                    chem=chem, massp=float(np.median(hs.mass)))
    n = len(STELLAR_FIELDS)
    assert out.shape == (hs.nhalos, n + 9)
    assert np.array_equal(np.isnan(out[:, n:]), np.isnan(ref[:, n:]))
    ok = ~np.isnan(ref[:, n:])
    assert np.all(np.abs(out[:, n:][ok] - ref[:, n:][ok]) <= 1e-12 * np.abs(ref[:, n:][ok]))
    assert np.isnan(out[:, -1]).all() and np.isnan(out[3, n:]).all()
