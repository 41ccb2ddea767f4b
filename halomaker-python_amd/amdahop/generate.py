"""Deterministic synthetic particle sets for the AdaptaHOP hot path.

The reference reads RAMSES snapshots (src/input_output.py:624-760): positions
arrive in [0,1) box units and are shifted by -0.5 (input_output.py:734), dark
matter comes first and stars after it (input_output.py:716-717), and the total
mass is normalised to 1 so that the mean density is 1 and ``rhot`` = 80 means
80x mean (compute_adaptahop.f90:2835).  The sets below reproduce those
conventions with clustered Plummer clumps on a uniform background.

Every array is a pure function of (config, seed) through
``numpy.random.default_rng``; ``digest`` gives a SHA-256 over the arrays so a
NumPy stream change is detected by the tests.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np

__all__ = ["ParticleSet", "clumpy_box", "mixed_family_box", "config", "digest",
           "CONFIGS", "HaloSet", "nfw_halos", "halo_config"]


@dataclass
class ParticleSet:
    pos: np.ndarray        # (N, 3) float64, C order, in [-L/2, L/2)
    mass: np.ndarray       # (N,) float64, sums to 1
    nvoisins: int
    nhop: int
    name: str
    n_dm: int

    @property
    def n(self) -> int:
        return int(self.mass.shape[0])

    @property
    def pos_f(self) -> np.ndarray:
        """(3, N) Fortran-ordered view, as the reference passes ``pos.T``."""
        return np.asfortranarray(self.pos.T)


def _pareto_counts(rng, nclumps, total, alpha, nmin):
    """Member counts: classic Pareto(alpha, x_m=nmin), rescaled to sum to total.

    A raw Pareto sample with x_m=64 sums to far less than the clumped fraction
    of N, so the counts are scaled up proportionally and the rounding residue
    is given to the largest clumps, which keeps the sum exact."""
    raw = nmin * (1.0 + rng.pareto(alpha, size=nclumps))
    scaled = raw * (total / raw.sum())
    counts = np.floor(scaled).astype(np.int64)
    counts = np.maximum(counts, 1)
    resid = total - int(counts.sum())
    order = np.argsort(-scaled, kind="stable")
    i = 0
    while resid != 0:
        j = order[i % nclumps]
        step = 1 if resid > 0 else -1
        if counts[j] + step >= 1:
            counts[j] += step
            resid -= step
        i += 1
    return counts


def _plummer_offsets(rng, n, a):
    """n offsets from a Plummer sphere of scale radius a, truncated at 10a.

    Enclosed-mass fraction M(<r)/M = r^3/(r^2+a^2)^{3/2}; inverse r = a/sqrt(u^{-2/3}-1).
    u is drawn below M(<10a)/M so the truncation is exact (no clipping)."""
    umax = 1000.0 / 101.0 ** 1.5
    u = rng.uniform(0.0, umax, size=n)
    u = np.maximum(u, 1e-300)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * r[:, None]


def _wrap_shift(p):
    """Wrap into [0,1) then shift by -0.5 as the reader does."""
    p = np.mod(p, 1.0)
    p[p >= 1.0] = 0.0
    return p - 0.5


def clumpy_box(n, nclumps, seed, nvoisins=32, nhop=16, clump_frac=0.6, alpha=1.8,
               nmin=64, amin=0.002, amax=0.01, name="clumpy"):
    """Equal-mass particles, ``clump_frac`` of them in Plummer clumps."""
    rng = np.random.default_rng(seed)
    nclump_tot = int(round(clump_frac * n))
    counts = _pareto_counts(rng, nclumps, nclump_tot, alpha, nmin)
    centres = rng.uniform(0.0, 1.0, size=(nclumps, 3))
    a = np.exp(rng.uniform(np.log(amin), np.log(amax), size=nclumps))
    pos = np.empty((n, 3), dtype=np.float64)
    k = 0
    for c in range(nclumps):
        m = int(counts[c])
        pos[k:k + m] = centres[c] + _plummer_offsets(rng, m, a[c])
        k += m
    pos[k:] = rng.uniform(0.0, 1.0, size=(n - k, 3))
    pos = _wrap_shift(pos)
    mass = np.full(n, 1.0 / n, dtype=np.float64)
    return ParticleSet(np.ascontiguousarray(pos), mass, nvoisins, nhop, name, n)


def mixed_family_box(n, nclumps, seed, nvoisins=64, nhop=16, clump_frac=0.6, star_frac=0.25,
                     alpha=1.8, nmin=64, amin=0.002, amax=0.01, star_sigma_dex=0.5,
                     name="mixed"):
    """Dark matter (equal mass) first, then stars (log-normal masses).

    Stars live only in clumps, with half the parent scale radius; the rest of
    the clumped fraction and the whole background are dark matter.  Star
    masses have median equal to the dark-matter particle mass before the
    global renormalisation to total mass 1."""
    rng = np.random.default_rng(seed)
    nclump_tot = int(round(clump_frac * n))
    n_star = int(round(star_frac * n))
    n_dm = n - n_star
    if n_star > nclump_tot:
        raise ValueError("stars must fit in the clumped fraction")
    counts = _pareto_counts(rng, nclumps, nclump_tot, alpha, nmin)
    centres = rng.uniform(0.0, 1.0, size=(nclumps, 3))
    a = np.exp(rng.uniform(np.log(amin), np.log(amax), size=nclumps))
    # split each clump's members between stars and dark matter, keeping the total exact
    sfrac = n_star / nclump_tot
    sc = np.floor(counts * sfrac).astype(np.int64)
    resid = n_star - int(sc.sum())
    order = np.argsort(-(counts * sfrac - sc), kind="stable")
    sc[order[:resid]] += 1
    dc = counts - sc
    pos_dm = np.empty((n_dm, 3), dtype=np.float64)
    pos_st = np.empty((n_star, 3), dtype=np.float64)
    kd = ks = 0
    for c in range(nclumps):
        md, ms = int(dc[c]), int(sc[c])
        pos_dm[kd:kd + md] = centres[c] + _plummer_offsets(rng, md, a[c])
        pos_st[ks:ks + ms] = centres[c] + _plummer_offsets(rng, ms, 0.5 * a[c])
        kd += md
        ks += ms
    pos_dm[kd:] = rng.uniform(0.0, 1.0, size=(n_dm - kd, 3))
    pos = _wrap_shift(np.concatenate([pos_dm, pos_st]))
    m_dm = 1.0 / n
    m_st = m_dm * 10.0 ** rng.normal(0.0, star_sigma_dex, size=n_star)
    mass = np.concatenate([np.full(n_dm, m_dm), m_st])
    mass /= mass.sum()
    return ParticleSet(np.ascontiguousarray(pos), mass, nvoisins, nhop, name, n_dm)


# The workloads the hot path is measured and tested on.
CONFIGS = {
    "c1": dict(kind="clumpy", n=262_144, nclumps=150, seed=0, nvoisins=32),
    "c2": dict(kind="clumpy", n=2_097_152, nclumps=1_200, seed=1, nvoisins=32),
    "c3": dict(kind="mixed", n=16_777_216, nclumps=9_000, seed=2, nvoisins=64),
    "c5": dict(kind="clumpy", n=2_097_152, nclumps=1_200, seed=1, nvoisins=32),
}


def config(name, n=None, nclumps=None, seed=None):
    """Build a named workload; ``n``/``nclumps``/``seed`` override for reduced cases."""
    c = dict(CONFIGS[name])
    if n is not None:
        scale = n / c["n"]
        c["nclumps"] = max(1, int(round(c["nclumps"] * scale))) if nclumps is None else nclumps
        c["n"] = n
    elif nclumps is not None:
        c["nclumps"] = nclumps
    if seed is not None:
        c["seed"] = seed
    kind = c.pop("kind")
    fn = clumpy_box if kind == "clumpy" else mixed_family_box
    return fn(c["n"], c["nclumps"], c["seed"], nvoisins=c["nvoisins"], name=name)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


# ----------------------------------------------------------------------------
# Batched halo-property workload: NFW-like spheres of very uneven size.
# ----------------------------------------------------------------------------
@dataclass
class HaloSet:
    offsets: np.ndarray    # (H+1,) int64 CSR offsets into the member arrays
    pos: np.ndarray        # (M, 3) float64 physical positions (box units, L=1)
    vel: np.ndarray        # (M, 3) float64
    mass: np.ndarray       # (M,) float64
    name: str

    @property
    def nhalos(self) -> int:
        return int(self.offsets.shape[0] - 1)


def _nfw_x(y):
    """x >= 0 with ln(1+x) - x/(1+x) = y: the NFW enclosed-mass profile inverted through
    the principal branch of Lambert's W, 1 + x = -1 / W0(-exp(-1-y))."""
    from scipy.special import lambertw
    w = lambertw(-np.exp(-1.0 - y), 0).real
    return np.maximum(-1.0 / w - 1.0, 0.0)


def nfw_halos(nhalos, seed, alpha=1.9, nmin=20, nmax=2_000_000, total=None,
              mass_sigma_dex=0.0, gravconst=430.1, name="halos"):
    """``nhalos`` NFW spheres in a unit periodic box.

    Member counts are Pareto(alpha) with minimum ``nmin``, scaled to sum to about
    ``total`` when it is given, then truncated to [nmin, nmax].  Concentrations are
    logU[4, 15]; radii follow the NFW enclosed-mass profile up to R = 0.01 (n/1000)^(1/3);
    centres are U[0,1)^3, positions wrapped and shifted by -0.5; velocities are isotropic
    Gaussians with sigma^2 = G M / (2 R); masses are equal and sum to 1 (log-normal with
    ``mass_sigma_dex`` when it is positive).  Vectorised over all members.
    """
    rng = np.random.default_rng(seed)
    raw = nmin * (1.0 + rng.pareto(alpha, size=nhalos))
    if total is not None:
        raw *= float(total) / raw.sum()
    counts = np.clip(np.floor(raw).astype(np.int64), nmin, nmax)
    offsets = np.zeros(nhalos + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    mtot = int(offsets[-1])
    conc = np.exp(rng.uniform(np.log(4.0), np.log(15.0), size=nhalos))
    centres = rng.uniform(0.0, 1.0, size=(nhalos, 3))
    rvir = 0.01 * (counts / 1000.0) ** (1.0 / 3.0)
    mass = np.full(mtot, 1.0 / mtot)
    if mass_sigma_dex > 0:
        mass *= 10.0 ** rng.normal(0.0, mass_sigma_dex, size=mtot)
    hid = np.repeat(np.arange(nhalos, dtype=np.int32), counts)
    mu_c = np.log1p(conc) - conc / (1.0 + conc)
    r = _nfw_x(rng.uniform(0.0, 1.0, size=mtot) * mu_c[hid]) * (rvir / conc)[hid]
    pos = rng.normal(size=(mtot, 3))
    pos *= (r / np.linalg.norm(pos, axis=1))[:, None]
    del r
    pos += centres[hid]
    np.mod(pos, 1.0, out=pos)
    pos -= 0.5
    pos[pos >= 0.5] -= 1.0
    sig = np.sqrt(gravconst * np.add.reduceat(mass, offsets[:-1]) / (2.0 * rvir))
    vel = rng.normal(size=(mtot, 3))
    vel *= sig[hid][:, None]
    return HaloSet(offsets, pos, vel, mass, name)


def halo_config(nhalos=50_000, total=30_000_000, seed=4):
    """The batched halo-property workload: 50 000 NFW halos, about 30 M members."""
    return nfw_halos(nhalos, seed, total=total, name=f"nfw{nhalos}")
