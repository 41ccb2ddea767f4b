"""FP64 NumPy restatement of the SPH density with frozen neighbour rows and of its
gradient -- TEST INFRASTRUCTURE ONLY.

The density follows reference fortran/compute_adaptahop.f90 omp_compute_density
:2816-2838 (h = sqrt(d2_k)/2, sequential sum over the first k-1 neighbours of
m_j W(sqrt(d2_j)/h), rho = L^3 (sum + m_i) / (pi h^3) with the single-rounded pi of :58),
spline :3270-3287 and the distance of walk_tree2 :2931-2941, in the same operation order,
so that with the reference's own neighbour rows it reproduces the reference density bit
for bit.

The gradient is the analytic derivative of that density with the rows frozen (the
reference has no gradient).  Pinning (tests/test_adjoint_oracle.py): the density equals
the C oracle's (itself bit-exact against the compiled reference) on the same rows; the
gradient equals central finite differences of the density on every coordinate and mass
of small systems, including pairs across the periodic edge.
"""
from __future__ import annotations

import numpy as np

PI_REF = 3.1415927410125732421875


def _box(box):
    return np.array((box, box, box) if np.isscalar(box) else box, dtype=np.float64)


def d2_ref(a, b, L):
    """walk_tree2 :2931-2941: per axis min(|d|, L - |d|), summed ((x + y) + z)."""
    dx = np.abs(b[..., 0] - a[..., 0])
    dx = np.minimum(dx, L[0] - dx)
    d = dx * dx
    dy = np.abs(b[..., 1] - a[..., 1])
    dy = np.minimum(dy, L[1] - dy)
    d = d + dy * dy
    dz = np.abs(b[..., 2] - a[..., 2])
    dz = np.minimum(dz, L[2] - dz)
    return d + dz * dz


def spline(x):
    """spline :3270-3287."""
    out = np.zeros_like(x)
    a = x <= 1.0
    x2 = x[a] * x[a]
    out[a] = 1.0 - 1.5 * x2 + 0.75 * (x2 * x[a])
    b = (x > 1.0) & (x <= 2.0)
    t = 2.0 - x[b]
    out[b] = 0.25 * (t * t * t)
    return out


def density_frozen(pos, mass, nbr, box=1.0, idx=None):
    """Density of particles ``idx`` (default all) from their frozen rows nbr (N, k)."""
    L = _box(box)
    idx = np.arange(pos.shape[0]) if idx is None else np.asarray(idx)
    k = nbr.shape[1]
    P = pos[idx]
    d2k = d2_ref(P, pos[nbr[idx, k - 1]], L)
    r = np.sqrt(d2k) * 0.5
    r3 = r * r * r
    unsr = 1.0 / r
    contrib = np.zeros(idx.shape[0])
    for t in range(k - 1):
        j = nbr[idx, t]
        contrib = contrib + mass[j] * spline(np.sqrt(d2_ref(P, pos[j], L)) * unsr)
    return L.prod() * (contrib + mass[idx]) / (PI_REF * r3)


def _sep(pos, i, j, L):
    """Minimum-image separation pos[j] - pos[i]."""
    s = pos[j] - pos[i]
    s = np.where(s > L / 2, s - L, s)
    return np.where(s < -L / 2, s + L, s)


def _spline_d(q):
    w = np.zeros_like(q)
    wp = np.zeros_like(q)
    a = q <= 1.0
    w[a] = 1.0 - 1.5 * q[a] ** 2 + 0.75 * q[a] ** 3
    wp[a] = -3.0 * q[a] + 2.25 * q[a] ** 2
    b = (q > 1.0) & (q <= 2.0)
    t = 2.0 - q[b]
    w[b] = 0.25 * t ** 3
    wp[b] = -0.75 * t ** 2
    return w, wp


def density_grad(pos, mass, nbr, grad_rho, box=1.0):
    """(dLoss/dpos (N,3), dLoss/dmass (N,)) for dLoss/drho = grad_rho, rows frozen."""
    L = _box(box)
    n, k = nbr.shape
    A = L.prod() / PI_REF
    i = np.arange(n)
    jk = nbr[:, k - 1]
    sk = _sep(pos, i, jk, L)
    dk = np.sqrt(np.sum(sk * sk, axis=1))
    r = 0.5 * dk
    gA3 = grad_rho * A / r ** 3
    gpos = np.zeros((n, 3))
    gmass = np.zeros(n)
    own = np.zeros((n, 3))
    C = np.zeros(n)
    S = np.zeros(n)
    for t in range(k - 1):
        j = nbr[:, t]
        s = _sep(pos, i, j, L)
        d = np.sqrt(np.sum(s * s, axis=1))
        w, wp = _spline_d(d / r)
        C += mass[j] * w
        S += mass[j] * wp * d
        np.add.at(gmass, j, gA3 * w)
        c = np.where(d > 0, gA3 / r * mass[j] * wp / np.where(d > 0, d, 1.0), 0.0)
        f = c[:, None] * s
        own -= f
        np.add.at(gpos, j, f)
    gmass += gA3
    drho_dr = -A / r ** 4 * (3.0 * (C + mass) + S / r)
    f = (grad_rho * drho_dr * 0.5 / dk)[:, None] * sk
    own -= f
    np.add.at(gpos, jk, f)
    return gpos + own, gmass


def affected(nbr, p):
    """Particles whose density depends on particle p with the rows frozen: p itself and
    every particle that lists p among its k neighbours."""
    return np.union1d([p], np.nonzero(np.any(nbr == p, axis=1))[0])
