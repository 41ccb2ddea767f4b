"""amdahop -- MI355X-native AdaptaHOP hot path (kNN, SPH density, hop, saddle edges).

Host-side entry points over libamdahop.so:

* :func:`run_device` -- the hot path on device-resident arrays (torch tensors are used
  only as device allocations; the library sees raw pointers);
* the drop-in ``compute_adaptahop.neikdtree`` module (package ``compute_adaptahop``);
* :mod:`amdahop.generate` -- the synthetic workloads.
"""
from __future__ import annotations

import ctypes
from typing import NamedTuple

from ._lib import HotParams, RunInfo, check, lib

__all__ = ["run_device", "device_saddle", "device_hop", "knn_device", "density_adjoint", "sph_density",
           "halo_props", "HALO_FIELDS", "halo_profiles", "PROFILE_FIELDS", "halo_stellar",
           "STELLAR_FIELDS", "AngMomShells", "write_ang_mom_of_r", "HotParams", "RunInfo"]


def run_device(pos, mass, nvoisins, nhop=16, rho_threshold=80.0, box=1.0, mixed=True,
               bucket=64, saddle=True, density_out=None, gid_out=None):
    """Run kNN + density + hop + group ids (+ saddle edges) on device tensors.

    pos: (N,3) float64 contiguous on the GPU; mass: (N,) float64.  Returns
    (density, gid, info) with density/gid in original particle order (torch tensors on
    the same device) and ``info`` the :class:`RunInfo` of the run."""
    import torch

    if pos.dtype != torch.float64 or mass.dtype != torch.float64:
        raise TypeError("pos and mass must be float64")
    if not (pos.is_cuda and mass.is_cuda):
        raise ValueError("pos and mass must be device tensors")
    pos = pos.contiguous()
    mass = mass.contiguous()
    n = int(mass.shape[0])
    if tuple(pos.shape) != (n, 3):
        raise ValueError("pos must have shape (N,3)")
    Lx, Ly, Lz = (box, box, box) if not isinstance(box, (tuple, list)) else box
    p = HotParams(int(nvoisins), int(nhop), float(rho_threshold), float(Lx), float(Ly), float(Lz),
                  1 if mixed else 0, int(bucket), 1 if saddle else 0)
    dens = density_out if density_out is not None else torch.empty(n, dtype=torch.float64, device=pos.device)
    gid = gid_out if gid_out is not None else torch.empty(n, dtype=torch.int32, device=pos.device)
    _check_out(dens, (n,), torch.float64, pos.device, "density_out")
    _check_out(gid, (n,), torch.int32, pos.device, "gid_out")
    info = RunInfo()
    torch.cuda.synchronize(pos.device)
    check(lib().ahop_run_device(ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()), n,
                                ctypes.byref(p), ctypes.c_void_p(dens.data_ptr()),
                                ctypes.c_void_p(gid.data_ptr()), ctypes.byref(info)))
    return dens, gid, info


def _check_out(t, shape, dtype, device, name):
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype or t.device != device or \
            not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {shape} on {device}")


def _box3(box):
    return (box, box, box) if not isinstance(box, (tuple, list)) else tuple(box)


def _check_particles(pos, mass):
    import torch
    if pos.dtype != torch.float64 or mass.dtype != torch.float64:
        raise TypeError("pos and mass must be float64")
    if not (pos.is_cuda and mass.is_cuda):
        raise ValueError("pos and mass must be device tensors")
    n = int(mass.shape[0])
    if tuple(pos.shape) != (n, 3) or mass.dim() != 1:
        raise ValueError("pos must have shape (N,3) and mass (N,)")
    return pos.contiguous(), mass.contiguous(), n


def knn_device(pos, mass, nvoisins, box=1.0, mixed=True, bucket=64, return_order=False):
    """kNN + SPH density with every neighbour row kept: returns (density (N,) float64,
    nbr (N, nvoisins) int32 original ids in ascending distance, RunInfo), and with
    ``return_order`` also the Morton order (N,) int32 (sorted position -> original index)."""
    import torch

    pos, mass, n = _check_particles(pos, mass)
    Lx, Ly, Lz = _box3(box)
    p = HotParams(int(nvoisins), 0, 0.0, float(Lx), float(Ly), float(Lz), 1 if mixed else 0,
                  int(bucket), 0)
    dens = torch.empty(n, dtype=torch.float64, device=pos.device)
    nbr = torch.empty((n, int(nvoisins)), dtype=torch.int32, device=pos.device)
    order = torch.empty(n, dtype=torch.int32, device=pos.device) if return_order else None
    info = RunInfo()
    torch.cuda.synchronize(pos.device)
    check(lib().ahop_knn_device(ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()),
                                n, ctypes.byref(p), ctypes.c_void_p(dens.data_ptr()),
                                ctypes.c_void_p(nbr.data_ptr()),
                                ctypes.c_void_p(order.data_ptr() if return_order else None),
                                ctypes.byref(info)))
    return (dens, nbr, info, order) if return_order else (dens, nbr, info)


def density_adjoint(pos, mass, nbr, grad_rho, box=1.0, order=None, out=None, no_alloc=False):
    """dLoss/dpos (N,3) and dLoss/dmass (N,) of the SPH density for the upstream gradient
    ``grad_rho`` = dLoss/drho, with the neighbour rows ``nbr`` (N,k) frozen.  ``order`` (N,)
    int32, optional: a permutation of the particles that keeps neighbours close, the Morton
    order from :func:`knn_device` (``return_order=True``); with it the neighbours' terms are
    summed per workgroup in LDS first, and far fewer global atomics are needed.  ``out``:
    optional (grad_pos (N,3), grad_mass (N,)) float64 device tensors to write into.
    ``no_alloc``: call the allocation-free entry point (its scratch is the library's fixed
    device array, N <= 4194304; no timing, so the milliseconds read 0).  Returns
    (grad_pos, grad_mass, device milliseconds)."""
    import torch

    pos, mass, n = _check_particles(pos, mass)
    if nbr.dtype != torch.int32 or nbr.dim() != 2 or nbr.shape[0] != n or not nbr.is_cuda:
        raise ValueError("nbr must be an (N, k) int32 device tensor")
    if order is not None:
        if (order.dtype != torch.int32 or tuple(order.shape) != (n,)
                or order.device != pos.device):
            raise ValueError("order must be an (N,) int32 tensor on the particles' device")
        order = order.contiguous()
    grad_rho = grad_rho.to(torch.float64).contiguous()
    if tuple(grad_rho.shape) != (n,) or grad_rho.device != pos.device:
        raise ValueError("grad_rho must have shape (N,) on the particles' device")
    nbr = nbr.contiguous()
    Lx, Ly, Lz = _box3(box)
    if out is not None:
        gpos, gmass = out
        _check_out(gpos, (n, 3), torch.float64, pos.device, "out[0]")
        _check_out(gmass, (n,), torch.float64, pos.device, "out[1]")
    else:
        gpos = torch.empty((n, 3), dtype=torch.float64, device=pos.device)
        gmass = torch.empty(n, dtype=torch.float64, device=pos.device)
    ms = ctypes.c_float(0.0)
    torch.cuda.synchronize(pos.device)
    if no_alloc:
        check(lib().ahop_density_adjoint_device_noalloc(
            ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()), n,
            int(nbr.shape[1]), ctypes.c_void_p(nbr.data_ptr()),
            ctypes.c_void_p(order.data_ptr() if order is not None else None),
            ctypes.c_void_p(grad_rho.data_ptr()), float(Lx), float(Ly), float(Lz),
            ctypes.c_void_p(gpos.data_ptr()), ctypes.c_void_p(gmass.data_ptr())))
        return gpos, gmass, 0.0
    check(lib().ahop_density_adjoint_device(
        ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(mass.data_ptr()), n, int(nbr.shape[1]),
        ctypes.c_void_p(nbr.data_ptr()),
        ctypes.c_void_p(order.data_ptr() if order is not None else None),
        ctypes.c_void_p(grad_rho.data_ptr()), float(Lx),
        float(Ly), float(Lz), ctypes.c_void_p(gpos.data_ptr()), ctypes.c_void_p(gmass.data_ptr()),
        ctypes.byref(ms)))
    return gpos, gmass, ms.value


def _density_function():
    import torch

    class SPHDensity(torch.autograd.Function):
        """rho = SPH density of omp_compute_density (f90:2816-2838); backward is the
        hand-written adjoint with the forward's neighbour rows frozen."""

        @staticmethod
        def forward(ctx, pos, mass, nvoisins, box, mixed):
            dens, nbr, _, order = knn_device(pos.detach(), mass.detach(), nvoisins, box, mixed,
                                             return_order=True)
            ctx.save_for_backward(pos, mass, nbr, order)
            ctx.box = box
            ctx.mark_non_differentiable(nbr)
            return dens, nbr

        @staticmethod
        def backward(ctx, grad_rho, grad_nbr):
            pos, mass, nbr, order = ctx.saved_tensors
            gpos, gmass, _ = density_adjoint(pos.detach(), mass.detach(), nbr, grad_rho, ctx.box,
                                             order=order)
            return gpos, gmass, None, None, None

    return SPHDensity


_SPHDensity = None


def sph_density(pos, mass, nvoisins, box=1.0, mixed=True):
    """Differentiable SPH density: returns (rho (N,), nbr (N,k) int32).  Gradients with
    respect to ``pos`` (N,3) and ``mass`` (N,) come from the hand-written adjoint kernel."""
    global _SPHDensity
    if _SPHDensity is None:
        _SPHDensity = _density_function()
    return _SPHDensity.apply(pos, mass, int(nvoisins), box, bool(mixed))


def device_saddle(info):
    """Host copy (numpy) of the saddle CSR of the last :func:`run_device` call."""
    import numpy as np

    from ._lib import ptr
    off = np.empty(info.ngroups + 1, dtype=np.int32)
    isad = np.empty(max(info.nsaddle, 1), dtype=np.int32)
    rho = np.empty(max(info.nsaddle, 1), dtype=np.float64)
    check(lib().ahop_device_saddle(ptr(off), ptr(isad), ptr(rho)))
    return off, isad[:info.nsaddle], rho[:info.nsaddle]


def device_hop(n, device=None, out=None):
    """Hop targets of the last :func:`run_device` call as an (n,) int32 device tensor: the
    0-based original index each particle hops to (-1 where inactive)."""
    import torch

    dev = device if device is not None else torch.device("cuda:0")
    hop = out if out is not None else torch.empty(int(n), dtype=torch.int32, device=dev)
    _check_out(hop, (int(n),), torch.int32, hop.device, "out")
    check(lib().ahop_device_hop(ctypes.c_void_p(hop.data_ptr()), int(n)))
    return hop


HALO_FIELDS = ["m", "px", "py", "pz", "vx", "vy", "vz", "sigma", "Lx", "Ly", "Lz", "r",
               "Ixx", "Ixy", "Ixz", "Iyy", "Iyz", "Izz", "sha", "shb", "shc", "ek", "ep", "et",
               "rvir", "mvir", "cvel", "tvir", "spin", "rho_0", "r_c"]


def halo_props(offsets, pos, vel, mass, density=None, *, lbox=1.0, hub_pt=70.3, gravconst=430.1,
               vir_overdens=340.0, rho_mean=1.0, cdm=True, nshell=1000, direct_pe_nmax=1000,
               tile=0, workgroup=0, sc=False, dpmmc=False, dcell_min=0.0, eps_sc=0.0,
               light=False, flag_period=1, virial_pe_profile="shell",
               virial_pe_norm_policy="profile_outer_for_large", virial_shell_self_mode="count",
               ang_mom_of_r=False, nshells=100):
    """Per-halo properties of a CSR member layout (device tensors), one row per halo in
    the order of HALO_FIELDS.  ``sc`` / ``dpmmc`` recentre each halo as the reference's
    SC / DPMMC options do (with ``dcell_min`` and ``eps_sc``; DPMMC needs ``density``).
    ``light`` skips the shell tables, energies and virial scan (those columns read NaN).
    The reference's switches keep its names and values: ``flag_period`` (FlagPeriod; 0 turns
    the periodic wrap off), ``virial_pe_profile`` ('shell' or 'baseline'),
    ``virial_pe_norm_policy`` ('profile_outer_for_large' or 'legacy') and
    ``virial_shell_self_mode`` ('count' or 'half_shell') of potential_profile_1b40
    (compute_halo_props.py:1065-1123); an unknown value is a ValueError as there.
    ``ang_mom_of_r`` (ANG_MOM_OF_R) also bins each halo's members into ``nshells`` (the
    reference's 100) ellipsoidal shells of the all-member ellipsoid and sums mass and
    angular momentum per shell, as det_ang_momentum_per_shell_1b5 (:1970-2013) does for the
    halos whose rvir is finite (:1257); ``write_ang_mom_of_r`` writes its file.
    Returns (out tensor (H, 31) float64, HaloInfo), and with ``ang_mom_of_r`` a third item,
    AngMomShells(shells (H, 4, nshells): per shell m, Lx, Ly, Lz; axes (H, 12): amax, bmax,
    cmax, then the principal axes v row-major), NaN rows for halos with no finite rvir."""
    import numpy as np
    import torch

    modes = {"virial_pe_profile": (virial_pe_profile, ("shell", "baseline")),
             "virial_pe_norm_policy": (virial_pe_norm_policy,
                                       ("profile_outer_for_large", "legacy")),
             "virial_shell_self_mode": (virial_shell_self_mode, ("count", "half_shell"))}
    sw = {}
    for name, (val, allowed) in modes.items():
        if val not in allowed:
            raise ValueError(f"unknown {name}={val!r}")
        sw[name] = allowed.index(val)
    if ang_mom_of_r and light:
        raise ValueError("ang_mom_of_r needs the full mode (light=False)")
    if ang_mom_of_r and not 1 <= int(nshells) <= 2000:
        raise ValueError("nshells must be in [1, 2000]")
    if flag_period == 0:
        lbox = float("inf")       # no periodic wrap anywhere (FlagPeriod = 0)

    from ._lib import HaloInfo, HaloParams, ptr
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nh = off.shape[0] - 1
    for t in (pos, vel, mass) + ((density,) if density is not None else ()):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise TypeError("pos, vel, mass, density must be contiguous float64 device tensors")
    nm = mass.shape[0]
    if pos.shape != (nm, 3) or vel.shape != (nm, 3) or (density is not None and
                                                         density.shape != (nm,)):
        raise ValueError("pos and vel must be (M, 3) and density (M,) for M = len(mass)")
    if off.ndim != 1 or nh < 1 or off[0] < 0 or off[-1] > nm:
        raise ValueError("offsets must be a 1-D CSR array inside [0, len(mass)]")
    out = torch.empty((nh, len(HALO_FIELDS)), dtype=torch.float64, device=pos.device)
    p = HaloParams(float(lbox), float(hub_pt), float(gravconst), float(vir_overdens),
                   float(rho_mean), 1 if cdm else 0, int(nshell), int(direct_pe_nmax), int(tile),
                   int(workgroup), 1 if sc else 0, 1 if dpmmc else 0, float(dcell_min),
                   float(eps_sc), 1 if light else 0, sw["virial_pe_profile"],
                   sw["virial_pe_norm_policy"], sw["virial_shell_self_mode"])
    ang = None
    if ang_mom_of_r:
        ang = AngMomShells(torch.empty((nh, 4, int(nshells)), dtype=torch.float64,
                                       device=pos.device),
                           torch.empty((nh, 12), dtype=torch.float64, device=pos.device))
        p.ang_mom_nshells = int(nshells)
        p.d_ang_mom = ang.shells.data_ptr()
        p.d_ang_axes = ang.axes.data_ptr()
    info = HaloInfo()
    torch.cuda.synchronize(pos.device)
    dptr = ctypes.c_void_p(density.data_ptr()) if density is not None else None
    check(lib().ahop_halo_props_device(nh, ptr(off), ctypes.c_void_p(pos.data_ptr()),
                                       ctypes.c_void_p(vel.data_ptr()),
                                       ctypes.c_void_p(mass.data_ptr()), dptr, ctypes.byref(p),
                                       ctypes.c_void_p(out.data_ptr()), ctypes.byref(info)))
    if ang is not None:
        return out, info, ang
    return out, info


class AngMomShells(NamedTuple):
    """Per-shell mass and angular momentum of halo_props(ang_mom_of_r=True)."""
    shells: object   # (H, 4, nshells) device tensor: m, Lx, Ly, Lz per shell
    axes: object     # (H, 12) device tensor: amax, bmax, cmax, v row-major


def write_ang_mom_of_r(path, ids, out, ang, nb_of_halos, nb_of_subhalos):
    """Write the ANG_MOM_OF_R file of compute_halo_props.py (:477-484, :1257-1258, :2014-2019)
    as a Fortran sequential unformatted file (4-byte record markers): a record
    (nb_of_halos, nb_of_subhalos) and a record (nshells), as int64 (the Python integers the
    reference passes to FortranFile.write_record), then one record per halo with a finite
    rvir, in the order of ``ids``: id (int32), px, py, pz, vx, vy, vz, rvir, mvir, m, r, spin,
    amax, bmax, cmax, v[0, :3], v[1, :3], v[2, :3], m[:nshells], Lx[:nshells], Ly[:nshells],
    Lz[:nshells] (float64).  The reference's own write of the last four (``L[0:nshells].x``
    on a Python list of vectors) raises AttributeError; this writes the per-shell arrays it
    names.  ``ids`` (H), ``out`` (H, 31) and ``ang`` (AngMomShells) as halo_props returns
    them (tensors or arrays).  Returns the number of halo records written."""
    import numpy as np

    def host(t):
        return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)

    ids = np.asarray(host(ids), dtype=np.int32)
    o = host(out)
    sh = host(ang.shells)
    ax = host(ang.axes)
    nsh = sh.shape[2]
    cols = [HALO_FIELDS.index(f) for f in ("px", "py", "pz", "vx", "vy", "vz", "rvir", "mvir",
                                            "m", "r", "spin")]

    def record(f, payload):
        n = np.array([len(payload)], dtype="<u4").tobytes()
        f.write(n + payload + n)

    nrec = 0
    with open(path, "wb") as f:
        record(f, np.array([nb_of_halos, nb_of_subhalos], dtype="<i8").tobytes())
        record(f, np.array([nsh], dtype="<i8").tobytes())
        for h in range(len(ids)):
            if not np.isfinite(o[h, HALO_FIELDS.index("rvir")]):
                continue
            payload = (np.array([ids[h]], dtype="<i4").tobytes()
                       + np.ascontiguousarray(o[h, cols], dtype="<f8").tobytes()
                       + np.ascontiguousarray(ax[h], dtype="<f8").tobytes()
                       + np.ascontiguousarray(sh[h], dtype="<f8").tobytes())
            record(f, payload)
            nrec += 1
    return nrec


PROFILE_FIELDS = ["vmaxcir", "rmaxcir", "cNFW", "cNFWerr", "inslope", "inslopeerr"]


def profile_bin_edges(first, last, nin):
    """Log-spaced bin edges of every halo eligible for the profile fits, computed as the
    reference's np.logspace(np.log10(first), np.log10(last), nbins + 1) with
    nbins = max(8, min(80, int(np.sqrt(nin)))) (compute_halo_props.py:1862-1863), vectorised
    over halos with the same element operations (linspace: i * step + start, last point =
    stop; then np.power(10., y)), so that the end members fall inside or outside the
    histogram exactly as there.  ``first`` / ``last`` / ``nin``: arrays over the eligible
    halos.  Returns (edges, per-halo edge counts)."""
    import numpy as np
    nin = np.asarray(nin, dtype=np.int64)
    nb = np.maximum(8, np.minimum(80, np.sqrt(nin.astype(np.float64)).astype(np.int64)))
    num = nb + 1
    start = np.log10(np.asarray(first, dtype=np.float64))
    stop = np.log10(np.asarray(last, dtype=np.float64))
    step = (stop - start) / (num - 1)
    eoff = np.zeros(len(num) + 1, dtype=np.int64)
    np.cumsum(num, out=eoff[1:])
    i = (np.arange(eoff[-1], dtype=np.int64) - np.repeat(eoff[:-1], num)).astype(np.float64)
    y = i * np.repeat(step, num)
    y += np.repeat(start, num)
    y[eoff[1:] - 1] = stop
    return np.power(10.0, y), num


def halo_profiles(offsets, pos, mass, props, *, lbox=1.0, gravconst=430.1):
    """Extended profiles (compute_extended_profiles_1e, compute_halo_props.py:1819-1926) of
    every halo of a CSR member layout whose members are all dark matter, about the centres
    and rvir of ``props`` (the rows of :func:`halo_props`).  Device tensors in, one row per
    halo in the order of PROFILE_FIELDS out.  Returns (tensor (H, 6) float64, device ms)."""
    import numpy as np
    import torch

    from ._lib import ptr
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nh = off.shape[0] - 1
    for t in (pos, mass, props):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise TypeError("pos, mass, props must be contiguous float64 device tensors")
    nm = mass.shape[0]
    if pos.shape != (nm, 3) or props.shape != (nh, len(HALO_FIELDS)):
        raise ValueError("pos must be (M, 3) and props (nhalos, %d)" % len(HALO_FIELDS))
    if off.ndim != 1 or nh < 1 or off[0] != 0 or off[-1] != nm:
        raise ValueError("offsets must be a 1-D CSR array over all of mass, from 0")
    dev = pos.device
    sdist = torch.empty(nm, dtype=torch.float64, device=dev)
    smass = torch.empty(nm, dtype=torch.float64, device=dev)
    scum = torch.empty(nm, dtype=torch.float64, device=dev)
    aux = torch.empty((nh, 4), dtype=torch.float64, device=dev)
    prof = torch.empty((nh, len(PROFILE_FIELDS)), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    ms1, ms2 = ctypes.c_float(0.0), ctypes.c_float(0.0)
    vp = ctypes.c_void_p
    check(lib().ahop_halo_profile_radii_device(
        nh, ptr(off), vp(pos.data_ptr()), vp(mass.data_ptr()), vp(props.data_ptr()),
        float(lbox), float(gravconst), vp(sdist.data_ptr()), vp(smass.data_ptr()),
        vp(scum.data_ptr()), vp(aux.data_ptr()), vp(prof.data_ptr()), ctypes.byref(ms1)))
    a = aux.cpu().numpy()
    rvir = props[:, HALO_FIELDS.index("rvir")].cpu().numpy()
    nin = a[:, 1].astype(np.int64)
    # the reference's early returns (:1842-1860): fewer than 8 non-zero distances, rvir not
    # > 0, fewer than 8 members inside rvir
    elig = (np.diff(off) - a[:, 0] >= 8) & (rvir > 0) & (nin >= 8)
    edges, num = profile_bin_edges(a[elig, 2], a[elig, 3], nin[elig])
    counts = np.zeros(nh, dtype=np.int64)
    counts[elig] = num
    eoff = np.zeros(nh + 1, dtype=np.int64)
    np.cumsum(counts, out=eoff[1:])
    d_edges = torch.from_numpy(edges).to(dev) if len(edges) else None
    torch.cuda.synchronize(dev)
    check(lib().ahop_halo_profile_fits_device(
        nh, ptr(off), vp(sdist.data_ptr()), vp(smass.data_ptr()), vp(scum.data_ptr()),
        vp(props.data_ptr()), vp(aux.data_ptr()), ptr(eoff), vp(d_edges.data_ptr()) if d_edges is not None else None,
        vp(prof.data_ptr()), ctypes.byref(ms2)))
    return prof, ms1.value + ms2.value


STELLAR_FIELDS = ["ndm", "nstar", "mdm", "m*", "mcontam", "sigma_dm", "sigma*",
                  "px*", "py*", "pz*", "Lx*", "Ly*", "Lz*", "r*", "r50", "r90", "age", "metal",
                  "SFR", "SFR_r50", "SFR_r90", "SFR10", "SFR10_r50", "SFR10_r90",
                  "vrot", "sig3d", "sigcyl", "vrot_r50", "sig3d_r50", "sigcyl_r50",
                  "vrot_r90", "sig3d_r90", "sigcyl_r90"]


CHEM_ELEMENTS = ("H", "O", "Fe", "Mg", "C", "N", "Si", "S", "D")


def halo_stellar(offsets, pos, vel, mass, density, is_star, age, metal, props, *, chem=None,
                 lbox=1.0, gravconst=430.1, massp=0.0, sc=False, dpmmc=False, dcell_min=0.0,
                 eps_sc=0.0):
    """Stellar properties (mode 'full') of every halo of a CSR member layout: device tensors
    pos/vel (M, 3), mass/density/age/metal (M,) float64, is_star (M,) bool or uint8 (the
    reference's indexps >= ndm), props the rows of :func:`halo_props` (the mean velocity).
    The SC / DPMMC options refine the stellar centre as the reference's do.  ``chem`` (M, nchem)
    float64, nchem <= 9: per-member abundances in CHEM_ELEMENTS order; their mass-weighted
    means over each halo's stars are appended as columns "<element>_star".  Returns (tensor
    (H, 33 + nchem) in the order of STELLAR_FIELDS then the chemistry, device ms)."""
    import numpy as np
    import torch

    from ._lib import HaloParams, ptr
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nh = off.shape[0] - 1
    nm = mass.shape[0]
    for t in (pos, vel, mass, density, age, metal, props):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise TypeError("pos, vel, mass, density, age, metal, props must be contiguous "
                            "float64 device tensors")
    if pos.shape != (nm, 3) or vel.shape != (nm, 3) or props.shape != (nh, len(HALO_FIELDS)):
        raise ValueError("pos and vel must be (M, 3) and props (nhalos, %d)" % len(HALO_FIELDS))
    for t in (density, age, metal, is_star):
        if tuple(t.shape) != (nm,):
            raise ValueError("density, age, metal and is_star must be (M,)")
    if off.ndim != 1 or nh < 1 or off[0] != 0 or off[-1] != nm:
        raise ValueError("offsets must be a 1-D CSR array over all of mass, from 0")
    star = is_star.to(torch.uint8).contiguous()
    nchem = 0
    if chem is not None:
        if (chem.dtype != torch.float64 or not chem.is_cuda or not chem.is_contiguous() or
                chem.dim() != 2 or chem.shape[0] != nm or not 1 <= chem.shape[1] <= 9):
            raise ValueError("chem must be a contiguous float64 device tensor (M, nchem <= 9)")
        nchem = int(chem.shape[1])
    out = torch.empty((nh, len(STELLAR_FIELDS)), dtype=torch.float64, device=pos.device)
    chem_out = torch.empty((nh, nchem), dtype=torch.float64, device=pos.device) if nchem else None
    p = HaloParams(float(lbox), 70.3, float(gravconst), 340.0, 1.0, 0, 1000, 1000, 0, 0,
                   1 if sc else 0, 1 if dpmmc else 0, float(dcell_min), float(eps_sc), 1)
    ms = ctypes.c_float(0.0)
    vp = ctypes.c_void_p
    torch.cuda.synchronize(pos.device)
    check(lib().ahop_halo_stellar_device(
        nh, ptr(off), vp(pos.data_ptr()), vp(vel.data_ptr()), vp(mass.data_ptr()),
        vp(density.data_ptr()), vp(star.data_ptr()), vp(age.data_ptr()), vp(metal.data_ptr()),
        vp(chem.data_ptr()) if nchem else None, nchem, vp(props.data_ptr()), ctypes.byref(p),
        float(massp), vp(out.data_ptr()), vp(chem_out.data_ptr()) if nchem else None,
        ctypes.byref(ms)))
    if nchem:
        out = torch.cat([out, chem_out], dim=1)
    return out, ms.value
