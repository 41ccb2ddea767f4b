#!/usr/bin/env python3
"""Throughput of the AdaptaHOP hot path (kNN + SPH density + hop/group ids) on one MI355X.

One step = one pass of the hot path over the resident 16,777,216-particle mixed-family
box (k=64, nhop=16): Morton sort, device tree, kNN + density, hop targets, group ids.
Inputs are in HBM before the timed region; outputs (density, group ids) stay in HBM.
Prints one JSON line.  Also reports, in the same line, the plain FP64 device path, the
saddle-edge stage, the roofline of the dominant kernel (kNN + density) and the
reference's own OpenMP Fortran path timed on this host's cores on a bounded sample, and
the stages beside the path: particles grouped by structure, the batched halo properties and
the density adjoint.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "halomaker-python_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mparticles/s kNN+density+hop at 16.8M; mixed vs FP64 GPU at ≤1e-6 rel ρ err"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
# FP64 operations (FMA as 2) of one pair of the direct potential: 3 differences, the squared
# distance (a multiply and 2 FMAs), 1/d (v_rsq_f64 and one Newton step of 2 multiplies and
# 2 FMAs: 7), the mass-weighted accumulation (an FMA)
PE_FLOPS_PER_PAIR = 17
FP32_VALU_PEAK_TFLOPS = 157.3
# algorithmic HBM bytes per particle of the kNN+density kernel: read position 24 B and
# mass 8 B, write density 8 B (the whole step adds the hop stage's density re-read 8 B and
# group-id write 8 B: 56 B per particle)
KNN_BYTES_PER_PARTICLE = 40
STEP_BYTES_PER_PARTICLE = 56
# per member of the batched halo properties: position 24 B, velocity 24 B, mass 8 B and a
# member index 4 B read once
HALO_BYTES_PER_MEMBER = 60
# per particle of the density adjoint (k=32): position 24, mass 8, upstream gradient 8,
# density 8, smoothing length 8, neighbour ids 4k = 128, position gradient 24, mass gradient 8
ADJOINT_BYTES_PER_PARTICLE = 216
# per particle of the grouping by structure: node id 4 B read, label 4 B and grouped particle
# index 4 B written
GROUP_BYTES_PER_PARTICLE = 12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpus():
    """The host's CPUs as this process sees them: threads for the CPU legs (OMP_NUM_THREADS
    when set, which is the CPU share a GPU job is given, else every CPU of the affinity mask),
    nproc, the affinity mask size and the CPU model."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": omp or aff, "nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads": omp or None, "cpu_model": model}


def cpu_baseline(ps, sample_n, nthreads, repeats, full_stages):
    """The reference's OpenMP Fortran (oracle/_ref, compiled from the reference source) on
    this host's cores: its tree + density + local-maxima stages (f90:2967, :682, :748), the
    reference's kNN+density+hop path as shipped, on the headline box itself (sample_n = N) or
    on the same generator at sample_n particles; best of `repeats`.  With full_stages, also
    one run of the whole public entry with the reference's own stage prints, with and without
    its optimize_nodes create-nodes path."""
    import oracle as O
    from amdahop import generate as G
    cpus = host_cpus()
    if sample_n != ps.n:
        ps = G.config("c3", n=sample_n)
    if O.ref_available():
        R = O.Reference(ps.pos, ps.mass, ps.nvoisins, nthreads=nthreads)
        ts = []
        for _ in range(repeats):
            ts.append(R.time_knn_density_hop())
            log(f"reference tree + density + local maxima on {ps.n} particles: {ts[-1]:.2f} s")
        kind = "reference"
    else:
        ts = []
        for _ in range(repeats):
            o = O.Oracle(ps.pos, ps.mass, ps.nvoisins, nthreads=nthreads)
            t0 = time.perf_counter()
            rho, _ = o.density()
            o.maxima(rho)
            ts.append(time.perf_counter() - t0)
            o.close()
        kind = "port"
    t = min(ts)
    what = ("the headline box itself" if ps.n == sample_n
            else f"the same generator at N={sample_n} (clumps scaled)")
    out = {"value": round(ps.n / t / 1e6, 4), "unit": "Mparticles/s", "cores": nthreads,
           "kind": kind,
           "sample": f"{what}, k=64, nhop=16: the reference's tree + density + local-maxima "
                     f"stages (its kNN+density+hop as shipped), best of {repeats} "
                     f"({', '.join(f'{x:.2f}' for x in ts)} s), OpenMP on {nthreads} threads",
           "host": cpus}
    if full_stages and kind == "reference":
        out["reference_stages_s"] = {}
        for opt in (False, True):
            st = R.timed_full(optimize_nodes=opt)
            out["reference_stages_s"]["optimize_nodes" if opt else "default"] = st
            log(f"reference full run (optimize_nodes={opt}): {st}")
    return out


def halo_props_stage(steps, cpu_sample_halos, with_cpu):
    """Batched per-halo properties on 50 000 NFW halos (about 30 M members): device time of
    all member passes and per-halo combines with inputs resident in HBM, the tuned schedule,
    the roofline of 60 algorithmic bytes per member, and the oracle's NumPy restatement of
    the reference's per-halo reductions timed on a bounded sample of the same halos."""
    import numpy as np
    import torch
    from amdahop import generate as G, halo_profiles, halo_props, halo_stellar

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    hs = G.halo_config()
    log(f"halo workload: {hs.nhalos} halos, {int(hs.offsets[-1])} members "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    pos = torch.from_numpy(hs.pos).to(dev)
    vel = torch.from_numpy(hs.vel).to(dev)
    mass = torch.from_numpy(hs.mass).to(dev)
    _, info = halo_props(hs.offsets, pos, vel, mass, hub_pt=70.3)      # autotune
    tile, wg = info.tile, info.workgroup
    dev_ms, pe_ms, vir_ms = [], [], []
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(steps):
        _, i2 = halo_props(hs.offsets, pos, vel, mass, hub_pt=70.3, tile=tile, workgroup=wg)
        dev_ms.append(i2.ms)
        pe_ms.append(i2.ms_direct_pe)
        vir_ms.append(i2.ms_virial)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - w0) * 1e3 / steps
    ms = float(np.median(dev_ms))
    M = int(hs.offsets[-1])
    achieved = HALO_BYTES_PER_MEMBER * M / (ms * 1e-3) / 1e9
    out = {"halos": hs.nhalos, "members": M, "ms": round(ms, 3), "wall_ms": round(wall_ms, 3),
           "Mmembers_per_s": round(M / (ms * 1e-3) / 1e6, 2), "tile": tile, "workgroup": wg,
           "autotune_trials": info.trials,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_member": HALO_BYTES_PER_MEMBER}}
    # the pair-sum potential of the halos below the pair-sum limit (det_halo_energies_1b3) is
    # FP64 arithmetic, O(n^2): its own roofline is the FP64 vector rate.  Pairs are the work
    # unit, PE_FLOPS_PER_PAIR FP64 operations each (the achieved rate); the operations the
    # kernel executes (masked pairs of the diagonal tiles included) come from the hardware
    # counters of the same workload (profiles/halo_fp64_c4.json, SQ_INSTS_VALU_FLOPS_FP64
    # summed over the launch), as do those of the shell tables / virial scan
    cnt = np.diff(hs.offsets)
    small = cnt[cnt < 1000].astype(np.int64)
    pairs = int(np.sum(small * (small - 1) // 2))
    pms = float(np.median(pe_ms))
    out["direct_pe"] = {"ms": round(pms, 3), "halos": int(small.size), "pairs": pairs,
                        "Gpairs_per_s": round(pairs / (pms * 1e-3) / 1e9, 2)}
    out["virial_ms"] = round(float(np.median(vir_ms)), 3)
    fpath = os.path.join(ROOT, "profiles", "halo_fp64_c4.json")
    if os.path.exists(fpath):
        with open(fpath) as f:
            fc = json.load(f)
        if fc.get("members") == M and fc.get("halos") == hs.nhalos:
            for kname, key, kms in (("k_direct_pe", "direct_pe", pms),
                                    ("k_virial_lanes", "virial", out["virial_ms"])):
                fl = fc["fp64_flops"][kname]
                alg = pairs * PE_FLOPS_PER_PAIR if kname == "k_direct_pe" else fl
                ach = alg / (kms * 1e-3) / 1e12
                out.setdefault("fp64_roofline", {})[key] = {
                    "bound": "fp64 valu", "achieved": round(ach, 3),
                    "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / FP64_VALU_PEAK_TFLOPS, 4), "fp64_flops_per_launch": alg,
                    "executed_fp64_flops_per_launch": fl,
                    "kernel_ms": round(kms, 3), "source": os.path.relpath(fpath, ROOT)}
    log(f"halo props: {ms:.2f} ms device ({wall_ms:.2f} ms wall), tile {tile} wg {wg}; "
        f"pair sums {pms:.2f} ms for {pairs} pairs, virial {out['virial_ms']:.2f} ms")
    # the same with the SC recentring of the reference's example parameter file (eps_SC 0.3,
    # dcell_min 0.001 Mpc: 1e-5 of a 100 Mpc box, the unit length here)
    sc = dict(sc=True, dcell_min=1e-5, eps_sc=0.3)
    sc_ms = [halo_props(hs.offsets, pos, vel, mass, hub_pt=70.3, tile=tile, workgroup=wg, **sc)[1].ms
             for _ in range(steps)]
    out["with_sc_recentring"] = {"ms": round(float(np.median(sc_ms)), 3), **sc}
    log(f"halo props with SC: {out['with_sc_recentring']['ms']:.2f} ms")
    # extended profiles (mode 'full'): peak circular velocity, NFW fit, inner slope
    props, _ = halo_props(hs.offsets, pos, vel, mass, hub_pt=70.3, tile=tile, workgroup=wg)
    halo_profiles(hs.offsets, pos, mass, props)
    pr_ms, pr_wall = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        w1 = time.perf_counter()
        prof, dms = halo_profiles(hs.offsets, pos, mass, props)
        torch.cuda.synchronize()
        pr_wall.append((time.perf_counter() - w1) * 1e3)
        pr_ms.append(dms)
    nfit = int(torch.isfinite(prof[:, 2]).sum().item())
    out["extended_profiles"] = {"ms": round(float(np.median(pr_ms)), 3),
                                "wall_ms": round(float(np.median(pr_wall)), 3),
                                "nfw_fits": nfit}
    log(f"extended profiles: {out['extended_profiles']['ms']:.2f} ms device, "
        f"{out['extended_profiles']['wall_ms']:.2f} ms wall, {nfit} NFW fits")
    # stellar properties: the first fifth of each halo's members (random positions) as stars,
    # ages 1 Myr - 10 Gyr log-uniform, random metallicities and densities (seed 6)
    rng = np.random.default_rng(6)
    local = np.arange(M) - np.repeat(hs.offsets[:-1], np.diff(hs.offsets))
    star_h = local < np.repeat(np.maximum(1, np.diff(hs.offsets) // 5), np.diff(hs.offsets))
    age_h = 10 ** rng.uniform(-3, 1, M)
    metal_h = rng.uniform(0, 0.04, M)
    dens_h = rng.uniform(1, 100, M)
    star_d = torch.from_numpy(star_h).to(dev)
    age_d, metal_d, dens_d = (torch.from_numpy(a).to(dev) for a in (age_h, metal_h, dens_h))
    halo_stellar(hs.offsets, pos, vel, mass, dens_d, star_d, age_d, metal_d, props)
    st_ms = [halo_stellar(hs.offsets, pos, vel, mass, dens_d, star_d, age_d, metal_d, props)[1]
             for _ in range(steps)]
    out["stellar"] = {"ms": round(float(np.median(st_ms)), 3), "stars": int(star_h.sum())}
    log(f"stellar: {out['stellar']['ms']:.2f} ms for {int(star_h.sum())} stars")
    del star_d, age_d, metal_d, dens_d
    del pos, vel, mass
    if with_cpu:
        import oracle.halo_props as HP
        rng = np.random.default_rng(0)
        pick = np.sort(rng.choice(hs.nhalos, size=cpu_sample_halos, replace=False))
        nmem = 0
        c0 = time.perf_counter()
        for h in pick:
            s, e = int(hs.offsets[h]), int(hs.offsets[h + 1])
            HP.halo_props(hs.pos[s:e], hs.vel[s:e], hs.mass[s:e], hub_pt=70.3)
            nmem += e - s
        t = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": round(nmem / t / 1e6, 4), "unit": "Mmembers/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_sample_halos} random halos of the same workload "
                                         f"({nmem} members), NumPy restatement, {t:.1f} s"}
        import oracle.halo_profiles as HPF
        pr = props.cpu().numpy()
        c0 = time.perf_counter()
        for h in pick:
            s, e = int(hs.offsets[h]), int(hs.offsets[h + 1])
            HPF.extended_profiles(hs.pos[s:e], hs.mass[s:e], pr[h, 1:4], pr[h, 24], 1.0, 430.1)
        t = time.perf_counter() - c0
        import oracle.halo_stellar as HSt
        pv = props.cpu().numpy()[:, 4:7]
        c0 = time.perf_counter()
        for h in pick:
            s, e = int(hs.offsets[h]), int(hs.offsets[h + 1])
            HSt.stellar_props(hs.pos[s:e], hs.vel[s:e], hs.mass[s:e], dens_h[s:e], star_h[s:e],
                              age_h[s:e], metal_h[s:e], pv[h])
        ts = time.perf_counter() - c0
        out["stellar"]["cpu_baseline"] = {
            "value": round(len(pick) / ts, 1), "unit": "halos/s", "cores": 1, "kind": "port",
            "sample": f"the same {cpu_sample_halos} halos, NumPy restatement, {ts:.1f} s"}
        out["extended_profiles"]["cpu_baseline"] = {
            "value": round(len(pick) / t, 1), "unit": "halos/s", "cores": 1, "kind": "port",
            "sample": f"the same {cpu_sample_halos} halos, NumPy/SciPy restatement, {t:.1f} s"}
    return out


def grouping_stage(gid, ngroups, steps, with_cpu):
    """Particles grouped by structure on the headline box: the step's group ids stand for the
    node ids of liste_parts (0 = in none), grouped by id and through a node-to-structure map
    onto a third as many structures (as MSM merges nodes); device time with the ids resident
    in HBM, roofline at 12 algorithmic bytes per particle; the reference's own NumPy grouping
    (argsort + np.unique, compute_neiKDtree_mod.py:109-119) on the same ids as the CPU leg."""
    import numpy as np
    import torch
    from amdahop.structures import group_members_device

    dev = gid.device
    n = int(gid.shape[0])
    nstruct = max(1, ngroups // 3)
    mp = torch.from_numpy(np.random.default_rng(5).integers(1, nstruct + 1, size=max(1, ngroups))
                          .astype(np.int32)).to(dev)
    labels = torch.empty(n, dtype=torch.int32, device=dev)
    pids = torch.empty(n, dtype=torch.int32, device=dev)
    off_n = torch.empty(ngroups + 2, dtype=torch.int64, device=dev)
    off_s = torch.empty(nstruct + 2, dtype=torch.int64, device=dev)
    group_members_device(gid, labels, pids, off_n, nnodes=ngroups)
    by_node, by_struct = [], []
    for _ in range(steps):
        by_node.append(group_members_device(gid, labels, pids, off_n, nnodes=ngroups))
        by_struct.append(group_members_device(gid, labels, pids, off_s, node_to_struct=mp,
                                              nstruct=nstruct))
    ms = float(np.median(by_struct))
    achieved = GROUP_BYTES_PER_PARTICLE * n / (ms * 1e-3) / 1e9
    out = {"n": n, "nodes": ngroups, "structures": nstruct, "ms": round(ms, 3),
           "ms_by_node_id": round(float(np.median(by_node)), 3),
           "Mparticles_per_s": round(n / (ms * 1e-3) / 1e6, 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_particle": GROUP_BYTES_PER_PARTICLE}}
    log(f"grouping: {ms:.3f} ms device on n={n}, {ngroups} nodes -> {nstruct} structures")
    if with_cpu:
        node = gid.cpu().numpy()
        lab = np.where(node > 0, mp.cpu().numpy()[np.maximum(node, 1) - 1], 0)
        c0 = time.perf_counter()
        order = np.argsort(lab)
        np.unique(lab[order], return_index=True, return_counts=True)
        t = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": round(n / t / 1e6, 4), "unit": "Mparticles/s",
                               "cores": 1, "kind": "port",
                               "sample": f"the whole workload ({n} labels), the reference's "
                                         f"NumPy argsort + np.unique, {t:.1f} s"}
    del labels, pids, off_n, off_s, mp
    return out


def ramses_stage(with_cpu):
    """The RAMSES particle reader on a 4,194,304-particle Ra3 output in 16 part files (a tenth
    stars), written to a temporary directory first (so read from the page cache): the native
    reader on 16 threads against the reference's per-file reading restated with NumPy/SciPy
    on one core."""
    import shutil
    import tempfile

    import numpy as np
    from amdahop import ramses

    n, ncpu, nst = 4_194_304, 16, 419_430
    rng = np.random.default_rng(7)
    top = tempfile.mkdtemp(prefix="amdahop_ramses_", dir="/tmp")
    rep = os.path.join(top, "output_00001")
    os.makedirs(rep)
    try:
        ids = np.concatenate([rng.permutation(n - nst) + 1, rng.permutation(nst) + 1])
        birth = np.concatenate([np.zeros(n - nst), -rng.uniform(0.5, 3.0, nst)])
        order = rng.permutation(n)
        nbytes = 0
        for icpu, part in enumerate(np.array_split(order, ncpu), start=1):
            m = len(part)
            path = os.path.join(rep, f"part_00001.out{icpu:05d}")
            ramses.write_part_file(path, ncpu, rng.uniform(0, 1, (m, 3)),
                                   rng.normal(size=(m, 3)), np.full(m, 1.0 / n), ids[part],
                                   "Ra3", birth[part], rng.uniform(0, 0.03, m))
            nbytes += os.path.getsize(path)
        ramses.read_particles(rep, "Ra3", "all", scale_l=1.0, scale_t=1.0, nthreads=16)
        t0 = time.perf_counter()
        r = ramses.read_particles(rep, "Ra3", "all", scale_l=1.0, scale_t=1.0, nthreads=16)
        t = time.perf_counter() - t0
        out = {"particles": n, "files": ncpu, "bytes": nbytes, "ms": round(t * 1e3, 2),
               "GB_per_s": round(nbytes / t / 1e9, 3), "threads": 16,
               "Mparticles_per_s": round(n / t / 1e6, 2)}
        assert r["npart"] == n
        log(f"ramses reader: {t * 1e3:.1f} ms for {nbytes / 1e6:.0f} MB ({n} particles)")
        if with_cpu:
            import oracle.ramses_reader as ORR
            c0 = time.perf_counter()
            ORR.read_particles(rep, "00001", ncpu, "Ra3", "all", True)
            tc = time.perf_counter() - c0
            out["cpu_baseline"] = {"value": round(n / tc / 1e6, 3), "unit": "Mparticles/s",
                                   "cores": 1, "kind": "port",
                                   "sample": f"the same files, NumPy/SciPy FortranFile reading "
                                             f"as the reference, {tc:.1f} s"}
    finally:
        shutil.rmtree(top, ignore_errors=True)
    return out


def adjoint_stage(steps, with_cpu):
    """SPH density adjoint on the 2,097,152-particle clumpy box (k=32): forward kNN rows
    frozen, upstream gradient of L = sum w log rho with w ~ N(0,1) (seed 3); device time
    of the adjoint (in the forward's Morton order) with inputs resident in HBM, roofline at
    216 algorithmic bytes per particle; the FP64 NumPy gradient timed on a 262,144-particle
    sample as the CPU leg."""
    import numpy as np
    import torch
    from amdahop import density_adjoint, generate as G, knn_device

    dev = torch.device("cuda:0")
    ps = G.config("c5")
    n, k = ps.n, ps.nvoisins
    pos = torch.from_numpy(ps.pos).to(dev)
    mass = torch.from_numpy(ps.mass).to(dev)
    dens, nbr, info, order = knn_device(pos, mass, k, return_order=True)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=n)).to(dev)
    grad_rho = w / dens
    # the forward's Morton order (as sph_density's backward passes it); the adjoint without
    # an order is timed beside it
    density_adjoint(pos, mass, nbr, grad_rho, order=order)
    dev_ms, plain_ms = [], []
    for _ in range(steps):
        dev_ms.append(density_adjoint(pos, mass, nbr, grad_rho, order=order)[2])
        plain_ms.append(density_adjoint(pos, mass, nbr, grad_rho)[2])
    ms = float(np.median(dev_ms))
    achieved = ADJOINT_BYTES_PER_PARTICLE * n / (ms * 1e-3) / 1e9
    out = {"n": n, "k": k, "ms": round(ms, 3), "Mparticles_per_s": round(n / (ms * 1e-3) / 1e6, 2),
           "forward_knn_ms": round(info.t_total, 3),
           "ms_without_order": round(float(np.median(plain_ms)), 3),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_particle": ADJOINT_BYTES_PER_PARTICLE}}
    log(f"adjoint: {ms:.3f} ms device on n={n} k={k}")
    del pos, mass, dens, nbr, order, w, grad_rho
    if with_cpu:
        import oracle.sph_adjoint as SA
        cs = G.config("c1")     # the same generator at 262,144 particles, k=32
        cpos = torch.from_numpy(cs.pos).to(dev)
        cmass = torch.from_numpy(cs.mass).to(dev)
        cd, cn, _ = knn_device(cpos, cmass, cs.nvoisins)
        g = np.random.default_rng(3).normal(size=cs.n) / cd.cpu().numpy()
        nb = cn.cpu().numpy().astype(np.int64)
        c0 = time.perf_counter()
        SA.density_grad(cs.pos, cs.mass, nb, g)
        t = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": round(cs.n / t / 1e6, 4), "unit": "Mparticles/s",
                               "cores": 1, "kind": "port",
                               "sample": f"same generator at N={cs.n}, k=32, FP64 NumPy "
                                         f"gradient with the same frozen rows, {t:.1f} s"}
    return out


def dropin_stage(ps):
    """One call of the drop-in compute_adaptahop.neikdtree.compute_adaptahop on the headline
    box, from the caller's host arrays to the module's outputs (node ids and node tables), the
    way compute_neiKDtree_mod.py drives it: wall time of the call (the first, which allocates,
    and a second), and the host structure tree's stage times and parallel branches."""
    import ctypes

    import numpy as np
    from amdahop._lib import check, lib, ptr
    from compute_adaptahop import neikdtree

    n, k = ps.n, ps.nvoisins
    nthreads = host_cpus()["threads"]
    neikdtree.sync_from_init_adaptahop(n, 20, 20, 0.7, 0.7, 0.3, 0.3, 1.0, 1.0, 0.7, 70.0, 1.0,
                                       1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5, 70.0, 1.0, [0, 0, 0])
    neikdtree.sync_from_change_pos(n, n, 0.001 / n ** (1 / 3), 0.001, 1.0, 1.0)
    neikdtree.sync_others(False, False, n, nthreads, 80.0, 1.0 / n, 0.0, 16, k, 4.0, 1.0, "MSM",
                          30, False)
    pos_f = np.asfortranarray(ps.pos.T)
    walls = []
    for _ in range(2):
        t0 = time.perf_counter()
        neikdtree.compute_adaptahop(pos_f, ps.mass)
        walls.append((time.perf_counter() - t0) * 1e3)
    st = np.zeros(4, dtype=np.int64)
    tms = np.zeros(5)
    check(lib().ahop_tree_stats(ptr(st), ptr(tms)))
    neikdtree.close()
    out = {"first_call_ms": round(walls[0], 2), "call_ms": round(walls[1], 2),
           "nodes": int(st[3]), "host_threads": nthreads,
           "host_tree_ms": round(float(tms.sum()), 3),
           "host_tree_stages_ms": {k_: round(float(v), 3) for k_, v in
                                   zip(("link", "colour", "create_nodes", "reorder", "tables"),
                                       tms)},
           "taskloop_levels": int(st[0]), "subtree_tasks": int(st[2])}
    log(f"drop-in call: {walls[1]:.1f} ms (first {walls[0]:.1f} ms), host tree "
        f"{tms.sum():.1f} ms, {int(st[3])} nodes")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--bucket", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    # the reference on the headline box itself, once (about 35 s on 16 host cores): a smaller
    # sample of the generator overstates its throughput on the full box (0.75 against 0.49
    # Mparticles/s at 2 097 152 particles), and best of 3 would hold the GPU for two minutes
    # more
    ap.add_argument("--cpu-sample-n", type=int, default=16_777_216,
                    help="particles of the reference's CPU run (16777216: the headline box)")
    ap.add_argument("--cpu-repeats", type=int, default=1)
    ap.add_argument("--cpu-full-stages", action="store_true",
                    help="also run the reference's whole entry point with its own stage "
                         "prints, with and without optimize_nodes")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-halo-props", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-grouping", action="store_true")
    ap.add_argument("--no-ramses", action="store_true")
    # one plain device copy of this many MiB before the timed steps: under a counter
    # profiler its FETCH_SIZE / WRITE_SIZE readings calibrate bytes per counter unit
    ap.add_argument("--copy-calibration-mib", type=int, default=0)
    ap.add_argument("--traffic-json", default=None,
                    help="JSON with measured HBM bytes per launch of the kNN kernel "
                         "({'fetch_bytes': F, 'write_bytes': W}) to report as roofline.traffic")
    ap.add_argument("--cpu-sample-halos", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(sharding design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from amdahop import device_hop, generate as G, run_device

    dev = torch.device("cuda:0")
    ps = G.config(args.workload)
    n, k = ps.n, ps.nvoisins
    log(f"workload {args.workload}: n={n} k={k}")
    pos = torch.from_numpy(ps.pos).to(dev)
    mass = torch.from_numpy(ps.mass).to(dev)
    dens = torch.empty(n, dtype=torch.float64, device=dev)
    gid = torch.empty(n, dtype=torch.int32, device=dev)

    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb} bytes read and written")
        del src, dst

    def step(mixed, saddle=False):
        return run_device(pos, mass, k, nhop=16, mixed=mixed, saddle=saddle, bucket=args.bucket,
                          density_out=dens, gid_out=gid)[2]

    for _ in range(args.warmup):
        step(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [step(True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the mixed step's outputs, kept for the comparison with the FP64 path
    dens_mixed = dens.clone()
    gid_mixed = gid.clone()
    hop_mixed = device_hop(n, dev)
    ms = (t1 - t0) * 1e3 / args.steps
    value = n / (ms * 1e-3) / 1e6
    knn_ms = sum(i.t_knn for i in infos) / len(infos)
    last = infos[-1]
    log(f"mixed: {ms:.2f} ms/step  sort {last.t_sort:.2f} tree {last.t_tree:.2f} knn {last.t_knn:.2f} "
        f"fallback {last.t_fallback:.2f} ({last.nfallback} q) hop {last.t_hop:.2f}")

    extra = {}
    # the drop-in boundary receives host arrays: the same step with the inputs copied from the
    # caller's pageable host memory and density + group ids copied back into pinned buffers
    # kept across calls, as the drop-in module keeps its outputs (PCIe-inclusive, not `value`)
    dens_h = torch.empty(n, dtype=torch.float64, pin_memory=True)
    gid_h = torch.empty(n, dtype=torch.int32, pin_memory=True)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    pos_h = torch.from_numpy(ps.pos).to(dev)
    mass_h = torch.from_numpy(ps.mass).to(dev)
    torch.cuda.synchronize()
    p1 = time.perf_counter()
    run_device(pos_h, mass_h, k, nhop=16, mixed=True, saddle=False, bucket=args.bucket,
               density_out=dens, gid_out=gid)
    p2 = time.perf_counter()
    dens_h.copy_(dens)
    gid_h.copy_(gid)
    torch.cuda.synchronize()
    p3 = time.perf_counter()
    del pos_h, mass_h, dens_h, gid_h
    extra["pcie_inclusive"] = {"value": round(n / (p3 - p0) / 1e6, 3), "unit": "Mparticles/s",
                               "h2d_ms": round((p1 - p0) * 1e3, 2),
                               "step_ms": round((p2 - p1) * 1e3, 2),
                               "d2h_ms": round((p3 - p2) * 1e3, 2)}
    if not args.no_fp64:
        step(False)
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        nf = max(2, args.steps // 2)
        finfo = [step(False) for _ in range(nf)]
        torch.cuda.synchronize()
        fms = (time.perf_counter() - f0) * 1e3 / nf
        extra["fp64_path"] = {"value": round(n / (fms * 1e-3) / 1e6, 3), "ms_per_step": round(fms, 3),
                              "knn_ms": round(sum(i.t_knn for i in finfo) / nf, 3)}
        extra["speedup_mixed_vs_fp64"] = round(fms / ms, 3)
        log(f"fp64: {fms:.2f} ms/step knn {extra['fp64_path']['knn_ms']:.2f}")
        # the metric's own claim, checked on every particle of the run: density error of the
        # mixed path against the FP64 path, and equal group ids and hop targets
        hop64 = device_hop(n, dev)
        rel = ((dens_mixed - dens).abs() / dens.abs()).max().item()
        gid_eq = bool(torch.equal(gid_mixed, gid))
        hop_eq = bool(torch.equal(hop_mixed, hop64))
        extra["rho_max_rel_err_vs_fp64"] = rel
        extra["ids_equal"] = gid_eq and hop_eq
        extra["mixed_vs_fp64_checked"] = {"particles": n, "density_bit_equal":
                                          bool(torch.equal(dens_mixed, dens)),
                                          "group_ids_equal": gid_eq, "hop_targets_equal": hop_eq}
        log(f"mixed vs fp64: max rel density error {rel:.3g}, group ids equal {gid_eq}, "
            f"hop targets equal {hop_eq}")
        del hop64
    del dens_mixed, gid_mixed, hop_mixed
    sinfo = step(True, saddle=True)
    extra["saddle_stage"] = {"ms": round(sinfo.t_saddle, 3), "directed_edges_kept": int(sinfo.nsaddle)}
    extra["stages_ms"] = {"sort": round(last.t_sort, 3), "tree": round(last.t_tree, 3),
                          "knn_density": round(last.t_knn, 3), "fp64_rerun": round(last.t_fallback, 3),
                          "hop_groups": round(last.t_hop, 3)}
    extra["nactive"], extra["ngroups"], extra["nfallback"] = int(last.nactive), int(last.ngroups), int(last.nfallback)
    w = max(1, last.knn_waves)
    extra["knn_work_per_wave"] = {"nodes": round(last.knn_nodes / w, 1),
                                  "buckets": round(last.knn_buckets / w, 1),
                                  "candidates_per_lane": round(last.knn_candidates / w, 1),
                                  "insert_steps": round(last.knn_insert_steps / w, 1),
                                  "inserts_per_lane": round(last.knn_inserts / w / 64, 1),
                                  "log_max": int(last.knn_log_max),
                                  "log_overflow": int(last.knn_log_overflow),
                                  "fp32_fp64_misorder": int(last.knn_misorder)}

    achieved = KNN_BYTES_PER_PARTICLE * n / (knn_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_knn_sel<64> (kNN + SPH density, mixed precision)", "kernel_ms": round(knn_ms, 3),
                "bytes_per_particle": KNN_BYTES_PER_PARTICLE,
                "step_achieved_GBs": round(STEP_BYTES_PER_PARTICLE * n / (ms * 1e-3) / 1e9, 2)}
    # the kernel is bound by instruction issue, not HBM: its FP32 candidate arithmetic (every
    # lane evaluates every staged candidate, 9 FP32 operations each: 3 differences, a square,
    # two fused multiply-adds counted as 2 each, the comparison) against the FP32 vector peak
    evals = int(last.knn_candidates) * 64
    fl = 9.0 * evals / (knn_ms * 1e-3) / 1e12
    roofline["fp32_candidates"] = {"evaluations": evals, "ops_per_evaluation": 9,
                                   "achieved_TFLOPs": round(fl, 3), "peak_TFLOPs": FP32_VALU_PEAK_TFLOPS,
                                   "frac": round(fl / FP32_VALU_PEAK_TFLOPS, 4)}
    tpath = args.traffic_json or os.path.join(ROOT, "profiles", f"knn_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("n") == n and tr.get("bucket") == args.bucket:
            roofline["traffic"] = int(tr["fetch_bytes"] + tr["write_bytes"])
            roofline["traffic_source"] = os.path.relpath(tpath, ROOT)
    # the bound it does sit on: VALU issue, the counted wave instructions of one launch over
    # this run's kernel time against the chip's rate (SIMDs x clock / 4 cycles per wave64
    # VALU instruction)
    vpath = os.path.join(ROOT, "profiles", f"knn_{args.workload}_valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            va = json.load(f)
        if va.get("n") == n and va.get("bucket") == args.bucket:
            rate = va["valu_wave_instructions_per_launch"] / (knn_ms * 1e-3)
            peak = va["simds"] * va["clock_ghz"] * 1e9 / va["cycles_per_wave64_valu_instruction"]
            roofline["valu_issue"] = {"wave_instructions": va["valu_wave_instructions_per_launch"],
                                      "achieved_per_s": float(f"{rate:.4g}"),
                                      "peak_per_s": float(f"{peak:.4g}"),
                                      "frac": round(rate / peak, 3),
                                      "source": os.path.relpath(vpath, ROOT)}

    if not args.no_grouping:
        extra["grouping"] = grouping_stage(gid, int(last.ngroups), max(3, args.steps),
                                           not args.no_cpu_baseline)
    del pos, mass, dens, gid
    if not args.no_halo_props:
        extra["halo_props"] = halo_props_stage(max(2, args.steps // 2), args.cpu_sample_halos,
                                               not args.no_cpu_baseline)
    if not args.no_adjoint:
        extra["density_adjoint"] = adjoint_stage(max(3, args.steps), not args.no_cpu_baseline)
    if not args.no_ramses:
        extra["ramses_reader"] = ramses_stage(not args.no_cpu_baseline)

    if not args.no_dropin:
        extra["dropin_call"] = dropin_stage(ps)

    cpu = None
    if not args.no_cpu_baseline:
        nthreads = host_cpus()["threads"]
        log(f"cpu baseline: {args.cpu_sample_n} particles on {nthreads} threads")
        cpu = cpu_baseline(ps, args.cpu_sample_n, nthreads, args.cpu_repeats,
                           args.cpu_full_stages)

    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mparticles/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": f"{args.workload}: {n} particles, periodic unit box, k={k}, nhop=16, "
                                  "Plummer clumps + uniform background, mixed precision",
                      "n": n, "nvoisins": k, "nhop": 16, "precision": "mixed", "bucket": args.bucket},
           "roofline": roofline, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
