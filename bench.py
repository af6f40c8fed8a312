#!/usr/bin/env python3
"""Benchmark of the MI355X forward model (one GPU).

Workload (limb sounder): 4096 atmospheres (81 levels, standard temperature
plus a seeded Gaussian random field of 3 K and 5 km correlation length; CO2,
H2O and O3), 32 tangent heights 5..67 km seen from 800 km, 12 channels
700..1400 cm^-1, EGA with all continua and refraction: 131,072 rays and
1,572,864 radiances per step.  One step is one forward pass over all rays
with atmospheres and geometry already resident in HBM.

The JSON line reports whole-job radiances/s of the FP32 path (FP64 ray
geometry), the FP64 path's rate beside it, the roofline of the dominant
kernel (the FP32 table lookups; its time from HIP events on the library's
stream, its instruction and HBM byte counts from a committed rocprofv3
counter profile), and the CPU FP64 oracle timed on this host's cores on a
bounded sample of the same workload.

Jacobians (the "jacobian" object): nadir sounder, 16,384 atmospheres x 8
view angles 0..50 deg x 96 channels 650..1600 cm^-1 (CO2, H2O, O3; EGA,
continua, refraction), state T and the three volume mixing ratios on 81
levels (n = 324, m = 768 per atmosphere).  One step is one adjoint pass
(forward model and all K rows, K written to HBM).  Beside it: the FD path
that kernel() runs, n + 1 = 325 FP64 forward models per Jacobian, measured
on 64 of the atmospheres in one device pass (perturbations and differences
on the device), its K compared with the adjoint's, and the CPU oracle
kernel() (OpenMP over the 324 columns) on a sample of atmospheres.
The FP32 adjoint's error is stated against the FP64 adjoint on a subset:
relative L2 per state block and the brightness temperature of the forward
part.

Serving (the "serving" object): Poisson arrivals at 2000 requests/s (seeded)
into the request server for a fixed time; 60% forward, 25% Jacobian, 15%
Levenberg-Marquardt retrievals (CONV_ITMAX 10, S_a: 2 K and 20% vmr,
exponential correlation over 5 km; S_eps 0.1 K^2); 1..64 rays per request,
limb or nadir 50/50; state T plus 0..3 volume mixing ratios on 81 levels
(n = 81..324); 12 channels, brightness temperatures, FP32 lookups, adjoint
Jacobians.  Latency is arrival (submission) to result in host memory;
p50/p99 per kind and overall, with the throughput sustained.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "radiances/s fwd; Jacobians/s adjoint vs FD; p50/p99 retrieval latency"
HBM_PEAK = 8.0e12
# Vector-instruction issue peak of one MI355X: 256 CUs x 2.4 GHz, one wave64
# VALU instruction per CU per cycle (four SIMD16 units, four cycles each).
VALU_PEAK = 256 * 2.4e9
# Counters of one k_lookup_f32 launch at the default workload (4096
# profiles), from rocprofv3 --pmc runs of this script
# (profiles/k_lookup_f32_limb_4096x32x12_counters.md): SQ_INSTS_VALU and
# SQ_INSTS_VMEM_RD summed over the launch (wave instructions), and HBM bytes
# = FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 KiB x 1024 (factors from a 1 GiB
# device copy under the same counters).  Other workload sizes report null.
LOOKUP_F32_COUNTERS = {"nprof": 4096, "valu": 3.0198e10, "vmem_rd": 1.3640e9, "hbm_bytes": 39.37e9,
                       "profile": "profiles/k_lookup_f32_limb_4096x32x12_counters.md"}
NPROF, NRAY_PER, GASES = 4096, 32, ["CO2", "H2O", "O3"]


def build_workload(nprof):
    import numpy as np
    from jurassic_amd import abi, scenes
    ctl = scenes.limb_3gas_ctl()
    zt = scenes.limb_tangent_heights()
    geom = scenes.limb_geometry(800.0, zt)
    atms = (abi.atm_t * nprof)()
    for k in range(nprof):
        atms[k] = scenes.make_profile_atm(GASES, iprof=k)
    # ray order: tangent height major, profile minor -> each block of 64 rays
    # holds one tangent height over 64 atmospheres (similar paths)
    nray = nprof * len(zt)
    prof = np.empty(nray, dtype=np.int32)
    g = np.empty((nray, 7))
    for j in range(len(zt)):
        sl = slice(j * nprof, (j + 1) * nprof)
        prof[sl] = np.arange(nprof)
        g[sl, 0] = 0.0
        g[sl, 1] = geom["obsz"][j]
        g[sl, 2] = geom["obslon"][j]
        g[sl, 3] = geom["obslat"][j]
        g[sl, 4] = geom["vpz"][j]
        g[sl, 5] = geom["vplon"][j]
        g[sl, 6] = geom["vplat"][j]
    return ctl, atms, prof, g


def host_cpu():
    """The host's CPU model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": avail, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(nsample):
    """FP64 oracle on this host's cores: OpenMP across atmospheres after one
    warm call (the kernel() pattern of the reference), on nsample profiles of
    the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from jurassic_amd import abi, scenes
    ctl = scenes.limb_3gas_ctl()
    zt = scenes.limb_tangent_heights()
    atms = (abi.atm_t * nsample)()
    obss = (abi.obs_t * nsample)()
    for k in range(nsample):
        atms[k] = scenes.make_profile_atm(GASES, iprof=k)
        obss[k] = scenes.make_obs(scenes.limb_geometry(800.0, zt))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # warm: tables and source function
    O.formod(ctl, abi.copy_struct(atms[0]), abi.copy_struct(obss[0]))
    t0 = time.perf_counter()
    O.formod_batch(ctl, atms, obss, nthreads=threads)
    dt = time.perf_counter() - t0
    nrad = nsample * len(zt) * ctl.nd
    return dict(value=nrad / dt, unit="radiances/s", cores=threads, kind="port",
                sample="%d profiles x %d tangent heights x %d channels (%d radiances), FP64 oracle, %.1f s"
                % (nsample, len(zt), ctl.nd, nrad, dt), host=host_cpu())


def lookup_roofline(nprof, nray, nd, mean_np, kms):
    """Roofline of the dominant forward kernel, k_lookup_f32 (one launch per
    pass; its time from HIP events on the library's stream).

    `bound`/`achieved`/`frac` are HBM: the bytes the kernel must move at least
    (every LOS field it reads once, the tables once, the path-emissivity
    write) per second of its time, against 8 TB/s; `traffic` is the HBM bytes
    per launch of the committed counter profile.  The kernel's limiter is
    vector-instruction issue, reported beside it (`issue`): SQ_INSTS_VALU of
    the same profile per second of the measured time against VALU_PEAK.  The
    per-lookup gather model (176 B per (segment, channel, emitter): 4 corners
    x 2 float brackets of 16 B plus the p and two T brackets, 48 B) is mostly
    served by L2, so its rate is not an HBM figure."""
    t = kms[1] * 1e-3
    n_chain = nray * mean_np * nd * len(GASES)         # (ray, point, channel, emitter) steps
    unique = nray * mean_np * (2 + len(GASES)) * 8 + n_chain * 4 + 36 * 40 * 20 * 200 * 16
    gather = n_chain * 176
    cnt = LOOKUP_F32_COUNTERS if nprof == LOOKUP_F32_COUNTERS["nprof"] else None
    r = {"kernel": "k_lookup_f32<false, false>", "kernel_ms": kms[1],
         "kernel_ms_all": {"raytrace": kms[0], "lookup": kms[1], "accumulate": kms[2]},
         "bound": "hbm", "achieved": unique / t / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
         "frac": unique / t / HBM_PEAK, "traffic": cnt["hbm_bytes"] if cnt else None}
    if cnt:
        ach = cnt["valu"] / t
        r["issue"] = {"limiter": "VALU issue", "achieved": ach / 1e9, "peak": VALU_PEAK / 1e9,
                      "unit": "G wave-instructions/s", "frac": ach / VALU_PEAK,
                      "valu_per_launch": cnt["valu"], "vmem_rd_per_launch": cnt["vmem_rd"]}
    r["hbm"] = {"unique_bytes": unique, "unique_GBps": unique / t / 1e9,
                "unique_frac": unique / t / HBM_PEAK,
                "measured_bytes": cnt["hbm_bytes"] if cnt else None,
                "measured_GBps": cnt["hbm_bytes"] / t / 1e9 if cnt else None,
                "peak_GBps": HBM_PEAK / 1e9}
    r["gather_model"] = {"bytes_per_lookup": 176, "bytes_per_launch": gather, "GBps": gather / t / 1e9,
                         "note": "algorithmic gather bytes; L2 serves most of them, so this rate is "
                                 "not HBM traffic"}
    r["counters_profile"] = cnt["profile"] if cnt else None
    return r


NJAC_PROF, NJAC_ANG, NJAC_CH = 16384, 8, 96


def build_jac_workload(nprof):
    """Nadir Jacobian workload: ray order view angle major, atmosphere minor."""
    import numpy as np
    from jurassic_amd import abi, scenes
    ctl = scenes.nadir_jac_ctl(NJAC_CH)
    geom = scenes.nadir_geometry(705.0, np.arange(NJAC_ANG) * 50.0 / 7.0)
    atms = (abi.atm_t * nprof)()
    for k in range(nprof):
        atms[k] = scenes.make_profile_atm(GASES, iprof=k)
    nray = nprof * NJAC_ANG
    prof = np.empty(nray, dtype=np.int32)
    g = np.zeros((nray, 7))
    for j in range(NJAC_ANG):
        sl = slice(j * nprof, (j + 1) * nprof)
        prof[sl] = np.arange(nprof)
        for c, key in enumerate(("obsz", "obslon", "obslat", "vpz", "vplon", "vplat")):
            g[sl, 1 + c] = geom[key][j]
    return ctl, atms, prof, g


def cpu_kernel_baseline(nsample):
    """The oracle's kernel() (one-sided FD, OpenMP over the state columns, as
    the reference) on nsample atmospheres of the Jacobian workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from jurassic_amd import abi, scenes
    import numpy as np
    ctl = scenes.nadir_jac_ctl(NJAC_CH)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    obs = scenes.make_obs(scenes.nadir_geometry(705.0, np.arange(NJAC_ANG) * 50.0 / 7.0))
    O.formod(ctl, scenes.make_profile_atm(GASES, iprof=0), abi.copy_struct(obs))  # warm
    t0 = time.perf_counter()
    for k in range(nsample):
        O.kernel(ctl, scenes.make_profile_atm(GASES, iprof=k), abi.copy_struct(obs), nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=nsample / dt, unit="Jacobians/s", cores=threads, kind="port",
                sample="%d atmospheres x 768 x 324 one-sided FD (325 FP64 oracle formods each), %.1f s"
                % (nsample, dt))


def jacobian_bench(args, torch, gpu, abi):
    import numpy as np
    nprof = args.jac_nprof
    ctl, atms, prof, geom = build_jac_workload(nprof)
    nray, nd = len(prof), ctl.nd
    xm, n = gpu.state_map(ctl, atms[0])
    m = NJAC_ANG * nd
    log("Jacobian workload: %d atmospheres, %d rays, %d channels, n = %d, K %.1f GB"
        % (nprof, nray, nd, n, nray * nd * n * 8 / 1e9))
    ctx = gpu.Context(ctl)
    dev = torch.device("cuda:0")
    atm_d = torch.frombuffer(bytearray(bytes(atms)), dtype=torch.uint8).to(dev)
    prof_d = torch.from_numpy(prof).to(dev)
    geom_d = torch.from_numpy(np.ascontiguousarray(geom)).to(dev)
    pmap_d = torch.zeros(nprof, dtype=torch.int32, device=dev)
    rad_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    tau_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    k_d = torch.empty(nray * nd * n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def adj(precision, natm=nprof, nr=nray, prof_ptr=None, geom_ptr=None, rad=rad_d, tau=tau_d, kk=k_d):
        ctx.jacobian_device(natm, atm_d.data_ptr(), nr, prof_ptr or prof_d.data_ptr(),
                            geom_ptr or geom_d.data_ptr(), rad.data_ptr(), tau.data_ptr(), xm[None, :],
                            pmap_d.data_ptr(), n, kk.data_ptr(), precision=precision)

    steps = max(1, args.jac_steps)
    adj(abi.JUR_FP32)
    torch.cuda.synchronize()
    kms = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    for _ in range(steps):
        adj(abi.JUR_FP32)
        kms = [a + b for a, b in zip(kms, ctx.last_timing())]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    kms = [k / steps for k in kms]
    mean_np = ctx.last_mean_np()
    jac_rate = nprof / dt
    log("adjoint FP32: %.1f ms/step (%.0f Jacobians/s); kernels: ray tracing %.1f, lookups %.1f, "
        "reverse sweep + K store %.1f ms; mean LOS points %.1f" % (dt * 1e3, jac_rate, kms[0], kms[1], kms[2], mean_np))

    # the FD path of kernel() (src/jurassic.c:4004-4099), measured: for the
    # first nfd atmospheres, the n + 1 FP64 forward models of each Jacobian
    # (one-sided steps of :4046-4068: 1 K for T, max(|0.01 x|, 1e-15) for
    # vmr) in one device pass, the perturbed atmospheres built and the
    # differences formed on the device, all inside the timed region
    nfd = min(args.jac_fd_nprof, nprof)
    asz = C.sizeof(abi.atm_t)
    col_off = np.empty(n, dtype=np.int64)
    is_t = np.zeros(n, dtype=bool)
    off_t, off_q = abi.atm_t.t.offset // 8, abi.atm_t.q.offset // 8
    for j in range(n):
        blk, ip = divmod(j, 81)
        col_off[j] = off_t + ip if blk == 0 else off_q + (blk - 1) * abi.NP + ip
        is_t[j] = blk == 0
    col_off_d = torch.from_numpy(col_off).to(dev)
    is_t_d = torch.from_numpy(is_t).to(dev)
    natm_fd = nfd * (n + 1)
    nray_fd = NJAC_ANG * natm_fd
    # rays: angle major, then atmosphere (profile major, perturbation minor)
    prof_fd = torch.arange(natm_fd, dtype=torch.int32, device=dev).repeat(NJAC_ANG)
    ang_rows = torch.from_numpy(np.ascontiguousarray(geom[np.arange(NJAC_ANG) * nprof])).to(dev)
    geom_fd = ang_rows.repeat_interleave(natm_fd, dim=0).contiguous()
    rad_fd = torch.empty(nray_fd * nd, dtype=torch.float64, device=dev)
    tau_fd = torch.empty_like(rad_fd)
    jcols = torch.arange(n, device=dev)

    def fd_jacobians():
        base = atm_d.view(nprof, asz)[:nfd]
        pert = base[:, None, :].expand(nfd, n + 1, asz).contiguous()
        pf = pert.view(torch.float64).view(nfd, n + 1, asz // 8)
        x0 = pf[:, 0, col_off_d]
        h = torch.where(is_t_d[None, :], torch.ones_like(x0), torch.clamp(torch.abs(0.01 * x0), min=1e-15))
        pf[:, 1 + jcols, col_off_d] += h
        ctx.formod_device(natm_fd, pert.data_ptr(), nray_fd, prof_fd.data_ptr(), geom_fd.data_ptr(),
                          rad_fd.data_ptr(), tau_fd.data_ptr(), precision=abi.JUR_FP64)
        radv = rad_fd.view(NJAC_ANG, nfd, n + 1, nd)
        # K[profile][angle][channel][column]
        return ((radv[:, :, 1:, :] - radv[:, :, :1, :]) / h[None, :, :, None]).permute(1, 0, 3, 2).contiguous()
    fd_jacobians()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k_fd = fd_jacobians()
    torch.cuda.synchronize()
    dt_fd = time.perf_counter() - t0
    fd_rate = nfd / dt_fd
    log("FD FP64 (kernel()): %d Jacobians = %d forward models x %d rays in %.1f ms -> %.1f Jacobians/s"
        % (nfd, natm_fd, NJAC_ANG, dt_fd * 1e3, fd_rate))

    # FP32 adjoint against FP64 adjoint on the first atmospheres of every angle
    nsub = min(args.jac_check, nprof)
    sel = np.concatenate([np.arange(j * nprof, j * nprof + nsub) for j in range(NJAC_ANG)])
    prof_s = torch.from_numpy(prof[sel].copy()).to(dev)
    geom_s = torch.from_numpy(np.ascontiguousarray(geom[sel])).to(dev)
    ns = len(sel)
    outs = {}
    for p in (abi.JUR_FP32, abi.JUR_FP64):
        r = torch.empty(ns * nd, dtype=torch.float64, device=dev)
        t = torch.empty_like(r)
        kk = torch.empty(ns * nd * n, dtype=torch.float64, device=dev)
        adj(p, nsub, ns, prof_s.data_ptr(), geom_s.data_ptr(), r, t, kk)
        torch.cuda.synchronize()
        outs[p] = (r.cpu().numpy().reshape(ns, nd), kk.cpu().numpy().reshape(ns * nd, n))
    nu = np.array([ctl.nu[i] for i in range(nd)])
    c1, c2 = 1.19104259e-8, 1.43877506

    def bt(r):
        return c2 * nu / np.log1p(c1 * nu ** 3 / r)
    bt_err = float(np.abs(bt(outs[abi.JUR_FP32][0]) - bt(outs[abi.JUR_FP64][0])).max())
    k32, k64 = outs[abi.JUR_FP32][1], outs[abi.JUR_FP64][1]
    blocks = {}
    for b, name in enumerate(["T"] + GASES):
        s_ = slice(b * 81, (b + 1) * 81)
        blocks[name] = float(np.linalg.norm(k32[:, s_] - k64[:, s_]) / np.linalg.norm(k64[:, s_]))
    log("FP32 vs FP64 adjoint on %d atmospheres: BT %.2e K, K rel L2 %s" % (nsub, bt_err, blocks))
    # the adjoint (refracted path included) against kernel()'s one-sided
    # differences on the FD atmospheres
    ncmp = min(nfd, nsub)
    kf = k_fd[:ncmp].cpu().numpy()  # [atm][angle][channel][column]
    fd_vs = {}
    for p, k_adj in ((abi.JUR_FP32, k32), (abi.JUR_FP64, k64)):
        ka = k_adj.reshape(NJAC_ANG, nsub, nd, n)[:, :ncmp].transpose(1, 0, 2, 3)
        fd_vs["fp32" if p == abi.JUR_FP32 else "fp64"] = {
            name: float(np.linalg.norm(ka[..., b * 81:(b + 1) * 81] - kf[..., b * 81:(b + 1) * 81])
                        / np.linalg.norm(kf[..., b * 81:(b + 1) * 81]))
            for b, name in enumerate(["T"] + GASES)}
    log("adjoint vs kernel() FD on %d atmospheres (refraction on): %s" % (ncmp, fd_vs))
    del k_fd, rad_fd, tau_fd
    ctx.close()
    del k_d

    # HBM bytes a Jacobian pass must move at least, per Jacobian: per (ray,
    # LOS point, channel) and emitter the lookup's stored partials (16 B) and
    # path emissivity (4 B), each written once and read once by the reverse
    # chains; per (ray, point, channel) the accumulation's H_i (8 B) and the
    # path's point values d rad/d z, d rad/d ds (16 B), each written and
    # read once; per (ray, point) the LOS fields (16 FP64 fields with the
    # march's positions and directions, read twice) and the reverse march's
    # 53 FP64 coefficients (written and read once); the FP64 K written once
    bytes_per_jac = (m * mean_np * (len(GASES) * 2 * (16 + 4) + 2 * 8 + 2 * 16)
                     + NJAC_ANG * mean_np * (2 * 16 * 8 + 2 * 53 * 8) + m * n * 8)
    out = {
        "value": jac_rate, "unit": "Jacobians/s", "ms_per_step": dt * 1e3, "steps": steps,
        "workload": "nadir sounder: %d atmospheres x %d view angles x %d channels, n = %d, m = %d "
                    "(CO2, H2O, O3; EGA; continua and refraction on; the refracted path differentiated)"
                    % (nprof, NJAC_ANG, nd, n, m),
        "precision": "FP32 lookups, FP64 geometry and adjoint accumulation",
        "mean_los_points": mean_np,
        "kernel_ms": {"raytrace": kms[0], "lookup": kms[1], "reverse_sweep_and_store": kms[2]},
        "fd_fp64_gpu": {"value": fd_rate, "unit": "Jacobians/s",
                        "method": "measured: kernel()'s n + 1 = %d one-sided FP64 forward models per "
                                  "Jacobian for %d atmospheres x %d angles x %d channels (%d forward "
                                  "models, %d rays) in one device pass, perturbations and differences "
                                  "on the device" % (n + 1, nfd, NJAC_ANG, nd, natm_fd, nray_fd),
                        "ms": dt_fd * 1e3},
        "adjoint_over_fd_fp64_gpu": jac_rate / fd_rate,
        "adjoint_vs_kernel_fd_rel_l2": {"atmospheres": ncmp, "note": "difference to kernel()'s one-sided "
                                        "finite differences (its own step sizes), mostly their "
                                        "first-order truncation; the adjoint includes the "
                                        "refracted path's derivative",
                                        **fd_vs},
        "fp32_vs_fp64": {"atmospheres": nsub, "bt_max_abs_K": bt_err, "k_rel_l2": blocks},
        "roofline": {"bound": "hbm", "achieved": bytes_per_jac * jac_rate / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": bytes_per_jac * jac_rate / HBM_PEAK,
                     "bytes_per_jacobian": bytes_per_jac,
                     "model": "unique bytes: partials and path emissivities written and read once, "
                              "per-point accumulation and path values, LOS fields, reverse-march "
                              "coefficients, K"},
    }
    if not args.no_cpu:
        ns_cpu = args.jac_cpu_sample
        log("CPU kernel() baseline on %d atmospheres" % ns_cpu)
        out["cpu_baseline"] = cpu_kernel_baseline(ns_cpu)
        out["adjoint_over_cpu"] = jac_rate / out["cpu_baseline"]["value"]
    return out


NWIDE_CH = 2048
WIDE_GASES = ["CO2", "H2O", "O3", "CH4", "N2O", "CO", "HNO3", "F11"]


def wide_tables(torch, nu, gases):
    """Synthetic tables of the wide workload, built on the device in the
    reference's tbl_t order: per (channel, emitter) 40 p x 20 T x 200 u
    (table_grids), eps = 1 - exp(-k (u/1e20)^0.8) stored as float,
    k = s f(nu) (p/1013)^0.7 (T/250)^-1.5 with s = 1e-2 for every emitter and
    a channel factor f(nu) = 0.2 + 0.8 (0.5 + 0.5 sin(2 pi nu / 97)), so that
    channels hold different tables."""
    import numpy as np
    from jurassic_amd import scenes
    p, t, u = scenes.table_grids()
    nd, ng, tp, tt, tu = len(nu), len(gases), len(p), len(t), len(u)
    dev = torch.device("cuda:0")
    i32 = torch.int32
    tabs = dict(np=torch.full((nd, ng), tp, dtype=i32, device=dev),
                nt=torch.full((nd, ng, tp), tt, dtype=i32, device=dev),
                nu=torch.full((nd, ng, tp, tt), tu, dtype=i32, device=dev),
                p=torch.tensor(p, device=dev).expand(nd, ng, tp).contiguous(),
                t=torch.tensor(t, device=dev).expand(nd, ng, tp, tt).contiguous(),
                u=torch.tensor(u.astype(np.float32), device=dev).expand(nd, ng, tp, tt, tu).contiguous())
    eps = torch.empty((nd, ng, tp, tt, tu), dtype=torch.float32, device=dev)
    pk = (torch.tensor(p, device=dev) / 1013.0) ** 0.7
    tk = (torch.tensor(t, device=dev) / 250.0) ** -1.5
    xu = (torch.tensor(u, device=dev) / 1e20) ** 0.8
    base = 1e-2 * pk[:, None] * tk[None, :]
    nut = torch.tensor(np.asarray(nu), device=dev)
    f = 0.2 + 0.8 * (0.5 + 0.5 * torch.sin(2 * math.pi * nut / 97.0))
    for c0 in range(0, nd, 64):
        c1 = min(nd, c0 + 64)
        x = f[c0:c1, None, None, None] * base[None, :, :, None] * xu[None, None, None, :]
        eps[c0:c1] = (-torch.expm1(-x)).float()[:, None].expand(c1 - c0, ng, tp, tt, tu)
    tabs["eps"] = eps
    return tabs


def boxcar_filters(nu, width=1.0, step=0.0005):
    """In-memory copies of the boxcar filter files (write_boxcar_filter)."""
    import numpy as np
    half = int(round(width / 2 / step))
    n = 2 * half + 5
    i = np.arange(n)
    w = ((i >= 2) & (i < n - 3)).astype(float)
    xs = [np.round(v + (i - half - 2) * step, 4) for v in nu]
    off = np.arange(len(nu) + 1, dtype=np.int32) * n
    return off, np.concatenate(xs), np.tile(w, len(nu))


def tblgen_bench(args, gpu):
    """Table production (tblgen, src/tblgen.c:44-140) for one pressure and
    temperature: a broad-band filter over a 1,000,001-point optical-depth
    spectrum, i.e. 10^6 points x ~600 column densities of exp terms.  Times
    the whole C-ABI call (host selection of the points and weights, copies,
    kernels, row selection) and the oracle's sequential loop on the same
    input."""
    import numpy as np
    n, nu0, nu1 = 1000001, 600.0, 1100.0
    rng = np.random.default_rng(7)
    nu = np.linspace(nu0, nu1, n)
    tau = np.full(n, 1e-5)
    for c, st, w in zip(rng.uniform(nu0, nu1, 400), 10 ** rng.uniform(-3, 0, 400), rng.uniform(0.02, 0.2, 400)):
        lo, hi = np.searchsorted(nu, [c - 50 * w, c + 50 * w])
        tau[lo:hi] += st * w / np.pi / ((nu[lo:hi] - c) ** 2 + w * w)
    nuf = np.linspace(610.0, 1090.0, 481)
    filt = 0.5 + 0.5 * np.cos((nuf - 850.0) / 240.0 * np.pi / 2)
    press, temp = 300.0, 240.0
    gpu.tblgen(press, temp, tau, nu0, nu1, nuf, filt)  # warm
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        u, e = gpu.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
    dt = (time.perf_counter() - t0) / reps
    npt = int(((nu >= nuf[0]) & (nu <= nuf[-1])).sum())
    ku = 1
    while 1.122 ** ku <= 1e30:
        ku += 1
    out = {"workload": "one (p, T): %d-point spectrum, %d points under the filter, %d column densities"
                       % (n, npt, ku),
           "rows": int(len(u)), "ms_per_call": dt * 1e3,
           "terms_per_s": npt * ku / dt, "unit": "exp terms/s"}
    log("tblgen: %.2f ms per (p, T), %d rows, %.3g terms/s" % (dt * 1e3, len(u), npt * ku / dt))
    # a whole table: 40 pressures x 20 temperatures, each its own 4001-point
    # spectrum of one channel (2 cm^-1 at 0.0005 cm^-1), in one batched call
    nt_p, nt_t, m = 40, 20, 4001
    tp = np.repeat(np.logspace(np.log10(1100.0), np.log10(0.01), nt_p), nt_t)
    tt = np.tile(np.linspace(160.0, 320.0, nt_t), nt_p)
    g0, g1 = 700.0, 702.0
    gn = np.linspace(g0, g1, m)
    base = np.full(m, 1e-6)
    for c, st, w in zip(rng.uniform(g0, g1, 30), 10 ** rng.uniform(-3, 0, 30), rng.uniform(0.001, 0.01, 30)):
        base += st * w / np.pi / ((gn - c) ** 2 + w * w)
    ttau = base[None, :] * (tp / 300.0)[:, None] * (250.0 / tt)[:, None]
    tnuf = np.linspace(700.2, 701.8, 33)
    tfilt = np.exp(-0.5 * ((tnuf - 701.0) / 0.3) ** 2)
    gpu.tblgen_batch(tp, tt, ttau, g0, g1, tnuf, tfilt)  # warm
    t0 = time.perf_counter()
    trows = gpu.tblgen_batch(tp, tt, ttau, g0, g1, tnuf, tfilt)
    dtab = time.perf_counter() - t0
    out["table"] = {"workload": "%d x %d (p, T) spectra of %d points, one batched call" % (nt_p, nt_t, m),
                    "s_per_table": dtab, "rows_total": int(sum(len(r[0]) for r in trows)),
                    "spectra_per_s": nt_p * nt_t / dtab}
    log("tblgen table: %d spectra in %.1f ms (%.0f spectra/s)" % (nt_p * nt_t, dtab * 1e3, nt_p * nt_t / dtab))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        u0, e0 = O.tblgen(press, temp, tau, nu0, nu1, nuf, filt)
        dc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dc, "unit": "(p, T) rows per s", "cores": 1, "kind": "port",
                               "sample": "the same call, oracle's sequential loop, %.1f s" % dc}
        out["gpu_over_cpu"] = dc / dt
        out["rows_match"] = bool(len(u0) == len(u) and np.array_equal(u0, u))
        out["eps_max_abs_diff"] = float(np.abs(e0 - e).max()) if len(u0) == len(u) else None
        log("tblgen CPU oracle: %.1f s (GPU / CPU = %.0f)" % (dc, dc / dt))
        ns_cpu = 160
        t0 = time.perf_counter()
        for s_ in range(ns_cpu):
            O.tblgen(tp[s_ * 5], tt[s_ * 5], ttau[s_ * 5], g0, g1, tnuf, tfilt)
        dcs = (time.perf_counter() - t0) / ns_cpu
        out["table"]["cpu_baseline"] = {"value": 1.0 / dcs, "unit": "spectra/s", "cores": 1, "kind": "port",
                                        "sample": "%d of the table's spectra, oracle, %.1f s" % (ns_cpu, dcs * ns_cpu)}
        out["table"]["gpu_over_cpu"] = (nt_p * nt_t / dtab) * dcs
        log("tblgen table CPU oracle: %.1f ms per spectrum (GPU / CPU = %.0f)"
            % (dcs * 1e3, (nt_p * nt_t / dtab) * dcs))
    return out


def wide_bench(args, torch, gpu, abi):
    """2048 channels x 8 emitters (CO2, H2O, O3, CH4, N2O, CO, HNO3, F11),
    nadir view angles (i mod 49) deg, continua and refraction on: forward
    model by EGA and by CGA and adjoint Jacobians for T, H2O and O3
    (n = 243), FP32 lookups, tables ingested from device memory."""
    import numpy as np
    from jurassic_amd import scenes
    nprof = args.wide_nprof
    nu = [650.0 + i * 2050.0 / 2047.0 for i in range(NWIDE_CH)]
    t0 = time.perf_counter()
    tabs = wide_tables(torch, nu, WIDE_GASES)
    filt = boxcar_filters(nu)
    torch.cuda.synchronize()
    t_tab = time.perf_counter() - t0
    # rays: profile i looks down at (i mod 49) deg; ray order angle-major
    ang = np.arange(nprof) % 49
    order = np.argsort(ang, kind="stable").astype(np.int32)
    geom = np.zeros((nprof, 7))
    g49 = scenes.nadir_geometry(705.0, np.arange(49, dtype=float))
    for c, key in enumerate(("obsz", "obslon", "obslat", "vpz", "vplon", "vplat")):
        geom[:, 1 + c] = np.asarray(g49[key])[ang[order]]
    atms = (abi.atm_t * nprof)()
    for k in range(nprof):
        atms[k] = scenes.make_profile_atm(WIDE_GASES, iprof=k)
    dev = torch.device("cuda:0")
    atm_d = torch.frombuffer(bytearray(bytes(atms)), dtype=torch.uint8).to(dev)
    prof_d = torch.from_numpy(order).to(dev)
    geom_d = torch.from_numpy(geom).to(dev)
    nray, nd = nprof, NWIDE_CH
    rad_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    tau_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    res = {"workload": "%d atmospheres x 1 nadir view (angle i mod 49 deg) x %d channels x 8 emitters "
                       "(%s); continua and refraction on; tables 40 p x 20 T x 200 u per (channel, emitter)"
                       % (nprof, nd, ", ".join(WIDE_GASES)),
           "table_build_s": t_tab}
    for formod, name in ((1, "ega"), (0, "cga")):
        ctl = abi.make_ctl(WIDE_GASES, nu[:1], tblbase="-", tblfmt=2)
        ctl.formod = formod
        t0 = time.perf_counter()
        ctx = gpu.Context.from_device_tables(ctl, nu, tabs, filters=filt)
        torch.cuda.synchronize()
        t_ctx = time.perf_counter() - t0
        log("wide %s context: %.1f s (tables packed in HBM, source function for %d channels)"
            % (name, t_ctx, nd))

        def fwd(nr=nray):
            ctx.formod_device(nprof, atm_d.data_ptr(), nr, prof_d.data_ptr(), geom_d.data_ptr(),
                              rad_d.data_ptr(), tau_d.data_ptr(), precision=abi.JUR_FP32)
        fwd(min(nray, 4096))  # warm-up on the first rays
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms = ctx.last_timing()
        r = dict(radiances_per_s=nray * nd / dt, ms_per_pass=dt * 1e3, context_s=t_ctx,
                 kernel_ms={"raytrace": kms[0], "lookup": kms[1], "accumulate": kms[2]},
                 mean_los_points=ctx.last_mean_np())
        rad = rad_d.view(nray, nd)
        r["finite"] = bool(torch.isfinite(rad).all().item())
        log("wide %s forward FP32: %.1f ms per %d radiances (%.3g radiances/s); kernels %s"
            % (name, dt * 1e3, nray * nd, r["radiances_per_s"], r["kernel_ms"]))
        if formod == 1:
            ctl.rett_zmin, ctl.rett_zmax = -100, 100
            for ig in (1, 2):
                ctl.retq_zmin[ig], ctl.retq_zmax[ig] = -100, 100
            xm, n = gpu.state_map(ctl, atms[0])
            pmap_d = torch.zeros(nprof, dtype=torch.int32, device=dev)
            # K of all rays exceeds HBM at full size (131072 x 2048 x 243 x
            # 8 B = 522 GB): the rays go through in chunks of 4096, each
            # chunk's K written to one buffer (the caller would stream it out)
            chunk = min(nray, 4096)
            k_d = torch.empty(chunk * nd * n, dtype=torch.float64, device=dev)

            def adj(r0, nr):
                ctx.jacobian_device(nprof, atm_d.data_ptr(), nr, prof_d.data_ptr() + 4 * r0,
                                    geom_d.data_ptr() + 56 * r0, rad_d.data_ptr() + 8 * nd * r0,
                                    tau_d.data_ptr() + 8 * nd * r0, xm[None, :], pmap_d.data_ptr(), n,
                                    k_d.data_ptr(), precision=abi.JUR_FP32)
            adj(0, chunk)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            kms = [0.0, 0.0, 0.0]
            finite = True
            for r0 in range(0, nray, chunk):
                adj(r0, min(chunk, nray - r0))
                kms = [a + b for a, b in zip(kms, ctx.last_timing())]
                if r0 == 0:
                    finite = bool(torch.isfinite(k_d).all().item())
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            r["jacobian"] = dict(jacobians_per_s=nprof / dt, ms_per_pass=dt * 1e3, n=n, m=nd, chunk_rays=chunk,
                                 kernel_ms={"raytrace": kms[0], "lookup": kms[1],
                                            "reverse_sweep_and_store": kms[2]},
                                 finite=finite)
            log("wide adjoint FP32 (T, H2O, O3; n = %d): %.1f ms per %d Jacobians (%.1f Jacobians/s)"
                % (n, dt * 1e3, nprof, nprof / dt))
            del k_d
        ctx.close()
        res[name] = r
    del tabs
    torch.cuda.empty_cache()
    return res


def serving_bench(args, gpu, abi):
    """Poisson stream of mixed requests into the request server."""
    import threading
    import numpy as np
    from jurassic_amd import scenes
    rng = np.random.default_rng(2024)
    ctl = scenes.limb_3gas_ctl()
    ctl.write_bbt = 1
    nd = ctl.nd
    ret = abi.ret_t()
    ret.kernel_recomp, ret.conv_itmax, ret.conv_dmin = 3, 10, 0.1
    for i in range(nd):
        ret.err_noise[i] = math.sqrt(0.1)
    ret.err_temp, ret.err_temp_cz, ret.err_temp_ch = 2.0, 5.0, 500.0
    for ig in range(3):
        ret.err_q[ig], ret.err_q_cz[ig], ret.err_q_ch[ig] = 20.0, 5.0, 500.0
    # geometry templates: (limb | nadir) x 1..64 rays
    geo = {}
    for kind in ("limb", "nadir"):
        for nr in range(1, 65):
            if kind == "limb":
                g = scenes.limb_geometry(800.0, np.sort(rng.uniform(5.0, 67.0, nr)))
            else:
                g = scenes.nadir_geometry(705.0, np.sort(rng.uniform(0.0, 50.0, nr)))
            geo[kind, nr] = scenes.make_obs(g)
    atms = [scenes.make_profile_atm(GASES, iprof=1000 + k) for k in range(64)]
    apr = scenes.make_profile_atm(GASES, perturb=False)
    # measurements for the retrievals: brightness temperatures of the pool
    # atmospheres on every template (one batched GPU forward model)
    keys = list(geo.keys())
    ma = (abi.atm_t * len(keys))()
    mo = (abi.obs_t * len(keys))()
    for i, key in enumerate(keys):
        ma[i] = atms[i % len(atms)]
        mo[i] = geo[key]
    gpu.formod_batch(ctl, ma, mo, abi.JUR_FP32)
    meas = {key: abi.copy_struct(mo[i]) for i, key in enumerate(keys)}
    # the schedule
    nreq = int(args.serve_rate * args.serve_seconds)
    t_arr = np.cumsum(rng.exponential(1.0 / args.serve_rate, nreq))
    kinds = rng.choice([abi.JUR_REQ_FORWARD, abi.JUR_REQ_JACOBIAN, abi.JUR_REQ_RETRIEVAL], nreq,
                       p=[0.6, 0.25, 0.15])
    nrays = rng.integers(1, 65, nreq)
    geos = rng.choice(["limb", "nadir"], nreq)
    ngas = rng.integers(0, 4, nreq)
    which = rng.integers(0, len(atms), nreq)
    reqs = []
    for i in range(nreq):
        key = (str(geos[i]), int(nrays[i]))
        retq = {ig: (-100.0, 100.0) for ig in range(int(ngas[i]))}
        if kinds[i] == abi.JUR_REQ_RETRIEVAL:
            reqs.append(gpu.Server.request(int(kinds[i]), apr, meas[key], rett=(-100.0, 100.0), retq=retq))
        else:
            reqs.append(gpu.Server.request(int(kinds[i]), atms[which[i]], geo[key], rett=(-100.0, 100.0),
                                           retq=retq))
    srv = gpu.Server(ctl, ret, precision=abi.JUR_FP32, lanes=args.serve_lanes,
                     forward_lane=1 if args.serve_forward_lane else 0,
                     path_split=1 if args.serve_path_split else 0)
    # warm: tables, kernels, allocations (every lane: both geometries), then
    # the first serve_warm seconds of the schedule, untimed, so that the work
    # buffers have grown to their steady-state sizes
    for g in ("limb", "nadir"):
        for kind in (abi.JUR_REQ_FORWARD, abi.JUR_REQ_JACOBIAN, abi.JUR_REQ_RETRIEVAL):
            i = srv.submit(gpu.Server.request(kind, apr if kind == abi.JUR_REQ_RETRIEVAL else atms[0],
                                              meas[g, 8] if kind == abi.JUR_REQ_RETRIEVAL else geo[g, 8],
                                              rett=(-100.0, 100.0)))
            srv.wait(i, 300.0)
            srv.release(i)
    nwarm = int(np.searchsorted(t_arr, args.serve_warm))
    if nwarm:
        t0 = time.perf_counter()
        wid = []
        for i in range(nwarm):
            dt = t_arr[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            wid.append(srv.submit(reqs[i]))
        for i in wid:
            srv.wait(i, 600.0)
            srv.release(i)
    r0 = srv.stats()[0]
    prof0 = srv.profile()
    log("serving: %d requests over %.1f s (%.0f/s), %d forward, %d Jacobian, %d retrieval"
        % (nreq, t_arr[-1], args.serve_rate, (kinds == 0).sum(), (kinds == 1).sum(), (kinds == 2).sum()))
    ids = [None] * nreq
    lag = [0.0]

    def submitter():
        t0 = time.perf_counter()
        for i in range(nreq):
            dt = t_arr[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            else:
                lag[0] = max(lag[0], -dt)
            ids[i] = srv.submit(reqs[i])

    th = threading.Thread(target=submitter)
    t_start = time.perf_counter()
    th.start()
    lat = np.zeros(nreq)
    t_sub = np.zeros(nreq)
    t_done = np.zeros(nreq)
    status = np.zeros(nreq, dtype=int)
    niter = np.zeros(nreq, dtype=int)
    fails = {}
    done = 0
    last_log = time.perf_counter()
    for i in range(nreq):
        while ids[i] is None:
            time.sleep(0.0005)
        res = srv.wait(ids[i], 600.0)
        status[i] = res.status
        niter[i] = res.niter
        if res.status:
            key = "%d %s" % (int(kinds[i]), res.message.decode(errors="replace"))
            fails[key] = fails.get(key, 0) + 1
        t_sub[i], t_done[i] = res.t_submit, res.t_done
        lat[i] = res.t_done - res.t_submit
        srv.release(ids[i])
        done += 1
        if time.perf_counter() - last_log > 5:
            log("serving: %d/%d results" % (done, nreq))
            last_log = time.perf_counter()
    th.join()
    wall = time.perf_counter() - t_start
    rounds, max_active = srv.stats()
    prof = srv.profile()
    srv.close()
    ok = status == 0
    out = {
        "workload": "Poisson %.0f requests/s for %.1f s: 60%% forward, 25%% Jacobian, 15%% retrieval; 1..64 rays; "
                    "limb/nadir 50/50; n = 81..324; 12 channels, BT; FP32 lookups, adjoint Jacobians"
                    % (args.serve_rate, args.serve_seconds),
        "requests": nreq, "failed": int((~ok).sum()), "failures": fails,
        "throughput_req_per_s": nreq / (t_done.max() - t_sub.min()),
        "offered_req_per_s": nreq / t_arr[-1],
        "submit_lag_max_s": lag[0],
        "device_rounds": rounds - r0, "max_in_flight": max_active, "lanes": args.serve_lanes, "forward_lane": bool(args.serve_forward_lane),
        "path_split": bool(args.serve_path_split), "warm_s": args.serve_warm,
        "latency_ms": {"p50": float(np.percentile(lat[ok], 50) * 1e3), "p99": float(np.percentile(lat[ok], 99) * 1e3),
                       "max": float(lat[ok].max() * 1e3)},
        "wall_s": wall,
        "worker_seconds": {"admission": prof[0], "jacobian_passes": prof[1], "forward_passes": prof[2],
                           "dense_and_bookkeeping": prof[3]},
        "timed_allocations": {"calls": int(prof[5] - prof0[5]), "seconds": prof[4] - prof0[4],
                              "regrow_seconds": prof[6] - prof0[6]},
        "latency_ms_p99_by_second": [
            float(np.percentile(lat[ok & (t_arr >= s0) & (t_arr < s0 + 1)], 99) * 1e3)
            for s0 in range(int(np.ceil(t_arr[-1])))
            if (ok & (t_arr >= s0) & (t_arr < s0 + 1)).any()],
    }
    for k, name in ((0, "forward"), (1, "jacobian"), (2, "retrieval")):
        sel = ok & (kinds == k)
        if sel.any():
            out["latency_ms_" + name] = {"p50": float(np.percentile(lat[sel], 50) * 1e3),
                                         "p99": float(np.percentile(lat[sel], 99) * 1e3),
                                         "count": int(sel.sum())}
    for g in ("limb", "nadir"):
        for k, name in ((0, "forward"), (2, "retrieval")):
            sel = ok & (kinds == k) & (geos == g)
            if sel.any():
                out["latency_ms_%s_%s" % (name, g)] = {"p50": float(np.percentile(lat[sel], 50) * 1e3),
                                                       "p99": float(np.percentile(lat[sel], 99) * 1e3),
                                                       "count": int(sel.sum())}
    sel = ok & (kinds == 2)
    if sel.any():
        out["retrieval_mean_iterations"] = float(niter[sel].mean())
    log("serving: p50 %.1f ms, p99 %.1f ms, %.0f req/s, %d device rounds, up to %d in flight"
        % (out["latency_ms"]["p50"], out["latency_ms"]["p99"], out["throughput_req_per_s"], out["device_rounds"],
           max_active))
    return out


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nprof", type=int, default=NPROF)
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="profiles for the CPU baseline (default: 128 per thread, about 10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-jac", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-warm", type=float, default=1.0,
                    help="seconds of the request schedule run untimed before the timed run")
    ap.add_argument("--serve-lanes", type=int, default=2,
                    help="server lanes for the requests with limb rays (with --serve-path-split 1; "
                         "one more lane takes the others)")
    ap.add_argument("--serve-forward-lane", action="store_true",
                    help="one more server lane for forward requests")
    ap.add_argument("--serve-path-split", type=int, default=1,
                    help="1: --serve-lanes lanes for the requests with limb rays, one more for the others")
    ap.add_argument("--jac-only", action="store_true", help="skip the forward-model section")
    ap.add_argument("--jac-nprof", type=int, default=NJAC_PROF)
    ap.add_argument("--jac-steps", type=int, default=2)
    ap.add_argument("--jac-check", type=int, default=64,
                    help="atmospheres per angle for the FP32-vs-FP64 adjoint check")
    ap.add_argument("--jac-cpu-sample", type=int, default=2)
    ap.add_argument("--jac-fd-nprof", type=int, default=64,
                    help="atmospheres of the measured FD Jacobian run")
    ap.add_argument("--no-wide", action="store_true")
    ap.add_argument("--no-tblgen", action="store_true")
    ap.add_argument("--tblgen-only", action="store_true")
    ap.add_argument("--wide-only", action="store_true")
    ap.add_argument("--wide-nprof", type=int, default=4096)
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a 1 GiB device-to-device copy (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(profiles shard across GPUs without any exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from jurassic_amd import abi, gpu

    torch.cuda.set_device(0)
    if args.copy_calib:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del src, dst
    if args.serve_only:
        print(json.dumps({"serving": serving_bench(args, gpu, abi)}))
        return
    if args.jac_only:
        print(json.dumps({"jacobian": jacobian_bench(args, torch, gpu, abi)}))
        return
    if args.wide_only:
        print(json.dumps({"wide": wide_bench(args, torch, gpu, abi)}))
        return
    if args.tblgen_only:
        print(json.dumps({"tblgen": tblgen_bench(args, gpu)}))
        return
    ctl, atms, prof, geom = build_workload(args.nprof)
    nprof, nray, nd = args.nprof, len(prof), ctl.nd
    log("workload built: %d atmospheres, %d rays, %d channels" % (nprof, nray, nd))
    ctx = gpu.Context(ctl)
    log("device context ready (tables and source function in HBM)")

    dev = torch.device("cuda:0")
    atm_d = torch.frombuffer(bytearray(bytes(atms)), dtype=torch.uint8).to(dev)
    prof_d = torch.from_numpy(prof).to(dev)
    geom_d = torch.from_numpy(np.ascontiguousarray(geom)).to(dev)
    rad_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    tau_d = torch.empty(nray * nd, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step(precision):
        ctx.formod_device(nprof, atm_d.data_ptr(), nray, prof_d.data_ptr(), geom_d.data_ptr(),
                          rad_d.data_ptr(), tau_d.data_ptr(), precision=precision)

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        torch.cuda.synchronize()
        kms = [0.0, 0.0, 0.0]
        t0 = time.perf_counter()
        for _ in range(steps):
            step(precision)
            kms = [a + b for a, b in zip(kms, ctx.last_timing())]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt, [k / steps for k in kms]

    nrad_all = nray * nd
    dt32, k32 = timed(abi.JUR_FP32, args.steps, args.warmup)
    mean_np = ctx.last_mean_np()
    log("FP32: %.1f ms/step; kernels: ray tracing %.1f ms, lookups %.1f ms, accumulation %.1f ms; "
        "mean LOS points %.1f" % (dt32 * 1e3 / args.steps, k32[0], k32[1], k32[2], mean_np))
    res = {}
    if not args.no_fp64:
        rad32 = rad_d.clone()
        dt64, k64 = timed(abi.JUR_FP64, max(1, args.steps // 2), 1)
        res["fp64_value"] = nray * nd * max(1, args.steps // 2) / dt64
        res["fp64_kernel_ms"] = {"raytrace": k64[0], "lookup": k64[1], "accumulate": k64[2]}
        log("FP64: %.1f ms/step; kernels: ray tracing %.1f ms, lookups %.1f ms, accumulation %.1f ms"
            % (dt64 * 1e3 / max(1, args.steps // 2), k64[0], k64[1], k64[2]))
        # brightness-temperature error of the FP32 path against the FP64 path
        # on every radiance of the workload
        nu_d = torch.tensor([ctl.nu[i] for i in range(nd)], dtype=torch.float64, device=dev)
        c1, c2 = 1.19104259e-8, 1.43877506

        def bt(r):
            return c2 * nu_d / torch.log1p(c1 * nu_d ** 3 / r.view(nray, nd))
        err = torch.abs(bt(rad32) - bt(rad_d))
        res["fp32_vs_fp64_bt_K"] = {"max": float(err.max().item()), "mean": float(err.mean().item()),
                                    "radiances": nrad_all}
        log("FP32 vs FP64 brightness temperature over %d radiances: max %.2e K, mean %.2e K"
            % (nrad_all, res["fp32_vs_fp64_bt_K"]["max"], res["fp32_vs_fp64_bt_K"]["mean"]))
        del rad32, err
    nrad = nray * nd
    value = nrad * args.steps / dt32
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "radiances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "limb sounder: %d atmospheres x 32 tangent heights x 12 channels "
                        "(CO2, H2O, O3; EGA; continua and refraction on), %d radiances per step"
                        % (nprof, nrad),
            "profiles": nprof, "rays": nray, "channels": nd, "emitters": len(GASES),
            "mean_los_points": mean_np, "precision": "FP32 lookups, FP64 ray geometry",
        },
        "roofline": lookup_roofline(nprof, nray, nd, mean_np, k32),
    }
    out.update(res)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ns = args.cpu_sample or max(128, 128 * threads)
        log("CPU baseline on %d profiles, %d threads" % (ns, threads))
        out["cpu_baseline"] = cpu_baseline(ns)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    ctx.close()
    del rad_d, tau_d, atm_d
    if not args.no_jac:
        out["jacobian"] = jacobian_bench(args, torch, gpu, abi)
    if not args.no_serve:
        out["serving"] = serving_bench(args, gpu, abi)
    if not args.no_tblgen:
        out["tblgen"] = tblgen_bench(args, gpu)
    if not args.no_wide:
        try:
            torch.cuda.empty_cache()
            out["wide"] = wide_bench(args, torch, gpu, abi)
        except Exception as e:  # the wide section reports its failure instead of the line's
            out["wide"] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("wide section failed: %s" % e)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
