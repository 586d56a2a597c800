#!/usr/bin/env python3
"""Benchmark of the earthquake-cycle hot path on one MI355X.

Workload: transform fault 80 x 20 km on 128 x 32 = 4096 Okada patches
(down-dip cells graded finer toward the surface) coupled to a hex8 mantle
block of 16 x 16 x 8 = 2048 cells (12288 strain dofs), linear Maxwell
viscosity 1e19 Pa s; K = 16384^2 FP64 (2.15 GB), assembled on the device
from the mesh.  One "step" = one evaluation of the ODE right-hand side
(prologue + the fused K-streaming/friction kernel) on a state resident in
HBM.  Also reported: the FP32- and BF16-streamed operator (throughput and
error against FP64), simulated years per hour of the on-device Tsit5, the
HBM roofline of the dominant kernel, and the reference algorithm (Toeplitz
DFT K_ff + dense mantle products) timed on the host cores with the oracle.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oetqf.jl_amd"))

HBM_PEAK_GBPS = 8000.0
METRIC = "RHS evals/s & simulated yr/hr vs plain FP64 GPU at stated err; HBM GB/s"


def algorithmic_bytes(nf, ne, kbytes_per_entry):
    """Bytes one RHS must move: K once, the state it reads (v, theta, sigma),
    the derivative it writes (dv, dtheta, deps, dsigma, ddelta) and the
    per-patch/per-cell parameters."""
    nm = 6 * ne
    n = nf + nm
    return (kbytes_per_entry * n * n + 8 * (2 * nf + nm) + 8 * (3 * nf + 2 * nm)
            + 8 * (4 * nf + 2 * ne))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(w, u0, sample_seconds):
    """The reference algorithm (equation.jl:185-205: Toeplitz DFT K_ff and
    dense FP64 mantle products, OpenMP) timed on the host with the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import oetqf_amd as oq

    mf, ma, lam, mu = w["mf"], w["ma"], w["lam"], w["mu"]
    st = oq.stress_greens_function(mf, lam, mu, fourier=False)
    g12 = oq.stress_greens_function(mf, ma, lam, mu)
    g21 = oq.stress_greens_function(ma, mf, lam, mu)
    g22 = oq.stress_greens_function(ma, lam, mu, checkeigvals=False)
    v, th, sig = u0.x[0], u0.x[1], u0.x[3]
    out = oracle.rhs_coupled(w["pf"], w["pa"], st, g12, g21, g22, v, th, sig)  # warm
    n = 0
    t0 = time.perf_counter()
    while True:
        out = oracle.rhs_coupled(w["pf"], w["pa"], st, g12, g21, g22, v, th, sig)
        n += 1
        el = time.perf_counter() - t0
        if el >= sample_seconds:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model, ncpu = cpu_model()
    return dict(value=n / el, unit="RHS evals/s", cores=cores, kind="port",
                cpu_model=model, host_logical_cpus=ncpu,
                sample="%d RHS evaluations of the same workload (%.1f s)" % (n, el)), out


def cpu_extras(w_c1, seconds):
    """More of the CPU restatement beside the GPU figures (bounded samples):
    the 512-patch fault's RHS (Toeplitz DFT, the reference algorithm) and a
    70-year stretch of its cycle (C Tsit5 over the dense operator), Green's
    assembly of the 512-patch fault, and the composite model's assembly
    extrapolated from a 64-receiver column sample of each block."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from types import SimpleNamespace
    from oetqf_amd.workloads import YEAR, strike_slip_512

    out = {}
    w = strike_slip_512()
    mf, pf, u0 = w["mf"], w["pf"], w["u0"]
    t0 = time.perf_counter()
    st = oracle.gf_ff(mf, w["lam"], w["mu"])
    out["c0_assembly_s"] = time.perf_counter() - t0
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < min(3.0, seconds):
        oracle.rhs_fault(pf, st, u0.x[0], u0.x[1])
        n += 1
    out["c0_rhs_per_s"] = n / (time.perf_counter() - t0)
    nf = mf.nx * mf.nxi
    K = np.zeros((nf, nf))
    for j in range(mf.nxi):
        for l in range(mf.nxi):
            blk = st[np.abs(np.arange(mf.nx)[:, None] - np.arange(mf.nx)[None, :]), j, l]
            K[j * mf.nx:(j + 1) * mf.nx, l * mf.nx:(l + 1) * mf.nx] = blk
    yrs = 70.0
    t0 = time.perf_counter()
    _, stt, _, _ = oracle.cycle_dense(pf, None, K, nf, 0, u0.pack(), 0.0, yrs * YEAR)
    el = time.perf_counter() - t0
    out["c0_cycle"] = dict(sim_years=yrs, wall_s=el, sim_yr_per_hr=yrs / (el / 3600.0),
                           accepted=int(stt.naccept),
                           note="first 70 years (two events) of the 300-year run")
    # composite assembly: time every block on 64 receivers, scale by rows
    mf1, ma1 = w_c1["mf"], w_c1["ma"]
    sel = np.linspace(0, ma1.ne - 1, 64).astype(int)
    msub = SimpleNamespace(**{k: getattr(ma1, k)[sel] for k in
                              ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")})
    I, J = np.linspace(0, mf1.nx - 1, 16).astype(int), np.linspace(0, mf1.nxi - 1, 4).astype(int)
    fsub = SimpleNamespace(x=mf1.x[I], ax=mf1.ax[I], y=mf1.y[J], z=mf1.z[J], axi=mf1.axi[J],
                           nx=len(I), nxi=len(J), dx=mf1.dx, dep=mf1.dep, dip=mf1.dip)
    t0 = time.perf_counter()
    oracle.gf_ff(mf1, w_c1["lam"], w_c1["mu"])
    t_ff = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.gf_mf(mf1, msub, w_c1["lam"], w_c1["mu"])
    t_mf = (time.perf_counter() - t0) * ma1.ne / 64
    t0 = time.perf_counter()
    oracle.gf_fm(ma1, fsub, w_c1["lam"], w_c1["mu"])
    t_fm = (time.perf_counter() - t0) * (mf1.nx * mf1.nxi) / 64
    t0 = time.perf_counter()
    oracle.gf_mm_rows(ma1, msub, w_c1["lam"], w_c1["mu"])
    t_mm = (time.perf_counter() - t0) * ma1.ne / 64
    out["c1_assembly_s_extrapolated"] = dict(ff=t_ff, mf=t_mf, fm=t_fm, mm=t_mm,
                                             total=t_ff + t_mf + t_fm + t_mm,
                                             note="K_ff complete; the other blocks timed on 64 "
                                                  "receivers and scaled by the receiver count")
    out.update(cpu_extras_65536(oracle))
    return out


def cpu_extras_65536(oracle, nsample=16, nevals=2):
    """The reference algorithm at N = 65536 (256x64 fault + 32x32x8 mantle)
    on the host: RHS evals/s over a few evaluations, and Green's assembly
    extrapolated from row samples.  The RHS runs the exact arithmetic of
    equation.jl:185-205 (Toeplitz DFT K_ff from the real st, dense FP64
    products with the three mantle blocks); the mantle blocks (32 GB) hold
    synthetic values, since assembling them on the host would take hours --
    a dense product's cost does not depend on its values."""
    from types import SimpleNamespace
    from oetqf_amd.workloads import precision_scale_65536

    w = precision_scale_65536()
    mf, ma = w["mf"], w["ma"]
    nf, nm = mf.nx * mf.nxi, 6 * ma.ne
    out = {}
    t0 = time.perf_counter()
    st = oracle.gf_ff(mf, w["lam"], w["mu"])
    t_ff = time.perf_counter() - t0
    # Fortran-ordered blocks: the oracle's column-major views need no copy
    g12 = np.full((nm, nf), 1e3, order="F")
    g21 = np.full((nf, nm), 1e3, order="F")
    g22 = np.full((nm, nm), -1e3, order="F")
    v, th, sig = w["u0"].x[0], w["u0"].x[1], w["u0"].x[3]
    t0 = time.perf_counter()  # the blocks are already paged in by np.full
    for _ in range(nevals):
        oracle.rhs_coupled(w["pf"], w["pa"], st, g12, g21, g22, v, th, sig)
    el = time.perf_counter() - t0
    del g12, g21, g22
    out["c2_rhs_per_s"] = dict(value=nevals / el, evals=nevals, wall_s=el,
                               note="Toeplitz DFT K_ff (real st) + dense FP64 mantle blocks "
                                    "(synthetic values, 32.2 GB)")
    sel = np.linspace(0, ma.ne - 1, nsample).astype(int)
    msub = SimpleNamespace(**{k: getattr(ma, k)[sel] for k in
                              ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz")})
    I = np.linspace(0, mf.nx - 1, nsample // 4).astype(int)
    J = np.linspace(0, mf.nxi - 1, 4).astype(int)
    fsub = SimpleNamespace(x=mf.x[I], ax=mf.ax[I], y=mf.y[J], z=mf.z[J], axi=mf.axi[J],
                           nx=len(I), nxi=len(J), dx=mf.dx, dep=mf.dep, dip=mf.dip)
    t0 = time.perf_counter()
    oracle.gf_mf(mf, msub, w["lam"], w["mu"])
    t_mf = (time.perf_counter() - t0) * ma.ne / nsample
    t0 = time.perf_counter()
    oracle.gf_fm(ma, fsub, w["lam"], w["mu"])
    t_fm = (time.perf_counter() - t0) * nf / (len(I) * len(J))
    t0 = time.perf_counter()
    oracle.gf_mm_rows(ma, msub, w["lam"], w["mu"])
    t_mm = (time.perf_counter() - t0) * ma.ne / nsample
    out["c2_assembly_s_extrapolated"] = dict(
        ff=t_ff, mf=t_mf, fm=t_fm, mm=t_mm, total=t_ff + t_mf + t_fm + t_mm,
        note="K_ff complete; the other blocks timed on %d receivers and scaled by the "
             "receiver count" % nsample)
    return out


def ensemble_leg(oq, args, w, u0, value):
    """The 256-member ensemble on the composite geometry, per operator type:
    FP32 (FP32 MFMA) and the split-FP16 pair (three FP16 MFMA products per
    entry, FP32-class error).  Per type: ms per ensemble RHS, member-RHS/s
    and its ratio to the FP64 plain rate, the GEMM's useful TFLOP/s
    (2 N^2 M) and its matrix-core rate, and the error of two members against
    their own FP64 RHS.  The lockstep stepper then runs on the fastest type
    whose error is within 1e-5, over a bounded wall-clock window."""
    from oetqf_amd.workloads import YEAR, ensemble_members

    mf, ma = w["mf"], w["ma"]
    nf, ne = mf.nx * mf.nxi, ma.ne
    nn = nf + 6 * ne
    M = args.ensemble
    members, U0 = ensemble_members(w, M)
    refs = {}
    for m in (0, M - 1):
        pm = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], members[m], w["pa"], u0, w["tspan"],
                                   nrept=w["nrept"], buffer_ratio=w["buffer_ratio"])
        um = u0.similar().unpack(U0[m])
        dm = oq.similar(um)
        pm.f(dm, um, None, 0.0)
        refs[m] = dm.pack()
        pm.f.close()
    out = dict(members=M, by_kdtype={})
    best = None
    for kd in ("fp32", "fp16x2"):
        log("ensemble: %d members, %s operator" % (M, kd))
        ens = oq.Ensemble(mf, ma, w["lam"], w["mu"], members, w["pa"], nrept=w["nrept"],
                          buffer_ratio=w["buffer_ratio"], kdtype=kd)
        dU = ens.eval(U0)
        ens.time(2)
        ms_ens = ens.time(max(5, args.steps // 10))
        ens.close()
        member_rps = M * 1000.0 / ms_ens
        useful = 2.0 * nn * nn * M / (ms_ens * 1e-3) / 1e12
        errs = [float(np.linalg.norm(dU[m] - refs[m]) / np.linalg.norm(refs[m])) for m in (0, M - 1)]
        r = dict(ms_per_ensemble_rhs=ms_ens, member_rhs_per_s=member_rps,
                 speedup_vs_fp64_plain=member_rps / value, gemm_tflops=useful,
                 rel_err_vs_fp64_members=errs)
        if kd == "fp32":
            r.update(mfma_peak_tflops=157.3, mfma_frac=useful / 157.3)
        else:
            # three FP16 products per entry on the padded column space
            nfp = -(-nf // 512) * 512
            ldc = -(-(nfp + 6 * ne) // 512) * 512
            mp = -(-M // 256) * 256
            hw = 3 * 2.0 * nn * ldc * mp / (ms_ens * 1e-3) / 1e12
            r.update(mfma_fp16_tflops=hw, mfma_peak_tflops=2500.0, mfma_frac=hw / 2500.0)
        out["by_kdtype"][kd] = r
        log("ensemble %s: %.3f ms per RHS of %d members -> %.0f member-RHS/s (%.1fx FP64 plain), "
            "%.1f useful TFLOP/s, err %s" % (kd, ms_ens, M, member_rps, member_rps / value, useful,
                                             ", ".join("%.1e" % e for e in errs)))
        if max(errs) <= 1e-5 and (best is None or ms_ens < out["by_kdtype"][best]["ms_per_ensemble_rhs"]):
            best = kd
    best = best or "fp32"
    out.update(out["by_kdtype"][best])
    out["kdtype"] = best
    if args.cycle_seconds > 0:
        # the lockstep stepper from the members' interseismic start, for a
        # bounded wall-clock window
        ens = oq.Ensemble(mf, ma, w["lam"], w["mu"], members, w["pa"], nrept=w["nrept"],
                          buffer_ratio=w["buffer_ratio"], kdtype=best)
        ens.start(U0, w["tspan"], reltol=w["reltol"], abstol=w["abstol"], dt0=1.0)
        t0 = time.perf_counter()
        iters = 0
        while time.perf_counter() - t0 < args.cycle_seconds:
            left = ens.run(32)
            iters += 32
            if left == 0:
                break
        el = time.perf_counter() - t0
        tm, _, na, nr, _ = ens.state()
        ens.close()
        out["cycle"] = dict(
            kdtype=best, wall_s=el, lockstep_iterations=iters, mean_member_sim_years=float(tm.mean() / YEAR),
            member_sim_yr_per_hr=float(tm.sum() / YEAR / (el / 3600.0)),
            accepted_mean=float(na.mean()), rejection_rate=float(nr.sum() / (na.sum() + nr.sum())),
            note="all members from their interseismic start, dt0 = 1 s, Tsit5 per member")
        log("ensemble Tsit5 (%s): %d lockstep iterations in %.1f s, mean %.3f yr/member, "
            "%.0f member-yr/hr" % (best, iters, el, tm.mean() / YEAR, out["cycle"]["member_sim_yr_per_hr"]))
    return out


def scale_leg(oq, args):
    """The precision variant at scale: 256x64 fault + 32x32x8 mantle
    (N = 65536, K 34.4 GB in FP64; fft_grid: FP64 with K_ff by DFT, K_mm on
    the mantle grid and K_fm, K_mf along strike, 0.88 GB).  Per storage type: device assembly time,
    RHS rate and HBM rate, RHS error against FP64 on a perturbed state, and
    a trajectory of `scale_years` from the interseismic start with its
    event catalogue (max V > 1e-3 m/s, crossings within 1 yr merged); when
    FP64 is among the types, the others' end state and catalogue are
    compared with it."""
    from oetqf_amd.cycle import events, recurrence, run_cycle
    from oetqf_amd.workloads import YEAR, precision_scale_65536

    w = precision_scale_65536()
    mf, ma = w["mf"], w["ma"]
    nf, ne = mf.nx * mf.nxi, ma.ne
    rng = np.random.default_rng(12)
    ug = w["u0"].similar()
    ug.x[0][:] = w["pf"].vpl * np.exp(0.5 * rng.standard_normal(ug.x[0].shape))
    ug.x[1][:] = w["pf"].L / ug.x[0]
    ug.x[3][:] = 1e6 * rng.standard_normal(ug.x[3].shape)
    span = (0.0, args.scale_years * YEAR)
    out = dict(workload=w["name"], N=nf + 6 * ne, k_bytes_fp64=8 * (nf + 6 * ne) ** 2,
               trajectory_years=args.scale_years, reltol=w["reltol"], abstol=w["abstol"])
    ref = None
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    esz = {"fp64": 8, "fp32": 4, "bf16": 2}
    for kd in args.scale_kdtypes.split(","):
        t0 = time.perf_counter()
        p = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], w["pf"], w["pa"], w["u0"], span,
                                  nrept=w["nrept"], buffer_ratio=w["buffer_ratio"], kdtype=kd)
        ta = time.perf_counter() - t0
        du = oq.similar(ug)
        p.f(du, ug, None, 0.0)
        p.f.time(3)
        me, mm = p.f.time(20)
        kbytes = p.f.info()[2]
        last = [time.perf_counter()]

        def prog(t, kd=kd):
            if time.perf_counter() - last[0] > 30 and t is not None:
                last[0] = time.perf_counter()
                log("N=65536 %s: t = %.2f yr" % (kd, t / YEAR))
        r = run_cycle(p, reltol=w["reltol"], abstol=w["abstol"], chunk=2048, progress=prog)
        p.f.close()
        ev = events(r["t"], r["vmax"], 1e-3, merge_gap=YEAR)
        # dense operators: K bytes over the matvec time; the grid operator: its bytes over the RHS
        hbm = (algorithmic_bytes(nf, ne, esz[kd]) / (mm * 1e-3) if kd in esz else kbytes / (me * 1e-3)) / 1e9
        res = dict(assembly_s=ta, ms_per_rhs=me, ms_matvec=mm, rhs_per_s=1000.0 / me, operator_bytes=int(kbytes),
                   hbm_GBps=hbm,
                   trajectory_wall_s=r["wall"], sim_yr_per_hr=args.scale_years / (r["wall"] / 3600.0),
                   accepted=int(r["naccept"]), rejected=int(r["nreject"]), retcode=r["retcode"],
                   events=len(ev), onsets_yr=[e["onset"] / YEAR for e in ev],
                   peak_v=[e["peak_v"] for e in ev],
                   recurrence_yr=[float(x) / YEAR for x in recurrence(ev)])
        if kd == "fp64":
            ref = (du, r["u"], ev)
        elif ref is not None:
            res.update(rel_err_dv=rel(du.x[0], ref[0].x[0]), rel_err_dsigma=rel(du.x[3], ref[0].x[3]),
                       traj_rel_err_v=rel(r["u"].x[0], ref[1].x[0]),
                       traj_rel_err_slip=rel(r["u"].x[4], ref[1].x[4]))
            if len(ev) == len(ref[2]) and len(ev) > 0:
                res["onset_rel_err_max"] = max(abs(a["onset"] - b["onset"]) / b["onset"]
                                               for a, b in zip(ev, ref[2]))
                if len(ev) > 1:
                    rg, rc = recurrence(ev), recurrence(ref[2])
                    res["recurrence_rel_err_max"] = float(np.max(np.abs(rg - rc) / rc))
        out[kd] = res
        log("N=65536 %s: assembly %.1f s, %.3f ms/RHS (%.0f GB/s), %.1f yr in %.1f s (%d+%d), %d events" % (
            kd, ta, me, res["hbm_GBps"], args.scale_years, r["wall"], r["naccept"], r["nreject"], len(ev)))
    return out


def toeplitz_leg(oq, args):
    """Fault-only 256x64 = 16384 patches (80 x 20 km, the N = 65536 fault
    without its mantle): the RHS with K_ff kept as its Toeplitz kernel,
    applied directly and by DFT convolution (the reference's algorithm),
    against the dense FP64 operator."""
    from oetqf_amd.workloads import precision_scale_65536

    w = precision_scale_65536()
    mf, pf = w["mf"], w["pf"]
    nf = mf.nx * mf.nxi
    rng = np.random.default_rng(13)
    u = oq.ArrayPartition(pf.vpl * np.exp(0.5 * rng.standard_normal((mf.nx, mf.nxi))),
                          np.zeros((mf.nx, mf.nxi)), np.zeros((mf.nx, mf.nxi)))
    u.x[1][:] = pf.L / u.x[0]
    out = dict(workload="fault only %dx%d = %d patches" % (mf.nx, mf.nxi, nf))
    ref = None
    for kd in ("fp64", "toeplitz", "fft"):
        p = oq.assemble_from_mesh(mf, None, w["lam"], w["mu"], pf, None, u, (0.0, 1.0), kdtype=kd)
        du = oq.similar(u)
        p.f(du, u, None, 0.0)
        p.f.time(3)
        me, mm = p.f.time(args.steps)
        p.f.close()
        # bytes of the operator the RHS reads: dense K, st, or G (the real
        # DFT of the circulant at the L/2 + 1 bins of the power-of-two length
        # L >= 2nx - 1)
        L = 2
        while L < 2 * mf.nx - 1:
            L *= 2
        kb = {"fp64": 8 * nf * nf, "toeplitz": 8 * mf.nx * mf.nxi * mf.nxi,
              "fft": 8 * (L // 2 + 1) * mf.nxi * mf.nxi}[kd]
        out[kd] = dict(ms_per_rhs=me, rhs_per_s=1000.0 / me, operator_bytes=kb)
        if ref is None:
            ref = du
        else:
            out[kd]["rel_err_dv_vs_dense"] = float(np.linalg.norm(du.x[0] - ref.x[0]) /
                                                   np.linalg.norm(ref.x[0]))
    out["speedup_vs_dense"] = out["fp64"]["ms_per_rhs"] / out["toeplitz"]["ms_per_rhs"]
    out["fft_speedup_vs_dense"] = out["fp64"]["ms_per_rhs"] / out["fft"]["ms_per_rhs"]
    if args.cycle_seconds > 0:
        # the fault-only cycle on the DFT operator (Tsit5 graph stepper) from
        # the workload's interseismic state, over a bounded wall-clock window
        from oetqf_amd.workloads import YEAR
        u0 = oq.ArrayPartition(w["u0"].x[0].copy(), w["u0"].x[1].copy(), w["u0"].x[4].copy())
        for kd in ("fft", "fp64"):
            p = oq.assemble_from_mesh(mf, None, w["lam"], w["mu"], pf, None, u0, w["tspan"], kdtype=kd)
            it = oq.Integrator(p, reltol=w["reltol"], abstol=w["abstol"])
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < min(args.cycle_seconds, 6.0):
                if it.step(256) != "Running":
                    break
            el = time.perf_counter() - t0
            t, _, _, na, nr = it.state()
            it.close()
            p.f.close()
            out[kd]["cycle"] = dict(sim_years=t / YEAR, wall_s=el, sim_yr_per_hr=(t / YEAR) / (el / 3600.0),
                                    accepted=int(na), rejected=int(nr))
            log("fault-only %d patches, %s Tsit5: %.3f yr in %.1f s (%.0f yr/hr)" %
                (nf, kd, t / YEAR, el, out[kd]["cycle"]["sim_yr_per_hr"]))
    log("fault-only %d patches: dense FP64 %.4f ms, Toeplitz %.4f ms (%.1fx, err %.1e), DFT %.4f ms "
        "(%.1fx, err %.1e) per RHS" % (
            nf, out["fp64"]["ms_per_rhs"], out["toeplitz"]["ms_per_rhs"], out["speedup_vs_dense"],
            out["toeplitz"]["rel_err_dv_vs_dense"], out["fft"]["ms_per_rhs"], out["fft_speedup_vs_dense"],
            out["fft"]["rel_err_dv_vs_dense"]))
    return out


def group_leg(oq, args):
    """Uneven-size grouped batch: 24 faults of 1024-8192 patches (half with
    a mantle of 512-4096 cells), 16 realizations each, FP32 operators."""
    from oetqf_amd.workloads import YEAR, grouped_batch

    ws = grouped_batch()
    t0 = time.perf_counter()
    grp = oq.Group([(w["mf"], w["ma"], w["pa"]) for w in ws], ws[0]["lam"], ws[0]["mu"],
                   [w["members"] for w in ws])
    t_asm = time.perf_counter() - t0
    U0 = grp.pack([w["U0"] for w in ws])
    dU = grp.eval(U0)
    grp.time(3)
    ms_default = grp.time(20)
    if args.group_shape:  # pinned (profiling runs): no tuning launches
        from oetqf_amd.group import SHAPES
        grp.set_shapes({b: args.group_shape for b in range(8)})
        def bucket(n):
            b = 0
            while b + 1 < 8 and n >= (2048 << b):
                b += 1
            return b
        choice = {bucket(int(n)): (args.group_shape, 0.0) for n in grp.N}
        grp.tune_table = {}
        assert args.group_shape in SHAPES
    else:
        choice = grp.autotune(5)
    grp.eval(U0)
    grp.time(3)
    ms = grp.time(args.steps)
    GM = grp.G * grp.M
    out = dict(geometries=grp.G, members_per_geometry=grp.M, N_min=int(min(grp.N)),
               N_max=int(max(grp.N)), k_bytes=grp.kbytes, assembly_s=t_asm,
               ms_per_group_rhs=ms, ms_per_group_rhs_default_shape=ms_default,
               member_rhs_per_s=GM * 1000.0 / ms, hbm_GBps=grp.kbytes / (ms * 1e-3) / 1e9,
               hbm_frac=grp.kbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
               tile_shape_by_bucket={"N<%d" % (2048 << b): c[0] for b, c in choice.items()},
               autotune_ms={"N<%d" % (2048 << b): t for b, t in grp.tune_table.items()})
    # error of the first member of the smallest and the largest geometry
    # against its own FP64 RHS
    Us, dUs = grp.unpack(U0), grp.unpack(dU)
    errs = {}
    for g in (int(np.argmin(grp.N)), int(np.argmax(grp.N))):
        w = ws[g]
        p = oq.assemble_from_mesh(w["mf"], w["ma"], w["lam"], w["mu"], w["members"][0], w["pa"],
                                  w["u0"], w["tspan"])
        u = w["u0"].similar().unpack(Us[g][0])
        d = oq.similar(u)
        p.f(d, u, None, 0.0)
        p.f.close()
        ref = d.pack()
        errs[w["name"]] = float(np.linalg.norm(dUs[g][0] - ref) / np.linalg.norm(ref))
    out["rel_err_vs_fp64"] = errs
    log("group: %d geometries x %d members, K %.2f GB: %.3f ms per RHS (default shape %.3f ms), "
        "%.0f member-RHS/s, %.0f GB/s; shapes %s" % (grp.G, grp.M, grp.kbytes / 1e9, ms, ms_default,
                                                    out["member_rhs_per_s"], out["hbm_GBps"],
                                                    out["tile_shape_by_bucket"]))
    if args.cycle_seconds > 0:
        grp.start(U0, (0.0, 300 * YEAR), dt0=1.0)
        t0 = time.perf_counter()
        iters = 0
        while time.perf_counter() - t0 < args.cycle_seconds:
            left = grp.run(32)
            iters += 32
            if left == 0:
                break
        el = time.perf_counter() - t0
        tm, _, na, nr, _ = grp.state()
        rej = nr / np.maximum(na + nr, 1)
        out["stepping"] = dict(wall_s=el, lockstep_iterations=iters,
                               mean_member_sim_years=float(tm.mean() / YEAR),
                               member_sim_yr_per_hr=float(tm.sum() / YEAR / (el / 3600.0)),
                               rejection_rate_mean=float(rej.mean()),
                               rejection_rate_max=float(rej.max()),
                               rejection_rate_overall=float(nr.sum() / (na.sum() + nr.sum())))
        log("group Tsit5: %d iterations in %.1f s, mean %.2f yr/member, rejection rate mean %.3f "
            "max %.3f" % (iters, el, tm.mean() / YEAR, rej.mean(), rej.max()))
    grp.close()
    return out


# ---------------- full simulated spans with event statistics ----------------

def _catalogue(t, vmax):
    from oetqf_amd.cycle import events, recurrence
    from oetqf_amd.workloads import YEAR
    ev = events(t, vmax, 1e-3, merge_gap=YEAR)
    rec = recurrence(ev) / YEAR
    return dict(events=len(ev), onsets_yr=[e["onset"] / YEAR for e in ev],
                peak_v=[e["peak_v"] for e in ev], recurrence_yr=[float(x) for x in rec],
                mean_recurrence_yr=float(rec.mean()) if len(rec) else None)


def _compare(cat, ref):
    """Event statistics of a run against the FP64 reference run: the same
    event count, and the recurrence intervals' relative error (interval by
    interval where the counts agree, and of the mean); and how many leading
    events agree (onsets within 0.1 % of each other, i.e. before the
    trajectories part: the cycle is chaotic, and two FP64 methods part as
    well), with the time the first disagreeing event starts."""
    out = dict(same_event_count=cat["events"] == ref["events"])
    on_a, on_b = cat["onsets_yr"], ref["onsets_yr"]
    k = 0
    while k < min(len(on_a), len(on_b)) and abs(on_a[k] - on_b[k]) <= 1e-3 * on_b[k]:
        k += 1
    out["leading_events_agree"] = k
    out["max_leading_onset_rel_err"] = max([abs(on_a[i] - on_b[i]) / on_b[i] for i in range(k)], default=0.0)
    if k < min(len(on_a), len(on_b)):
        out["first_difference_yr"] = min(on_a[k], on_b[k])
    if ref["mean_recurrence_yr"] and cat["mean_recurrence_yr"]:
        out["mean_recurrence_rel_err"] = abs(cat["mean_recurrence_yr"] - ref["mean_recurrence_yr"]) / \
            ref["mean_recurrence_yr"]
    if out["same_event_count"] and cat["recurrence_yr"]:
        a, b = np.array(cat["recurrence_yr"]), np.array(ref["recurrence_yr"])
        out["max_interval_rel_err"] = float(np.max(np.abs(a - b) / b))
        on_a, on_b = np.array(cat["onsets_yr"]), np.array(ref["onsets_yr"])
        out["max_onset_rel_err"] = float(np.max(np.abs(on_a - on_b) / on_b))
    return out


def _progress(tag):
    last = [time.perf_counter()]
    from oetqf_amd.workloads import YEAR

    def f(t):
        if t is not None and time.perf_counter() - last[0] > 20:
            last[0] = time.perf_counter()
            log("%s: t = %.1f yr" % (tag, t / YEAR))
    return f


def _load_refs(paths):
    """Legs of earlier spans JSON files (comma-separated paths), merged."""
    out = {}
    for path in [x for x in (paths or "").split(",") if x]:
        with open(path) as fh:
            for leg, v in json.load(fh).items():
                if leg == "composite" and leg in out:
                    out[leg]["runs"].update(v["runs"])
                elif leg == "ensemble" and leg in out:
                    for kd, e in v["by_kdtype"].items():
                        o = out[leg]["by_kdtype"].setdefault(kd, dict(members={}, sampled_catalogues={}))
                        o.setdefault("members", {}).update(e.get("members", {}))
                        o.setdefault("sampled_catalogues", {}).update(e.get("sampled_catalogues", {}))
                else:
                    out[leg] = v
    return out


def spans_composite(oq, runs, refs, save=None):
    """The composite workload (128x32 transform fault + 16x16x8 mantle) over
    its whole 500-year span from the seeded interseismic start, one run per
    (operator type, method) in `runs` (FP64 Tsit5, FP64 VCABM5, FP32 VCABM5,
    FP32 Tsit5 by default); event catalogues (max v > 1e-3 m/s, crossings
    within 1 yr merged) against the FP64 run of the same method, from this
    call or from an earlier spans JSON."""
    from oetqf_amd.cycle import run_cycle
    from oetqf_amd.workloads import YEAR, transform_fault_composite

    w = transform_fault_composite()
    out = dict(workload=w["name"], sim_years=(w["tspan"][1] - w["tspan"][0]) / YEAR,
               reltol=w["reltol"], abstol=w["abstol"], runs={})
    prior = refs.get("composite", {}).get("runs", {})
    for kd, alg in runs:
        p = oq.assemble_from_mesh(w["mf"], w["ma"], w["lam"], w["mu"], w["pf"], w["pa"], w["u0"],
                                  w["tspan"], nrept=w["nrept"], buffer_ratio=w["buffer_ratio"], kdtype=kd)
        tag = "%s %s" % (kd, alg)
        log("composite 500 yr: %s" % tag)
        r = run_cycle(p, reltol=w["reltol"], abstol=w["abstol"], chunk=4096,
                      alg=oq.VCABM5() if alg == "vcabm5" else oq.Tsit5(), progress=_progress(tag))
        p.f.close()
        cat = _catalogue(r["t"], r["vmax"])
        res = dict(wall_s=r["wall"], sim_yr_per_hr=out["sim_years"] / (r["wall"] / 3600.0),
                   accepted=int(r["naccept"]), rejected=int(r["nreject"]),
                   rejection_rate=float(r["nreject"] / max(r["naccept"] + r["nreject"], 1)),
                   retcode=r["retcode"], **cat)
        ref = out["runs"].get("fp64 " + alg) or prior.get("fp64 " + alg)
        if kd != "fp64" and ref is not None:
            res["vs_fp64"] = _compare(cat, ref)
            res["speedup_vs_fp64"] = ref["wall_s"] / r["wall"]
        out["runs"][tag] = res
        log("%s: %d events in %.0f s (%.0f yr/hr), %d accepted, %d rejected, mean recurrence %s yr%s" % (
            tag, cat["events"], r["wall"], res["sim_yr_per_hr"], res["accepted"], res["rejected"],
            cat["mean_recurrence_yr"], ", vs FP64 %s" % res["vs_fp64"] if "vs_fp64" in res else ""))
        if save is not None:
            save(out)
    return out


def _span_range(rng, n):
    a, b = (int(x) for x in (rng or "0:%d" % n).split(":"))
    assert 0 <= a < b <= n, "range %s outside 0:%d" % (rng, n)
    return a, b


def spans_ensemble(oq, kdtypes, sampled, member_range="", wall=None, method="tsit5"):
    """Members a:b (default all 256) of the composite-geometry ensemble over
    300 years with the lockstep Tsit5 or VCABM5 (per-member steps), per
    operator type: rejection rates, event counts and mean recurrence of
    every member, and the whole catalogues of every member that finished
    (for the members leg).  A member's trajectory does not depend on which
    other members share its batch (each GEMM output column and each
    member's scales and controller are its own), so a span can run as
    several member ranges; with a wall limit the members still running
    report how far they got."""
    from oetqf_amd.cycle import run_ensemble_cycle
    from oetqf_amd.workloads import YEAR, ensemble_members, transform_fault_composite

    w = transform_fault_composite()
    members, U0 = ensemble_members(w, 256)
    lo, hi = _span_range(member_range, 256)
    span = (0.0, 300 * YEAR)
    out = dict(workload=w["name"] + ", 256 frictional realizations", sim_years=300.0, method=method,
               reltol=w["reltol"], abstol=w["abstol"], by_kdtype={})
    for kd in kdtypes:
        log("ensemble 300 yr: %s, %s, members %d:%d" % (kd, method, lo, hi))
        ens = oq.Ensemble(w["mf"], w["ma"], w["lam"], w["mu"], members[lo:hi], w["pa"], nrept=w["nrept"],
                          buffer_ratio=w["buffer_ratio"], kdtype=kd)
        last = [time.perf_counter()]

        def progress(iters, left):
            if time.perf_counter() - last[0] > 20:
                last[0] = time.perf_counter()
                log("ensemble %s: %d iterations, %d members running" % (kd, iters, left))
        r = run_ensemble_cycle(ens, U0[lo:hi], span, reltol=w["reltol"], abstol=w["abstol"], dt0=1.0,
                               progress=progress, wall_limit=wall,
                               alg=oq.VCABM5() if method == "vcabm5" else oq.Tsit5())
        ens.close()
        M = hi - lo
        if not r["complete"]:  # stopped by the wall limit: where every member got to
            tend = r["t_end"] / YEAR
            out["by_kdtype"][kd] = dict(member_range=[lo, hi], complete=False, wall_s=r["wall"],
                                        lockstep_iterations=r["iters"], t_reached_yr_min=float(tend.min()),
                                        t_reached_yr_mean=float(tend.mean()),
                                        members_done=int(np.sum(tend >= 300.0)))
            log("ensemble %s: stopped after %.0f s at %d iterations; t reached min %.1f / mean %.1f yr, %d done" % (
                kd, r["wall"], r["iters"], tend.min(), tend.mean(), int(np.sum(tend >= 300.0))))
        cats = [_catalogue(r["t"][m], r["vmax"][m]) for m in range(M)]
        na, nr = r["naccept"], r["nreject"]
        def params(p):  # the member's scalars: L, sigma and the velocity-weakening a - b
            vw = p.b > p.a
            return dict(L_mm=float(p.L.flat[0]) * 1e3, sigma_MPa=float(p.σ.flat[0]) / 1e6,
                        a_minus_b=float((p.a - p.b)[vw].flat[0]) if vw.any() else 0.0)
        members_out = {str(lo + m): dict(accepted=int(na[m]), rejected=int(nr[m]), events=cats[m]["events"],
                                         mean_recurrence_yr=cats[m]["mean_recurrence_yr"],
                                         t_reached_yr=float(r["t_end"][m] / YEAR),
                                         complete=bool(r["t_end"][m] >= span[1]), **params(members[lo + m]))
                       for m in range(M)}
        if not r["complete"]:
            done_cats = {str(lo + m): cats[m] for m in range(M) if members_out[str(lo + m)]["complete"]}
            out["by_kdtype"][kd].update(members=members_out, sampled_catalogues=done_cats)
            continue
        out["by_kdtype"][kd] = dict(
            member_range=[lo, hi], complete=True, wall_s=r["wall"], lockstep_iterations=r["iters"],
            member_sim_yr_per_hr=M * 300.0 / (r["wall"] / 3600.0),
            all_success=all(x == "Success" for x in r["retcode"]),
            rejection_rate=float(nr.sum() / (na.sum() + nr.sum())),
            members=members_out, sampled_catalogues={str(lo + m): cats[m] for m in range(M)})
        log("ensemble %s: %d x 300 yr in %.0f s (%.0f member-yr/hr), rejection rate %.3f" % (
            kd, M, r["wall"], out["by_kdtype"][kd]["member_sim_yr_per_hr"], out["by_kdtype"][kd]["rejection_rate"]))
    return out


def spans_members(oq, members_fp64, ens, ref_path, method="tsit5", save=None):
    """Sampled ensemble members run alone in FP64 (Tsit5, the same
    tolerances and start) over the same 300 years; their catalogues are
    compared with the ensemble's catalogues of those members, read from the
    spans JSON of an ensemble leg."""
    from oetqf_amd.cycle import run_cycle
    from oetqf_amd.workloads import YEAR, ensemble_members, transform_fault_composite

    w = transform_fault_composite()
    members, U0 = ensemble_members(w, 256)
    span = (0.0, 300 * YEAR)
    out = dict(sim_years=300.0, ensemble_spans=ref_path or None, members={})
    for m in members_fp64:
        u0 = w["u0"].similar().unpack(U0[m])
        p = oq.assemble_from_mesh(w["mf"], w["ma"], w["lam"], w["mu"], members[m], w["pa"], u0, span,
                                  nrept=w["nrept"], buffer_ratio=w["buffer_ratio"])
        log("member %d alone, FP64 %s, 300 yr" % (m, method))
        r = run_cycle(p, reltol=w["reltol"], abstol=w["abstol"], chunk=4096, dt=1.0,
                      progress=_progress("member %d fp64" % m),
                      alg=oq.VCABM5() if method == "vcabm5" else oq.Tsit5())
        p.f.close()
        ref = _catalogue(r["t"], r["vmax"])
        res = dict(wall_s=r["wall"], accepted=int(r["naccept"]), rejected=int(r["nreject"]), **ref)
        for kd, e in (ens or {}).items():
            cat = e.get("sampled_catalogues", {}).get(str(m))
            if cat is not None:
                res["ensemble_%s_vs_fp64" % kd] = _compare(cat, ref)
            elif str(m) in e.get("members", {}):  # the ensemble kept only this member's summary
                s_ = e["members"][str(m)]
                c = dict(same_event_count=s_["events"] == ref["events"])
                if s_["mean_recurrence_yr"] and ref["mean_recurrence_yr"]:
                    c["mean_recurrence_rel_err"] = abs(s_["mean_recurrence_yr"] - ref["mean_recurrence_yr"]) / \
                        ref["mean_recurrence_yr"]
                c["events"] = [s_["events"], ref["events"]]
                res["ensemble_%s_vs_fp64" % kd] = c
        out["members"][str(m)] = res
        if save is not None:
            save(out)
        log("member %d: FP64 %d events, mean recurrence %s yr; %s" % (
            m, ref["events"], ref["mean_recurrence_yr"],
            "; ".join("%s %s" % (k[9:-8], v) for k, v in res.items() if k.startswith("ensemble_"))))
    return out


def spans_group(oq, geo_range="", shape="", wall=None, method="tsit5"):
    """Geometries a:b (default all 24) of the grouped batch (16 members each,
    FP32) over 300 years with the lockstep grouped Tsit5 or VCABM5 (method):
    per-member step-rejection rates and step counts.  With a pinned tile shape every
    member's trajectory is independent of which geometries share the run, so
    a span can run as several geometry ranges."""
    from oetqf_amd.workloads import YEAR, grouped_batch

    ws = grouped_batch()
    lo, hi = _span_range(geo_range, len(ws))
    ws = ws[lo:hi]
    grp = oq.Group([(w["mf"], w["ma"], w["pa"]) for w in ws], ws[0]["lam"], ws[0]["mu"],
                   [w["members"] for w in ws])
    if shape:
        grp.set_shapes({b: shape for b in range(8)})
        shapes = shape
    else:
        shapes = {b: c[0] for b, c in grp.autotune(3).items()}
    grp.start(grp.pack([w["U0"] for w in ws]), (0.0, 300 * YEAR), dt0=1.0,
              alg=oq.VCABM5() if method == "vcabm5" else oq.Tsit5())
    t0 = time.perf_counter()
    iters, last = 0, t0
    complete = False
    while True:
        left = grp.run(256)
        iters += 256
        if time.perf_counter() - last > 20:
            last = time.perf_counter()
            t, _, _, _, _ = grp.state()
            log("group 300 yr: %d iterations, %d members running, min t = %.1f yr" % (
                iters, left, t.min() / YEAR))
        if left == 0:
            complete = True
            break
        if wall is not None and time.perf_counter() - t0 > wall:
            break
    wall_s = time.perf_counter() - t0
    t, U, na, nr, rc = grp.state()
    grp.close()
    rate = nr / np.maximum(na + nr, 1)
    G = hi - lo
    tyr = t / YEAR
    return dict(geometry_range=[lo, hi], members_per_geometry=16, sim_years=300.0, wall_s=wall_s, method=method,
                complete=complete, members_done=int(np.sum(tyr >= 300.0)), t_reached_yr_min=float(tyr.min()),
                t_reached_yr_mean=float(tyr.mean()),
                lockstep_iterations=iters, member_sim_yr_per_hr=float(tyr.sum()) / (wall_s / 3600.0),
                all_success=all(x == "Success" for x in rc), tile_shapes=shapes,
                rejection_rate_mean=float(rate.mean()), rejection_rate_max=float(rate.max()),
                by_geometry={"%d: %s" % (lo + g, ws[g]["name"]): dict(index=lo + g, rejection_rate=float(rate[g * 16:(g + 1) * 16].mean()),
                                                 accepted_mean=float(na[g * 16:(g + 1) * 16].mean()),
                                                 rejected_mean=float(nr[g * 16:(g + 1) * 16].mean()),
                                                 t_reached_yr_min=float(tyr[g * 16:(g + 1) * 16].min()),
                                                 members_done=int(np.sum(tyr[g * 16:(g + 1) * 16] >= 300.0)))
                             for g in range(G)})


def spans_summary(paths):
    """Markdown tables of the span JSON files (comma-separated): composite
    runs, ensemble member ranges, members run alone in FP64, grouped
    geometry ranges."""
    lines = []
    runs, ens, alone, grp = {}, {}, {}, []
    for path in [x for x in paths.split(",") if x]:
        with open(path) as fh:
            d = json.load(fh)
        runs.update(d.get("composite", {}).get("runs", {}))
        for kd, e in d.get("ensemble", {}).get("by_kdtype", {}).items():
            ens.setdefault(kd, []).append(e)
        alone.update(d.get("members", {}).get("members", {}))
        if "group" in d:
            grp.append(d["group"])
    if runs:
        lines += ["| composite 500 yr | wall s | sim yr/hr | accepted | rejected | rejection rate | events | "
                  "mean recurrence yr | vs FP64 (same method) |", "|---|---:|---:|---:|---:|---:|---:|---:|---|"]
        for tag, r in runs.items():
            vs = r.get("vs_fp64")
            vs_s = "" if vs is None else ("same event count" if vs["same_event_count"] else "event count differs") + \
                "".join(", %s %.1e" % (k, v) for k, v in vs.items() if k != "same_event_count")
            lines.append("| %s | %.0f | %.0f | %d | %d | %.3f | %d | %s | %s |" % (
                tag, r["wall_s"], r["sim_yr_per_hr"], r["accepted"], r["rejected"],
                r.get("rejection_rate", r["rejected"] / max(1, r["accepted"] + r["rejected"])), r["events"],
                "%.2f" % r["mean_recurrence_yr"] if r["mean_recurrence_yr"] else "-", vs_s))
        lines.append("")
        pairs = [(a, b) for a in runs for b in runs if a < b]
        if pairs:
            lines += ["| composite pair | events | leading events that agree (onsets within 0.1 %) | max onset error "
                      "over them | first difference at yr | mean recurrence, relative difference |",
                      "|---|---|---:|---:|---:|---:|"]
            for a, b in pairs:
                c = _compare(runs[a], runs[b])
                ma, mb = runs[a]["mean_recurrence_yr"], runs[b]["mean_recurrence_yr"]
                lines.append("| %s vs %s | %d / %d | %d | %.1e | %s | %.3f |" % (
                    a, b, runs[a]["events"], runs[b]["events"], c["leading_events_agree"],
                    c["max_leading_onset_rel_err"], "%.1f" % c["first_difference_yr"]
                    if "first_difference_yr" in c else "-", abs(ma - mb) / mb if ma and mb else float("nan")))
            lines.append("")
    for kd, parts in ens.items():
        mem = {}
        for e in parts:
            mem.update(e.get("members", {}))
        lines += ["| ensemble %s, members | wall s | lockstep iterations | members at 300 yr | t reached (yr) min / mean | "
                  "member-yr/hr |" % kd, "|---|---:|---:|---:|---|---:|"]
        for e in parts:
            ms = [v for k, v in e.get("members", {}).items()]
            tr = [v.get("t_reached_yr", 300.0) for v in ms] or [300.0]
            lines.append("| %d:%d | %.0f | %d | %d of %d | %.1f / %.1f | %.0f |" % (
                e["member_range"][0], e["member_range"][1], e["wall_s"], e["lockstep_iterations"],
                sum(1 for v in ms if v.get("complete", True)), len(ms), min(tr), float(np.mean(tr)),
                sum(tr) / (e["wall_s"] / 3600.0)))
        lines.append("")
        done = [v for v in mem.values() if v.get("complete", True)]
        if done:
            na = sum(v["accepted"] for v in done)
            nr = sum(v["rejected"] for v in done)
            ev = [v["events"] for v in done]
            rec = [v["mean_recurrence_yr"] for v in done if v["mean_recurrence_yr"]]
            lines.append("%d members over the whole 300 yr: rejection rate %.3f; events per member min %d, median %d, "
                         "max %d; mean recurrence median %.1f yr." % (len(done), nr / max(1, na + nr), min(ev),
                                                                       int(np.median(ev)), max(ev),
                                                                       float(np.median(rec)) if rec else float("nan")))
            slow = sorted(((v.get("t_reached_yr", 300.0), k, v) for k, v in mem.items() if not v.get("complete", True)))
            if slow:
                lines.append("Members short of 300 yr when the wall limit stopped the run (t reached, accepted + rejected "
                             "steps, L, sigma, a - b): " + "; ".join(
                                 "%s: %.1f yr, %d + %d, %.1f mm, %.0f MPa, %.4f" % (
                                     k, t, v["accepted"], v["rejected"], v.get("L_mm", float("nan")),
                                     v.get("sigma_MPa", float("nan")), v.get("a_minus_b", float("nan")))
                                 for t, k, v in slow[:12]) + ("; ..." if len(slow) > 12 else "") + ".")
                lo_l = [v["L_mm"] for v in mem.values() if "L_mm" in v and not v.get("complete", True)]
                hi_l = [v["L_mm"] for v in done if "L_mm" in v]
                if lo_l and hi_l:
                    lines.append("L of the members short of 300 yr: median %.1f mm (max %.1f); of the finished ones: "
                                 "median %.1f mm (min %.1f)." % (float(np.median(lo_l)), max(lo_l),
                                                                  float(np.median(hi_l)), min(hi_l)))
            lines.append("")
    if alone:
        kds = sorted({k[9:-8] for r in alone.values() for k in r if k.startswith("ensemble_")})
        lines += ["| member | FP64 alone: wall s | accepted | rejected | events | mean recurrence yr | " +
                  " | ".join("ensemble %s vs FP64" % k for k in kds) + " |",
                  "|---:|---:|---:|---:|---:|---:|" + "---|" * len(kds)]
        for m, r in sorted(alone.items(), key=lambda x: int(x[0])):
            cells = []
            for k in kds:
                c = r.get("ensemble_%s_vs_fp64" % k)
                if c is None:
                    cells.append("-")
                    continue
                txt = "same events" if c["same_event_count"] else "events differ"
                if "events" in c:
                    txt += " (%d vs %d)" % tuple(c["events"])
                if "leading_events_agree" in c:
                    txt += ", %d leading agree (onsets within %.1e)" % (c["leading_events_agree"],
                                                                       c["max_leading_onset_rel_err"])
                if "first_difference_yr" in c:
                    txt += ", part at %.1f yr" % c["first_difference_yr"]
                if "mean_recurrence_rel_err" in c:
                    txt += ", mean recurrence %.1e" % c["mean_recurrence_rel_err"]
                cells.append(txt)
            lines.append("| %s | %.0f | %d | %d | %d | %s | %s |" % (
                m, r["wall_s"], r["accepted"], r["rejected"], r["events"],
                "%.2f" % r["mean_recurrence_yr"] if r["mean_recurrence_yr"] else "-", " | ".join(cells)))
        lines.append("")
    if grp:
        lines += ["| grouped geometries | wall s | lockstep iterations | member-yr/hr | rejection rate mean / max | whole span |",
                  "|---|---:|---:|---:|---|---|"]
        for g in grp:
            lines.append("| %d:%d | %.0f | %d | %.0f | %.3f / %.3f | %s |" % (
                g["geometry_range"][0], g["geometry_range"][1], g["wall_s"], g["lockstep_iterations"],
                g["member_sim_yr_per_hr"], g["rejection_rate_mean"], g["rejection_rate_max"],
                "yes" if g.get("complete", True) and g["all_success"] else
                "%d of %d members at 300 yr (min t %.1f yr)" % (g.get("members_done", 0), 16 * (
                    g["geometry_range"][1] - g["geometry_range"][0]), g.get("t_reached_yr_min", 0.0))))
        lines += ["", "| geometry | rejection rate | accepted (mean) | rejected (mean) | members at 300 yr | "
                  "min t reached (yr) |", "|---|---:|---:|---:|---:|---:|"]
        for g in grp:
            for name, r in sorted(g["by_geometry"].items(), key=lambda x: x[1]["index"]):
                lines.append("| %s | %.3f | %.0f | %.0f | %s | %s |" % (
                    name, r["rejection_rate"], r["accepted_mean"], r["rejected_mean"],
                    r.get("members_done", "-"), "%.1f" % r["t_reached_yr_min"] if "t_reached_yr_min" in r else "-"))
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cycle-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="members of the FP32 MFMA ensemble leg (0 = skip)")
    ap.add_argument("--no-cycle512", dest="cycle512", action="store_false",
                    help="skip the 300-year run of the 512-patch fault")
    ap.add_argument("--no-scale", dest="scale", action="store_false",
                    help="skip the N = 65536 precision-variant leg")
    ap.add_argument("--no-group", dest="group", action="store_false",
                    help="skip the uneven-size grouped batch leg")
    ap.add_argument("--group-shape", default="",
                    help="pin this grouped GEMM tile shape for every size bucket instead of autotuning")
    ap.add_argument("--scale-years", type=float, default=1.0,
                    help="trajectory length of the N = 65536 leg")
    ap.add_argument("--scale-kdtypes", default="fp64,fp32,bf16,fft_grid",
                    help="operator storage types of the N = 65536 leg (fp64 first: the reference)")
    ap.add_argument("--spans", default="",
                    help="instead of the bench line: whole simulated spans with event statistics, "
                         "comma-separated from composite, ensemble, members, group (JSON to --spans-out)")
    ap.add_argument("--spans-out", default="spans.json", help="where --spans writes its JSON")
    ap.add_argument("--spans-members", default="0,37,64,101,128,170,203,255",
                    help="ensemble members sampled by --spans ensemble and run alone in FP64 by --spans members")
    ap.add_argument("--spans-ref", default="",
                    help="spans JSON files of earlier calls (comma-separated): the FP64 composite runs and "
                         "the ensemble catalogues that this call's runs are compared with")
    ap.add_argument("--spans-summary", default="",
                    help="print the markdown tables of these span JSON files (comma-separated) and exit")
    ap.add_argument("--spans-wall", type=float, default=0.0,
                    help="wall-clock seconds after which an ensemble span stops and reports how far it got")
    ap.add_argument("--spans-method", default="tsit5", choices=["tsit5", "vcabm5"],
                    help="method of the ensemble, members and group spans")
    ap.add_argument("--spans-range", default="",
                    help="a:b, the ensemble members (of 256) or grouped-batch geometries (of 24) of this call")
    ap.add_argument("--spans-runs", default="fp64:tsit5,fp64:vcabm5,fp32:vcabm5,fp32:tsit5",
                    help="(operator type:method) runs of the composite leg")
    ap.add_argument("--spans-kdtypes", default="fp16x2,fp32",
                    help="operator types of the ensemble spans")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (known traffic) for counter scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; the sharding "
              "design is row blocks of K with an all-gather of s (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import oetqf_amd as oq
    from oetqf_amd.workloads import YEAR, transform_fault_composite

    if args.spans_summary:
        spans_summary(args.spans_summary)
        return
    if args.spans:
        res = {}
        refs = _load_refs(args.spans_ref)
        os.makedirs(os.path.dirname(os.path.abspath(args.spans_out)), exist_ok=True)

        def save(partial=None, leg=None):
            if leg is not None:
                res[leg] = partial
            with open(args.spans_out, "w") as fh:
                json.dump(res, fh, indent=1)

        sampled = [int(m) for m in args.spans_members.split(",") if m]
        for leg in args.spans.split(","):
            if leg == "composite":
                runs = [tuple(x.split(":")) for x in args.spans_runs.split(",") if x]
                res[leg] = spans_composite(oq, runs, refs, save=lambda o: save(o, "composite"))
            elif leg == "ensemble":
                res[leg] = spans_ensemble(oq, args.spans_kdtypes.split(","), sampled, args.spans_range,
                                          args.spans_wall or None, args.spans_method)
            elif leg == "members":
                ref = refs.get("ensemble", {}).get("by_kdtype")
                res[leg] = spans_members(oq, sampled, ref, args.spans_ref, args.spans_method,
                                         save=lambda o: save(o, "members"))
            elif leg == "group":
                res[leg] = spans_group(oq, args.spans_range, args.group_shape, args.spans_wall or None,
                                       args.spans_method)
            else:
                raise SystemExit("unknown span leg %r" % leg)
            save()
        log("spans written to %s" % args.spans_out)
        return

    if args.copy_calibration:
        # a device copy of known size (reads 1 GiB, writes 1 GiB) through the
        # HIP runtime the library is linked against: the reference for
        # scaling the FETCH_SIZE / WRITE_SIZE counters of the other kernels
        import ctypes
        oq.load_library()
        rt = ctypes.CDLL("libamdhip64.so.7")
        n = ctypes.c_size_t(1 << 30)
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(a), n) == 0 and rt.hipMalloc(ctypes.byref(b), n) == 0
        assert rt.hipMemset(a, 0, n) == 0
        for _ in range(3):
            assert rt.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
        assert rt.hipDeviceSynchronize() == 0
        rt.hipFree(a)
        rt.hipFree(b)

    w = transform_fault_composite()
    mf, ma = w["mf"], w["ma"]
    nf, ne = mf.nx * mf.nxi, ma.ne
    u0 = w["u0"]
    rng = np.random.default_rng(11)
    # a generic interseismic state: perturbed slip rates, nonzero mantle stress
    u0.x[0][:] = w["pf"].vpl * np.exp(0.5 * rng.standard_normal(u0.x[0].shape))
    u0.x[1][:] = w["pf"].L / u0.x[0]
    u0.x[3][:] = 1e6 * rng.standard_normal(u0.x[3].shape)

    log("assembling K on the device (%s)" % w["name"])
    t0 = time.perf_counter()
    prob = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], w["pf"], w["pa"], u0, w["tspan"],
                                 nrept=w["nrept"], buffer_ratio=w["buffer_ratio"])
    t_asm = time.perf_counter() - t0
    f = prob.f
    du64 = oq.similar(u0)
    f(du64, u0, None, 0.0)  # uploads the state; the timed loop reuses it in HBM
    log("assembly %.2f s; timing %d RHS evaluations" % (t_asm, args.steps))
    f.time(args.warmup)
    t0 = time.perf_counter()
    ms_eval, ms_mv = f.time(args.steps)
    wall = time.perf_counter() - t0
    value = 1000.0 / ms_eval
    nbytes = algorithmic_bytes(nf, ne, 8)
    achieved = nbytes / (ms_mv * 1e-3) / 1e9
    log("FP64: %.4f ms/eval, matvec %.4f ms, %.0f GB/s" % (ms_eval, ms_mv, achieved))

    variants = {}
    if not args.no_variants:
        for kd, esz in (("fp32", 4), ("bf16", 2)):
            p = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], w["pf"], w["pa"], u0, w["tspan"],
                                      nrept=w["nrept"], buffer_ratio=w["buffer_ratio"], kdtype=kd)
            du = oq.similar(u0)
            p.f(du, u0, None, 0.0)
            p.f.time(args.warmup)
            me, mm = p.f.time(args.steps)
            err_tau = float(np.linalg.norm(du.x[0] - du64.x[0]) / np.linalg.norm(du64.x[0]))
            err_sig = float(np.linalg.norm(du.x[3] - du64.x[3]) / np.linalg.norm(du64.x[3]))
            variants[kd] = dict(rhs_per_s=1000.0 / me, ms_matvec=mm,
                                hbm_GBps=algorithmic_bytes(nf, ne, esz) / (mm * 1e-3) / 1e9,
                                rel_err_dv=err_tau, rel_err_dsigma=err_sig,
                                speedup_vs_fp64=ms_eval / me)
            log("%s: %.4f ms/eval (%.2fx), err dv %.2e dsigma %.2e" % (kd, me, ms_eval / me, err_tau, err_sig))
            p.f.close()

    if not args.no_variants:
        # fft_coupled: the coupled model with K_ff by DFT convolution beside the dense FP64 mantle
        # blocks, as the reference's coupled ode applies it (equation.jl:200); fft_grid: the same
        # with K_mm as 2D convolutions over the mantle grid (all FP64)
        for name, kd in (("fft_coupled", "fft"), ("fft_grid", "fft_grid")):
            p = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], w["pf"], w["pa"], u0, w["tspan"],
                                      nrept=w["nrept"], buffer_ratio=w["buffer_ratio"], kdtype=kd)
            du = oq.similar(u0)
            p.f(du, u0, None, 0.0)
            p.f.time(args.warmup)
            me, _ = p.f.time(args.steps)
            errs = [float(np.linalg.norm(du.x[k] - du64.x[k]) / np.linalg.norm(du64.x[k])) for k in range(5)]
            _, _, kb = p.f.info()
            variants[name] = dict(rhs_per_s=1000.0 / me, ms_per_rhs=me, operator_bytes=int(kb),
                                  rel_err_vs_dense_fp64=max(errs), speedup_vs_fp64=ms_eval / me)
            log("coupled, %s: %.4f ms/eval (%.2fx dense FP64), operator %.3f GB, max rel diff %.1e" % (
                kd, me, ms_eval / me, kb / 1e9, max(errs)))
            p.f.close()
        variants["toeplitz_fault_only"] = toeplitz_leg(oq, args)

    ensemble = None
    if args.ensemble > 0:
        ensemble = ensemble_leg(oq, args, w, u0, value)

    cycle = None
    if args.cycle_seconds > 0:
        cycle = {}
        from oetqf_amd.cycle import events, run_cycle
        ev = {}
        for name, alg in (("tsit5", oq.Tsit5()), ("vcabm5", oq.VCABM5())):
            r = run_cycle(prob, reltol=w["reltol"], abstol=w["abstol"], alg=alg, chunk=1024,
                          wall_limit=args.cycle_seconds)
            t, el = float(r["t"][-1]), r["wall"]
            ev[name] = events(r["t"], r["vmax"], 1e-3, merge_gap=YEAR)
            cycle[name] = dict(sim_years=t / YEAR, wall_s=el, sim_yr_per_hr=(t / YEAR) / (el / 3600.0),
                               accepted=int(r["naccept"]), rejected=int(r["nreject"]),
                               rhs_per_step=6 if name == "tsit5" else 2, events=len(ev[name]),
                               onsets_yr=[e["onset"] / YEAR for e in ev[name]])
            log("%s: %.3f yr in %.1f s (%d+%d steps), %d events" % (name, t / YEAR, el, r["naccept"],
                                                                   r["nreject"], len(ev[name])))
        # VCABM5 on the FP32 operator, and on the FP64 grid operator, over the span the FP64 run
        # reached; their event catalogues against the FP64 one ("at stated error")
        t64 = cycle["vcabm5"]["sim_years"] * YEAR
        for name, kd, label in (("vcabm5_fp32", "fp32", "FP32 K"), ("vcabm5_fft_grid", "fft_grid", "grid K_mm")):
            pv = oq.assemble_from_mesh(mf, ma, w["lam"], w["mu"], w["pf"], w["pa"], u0, (0.0, t64),
                                       nrept=w["nrept"], buffer_ratio=w["buffer_ratio"], kdtype=kd)
            r = run_cycle(pv, reltol=w["reltol"], abstol=w["abstol"], alg=oq.VCABM5(), chunk=1024)
            pv.f.close()
            t, el = float(r["t"][-1]), r["wall"]
            ev_, e64 = events(r["t"], r["vmax"], 1e-3, merge_gap=YEAR), ev["vcabm5"]
            cycle[name] = dict(
                sim_years=t / YEAR, wall_s=el, sim_yr_per_hr=(t / YEAR) / (el / 3600.0),
                accepted=int(r["naccept"]), rejected=int(r["nreject"]),
                speedup_vs_fp64=(t / YEAR) / (el / 3600.0) / cycle["vcabm5"]["sim_yr_per_hr"],
                events=len(ev_), onsets_yr=[e["onset"] / YEAR for e in ev_],
                onset_rel_err_vs_fp64=[abs(a["onset"] - b["onset"]) / b["onset"] for a, b in zip(ev_, e64)],
                peak_v_rel_err_vs_fp64=[abs(a["peak_v"] - b["peak_v"]) / b["peak_v"] for a, b in zip(ev_, e64)])
            log("vcabm5 (%s): %.3f yr in %.1f s (%.2fx FP64), %d events (FP64 %d), onset err %s" % (
                label, t / YEAR, el, cycle[name]["speedup_vs_fp64"], len(ev_), len(e64),
                ["%.1e" % x for x in cycle[name]["onset_rel_err_vs_fp64"]]))
        cycle.update(reltol=w["reltol"], abstol=w["abstol"],
                     note="interseismic start of the cycle, FP64 K; each method over the same "
                          "wall-clock window (events: max v > 1e-3 m/s, crossings within 1 yr "
                          "merged); vcabm5_fp32 = the FP32 operator, vcabm5_fft_grid = the FP64 "
                          "operator with K_ff by DFT and K_mm on the mantle grid, each over the FP64 "
                          "VCABM5 span")
        cycle["sim_yr_per_hr"] = cycle["tsit5"]["sim_yr_per_hr"]

    cycle512 = None
    if args.cycle512:
        # the reference's CPU-runnable case: 300 years of the 512-patch fault
        from oetqf_amd.cycle import events, run_cycle
        from oetqf_amd.workloads import strike_slip_512
        w0 = strike_slip_512()
        p0 = oq.assemble_from_mesh(w0["mf"], None, w0["lam"], w0["mu"], w0["pf"], None, w0["u0"],
                                   w0["tspan"])
        yrs = (w0["tspan"][1] - w0["tspan"][0]) / YEAR
        cycle512 = dict(workload=w0["name"], sim_years=yrs, reltol=w0["reltol"], abstol=w0["abstol"],
                        event_def="max v > 1e-3 m/s, crossings within 1 yr merged")
        for name, alg in (("tsit5", None), ("vcabm5", oq.VCABM5())):
            r0 = run_cycle(p0, reltol=w0["reltol"], abstol=w0["abstol"], persistent="auto", alg=alg)
            ev = events(r0["t"], r0["vmax"], 1e-3, merge_gap=365 * 86400.0)
            cycle512[name] = dict(wall_s=r0["wall"], sim_yr_per_hr=yrs / (r0["wall"] / 3600.0),
                                  accepted=int(r0["naccept"]), rejected=int(r0["nreject"]),
                                  events=len(ev), retcode=r0["retcode"], stepper=r0["mode"])
            log("512-patch cycle (%s, %s): %.0f yr in %.1f s (%.0f yr/hr), %d events" %
                (name, r0["mode"], yrs, r0["wall"], cycle512[name]["sim_yr_per_hr"], len(ev)))
        cycle512["sim_yr_per_hr"] = max(cycle512["tsit5"]["sim_yr_per_hr"],
                                        cycle512["vcabm5"]["sim_yr_per_hr"])
        p0.f.close()

    scale = None
    if args.scale:
        scale = scale_leg(oq, args)

    group = None
    if args.group:
        group = group_leg(oq, args)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, reference algorithm)")
        cpu, ref = cpu_baseline(w, u0, args.cpu_seconds)
        cpu["extra"] = cpu_extras(w, args.cpu_seconds)
        def headline(v):
            if not isinstance(v, dict):
                return v
            return next((v[k] for k in ("total", "sim_yr_per_hr", "value") if k in v), None)
        log("CPU extras: %s" % json.dumps({k: headline(v) for k, v in cpu["extra"].items()}))
        cpu["rhs_rel_err_gpu_vs_cpu"] = float(np.linalg.norm(du64.x[0] - ref[0]) / np.linalg.norm(ref[0]))

    if ensemble and ensemble.get("cycle") and cycle:
        ensemble["cycle"]["vs_fp64_plain_sim_yr_per_hr"] = (
            ensemble["cycle"]["member_sim_yr_per_hr"] / cycle["tsit5"]["sim_yr_per_hr"])

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_rhs_matvec_fp64_transform_fault_4096.traffic.txt")
    if os.path.exists(tfile):
        for ln in open(tfile):
            if ln.startswith("traffic_bytes_per_launch:"):
                traffic = float(ln.split(":")[1])

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_eval,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded state; Green's functions assembled from the mesh)",
        "config": {"workload": w["name"], "n_fault": nf, "n_cells": ne, "N": nf + 6 * ne,
                   "k_bytes": 8 * (nf + 6 * ne) ** 2, "assembly_s": t_asm,
                   "wall_s_timed_loop": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_source": ("not measured in this run: read from the committed counter "
                                        "summary profiles/k_rhs_matvec_fp64_transform_fault_4096."
                                        "traffic.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of "
                                        "this bench, scaled by a 1 GiB copy)") if traffic else None,
                     "kernel": "k_rhs_matvec<FP64>", "ms_per_launch": ms_mv,
                     "algorithmic_bytes": nbytes},
        "variants": variants,
        "ensemble": ensemble,
        "cycle": cycle,
        "cycle_512": cycle512,
        "scale_65536": scale,
        "grouped_batch": group,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    # release the device handles while the runtime (and any profiler attached
    # to it) is still up, instead of from finalizers during interpreter exit
    prob.f.close()


if __name__ == "__main__":
    main()
