#!/usr/bin/env python3
"""Benchmark of the Stokes-matvec FMM (the layer potential GMRES applies every iteration).

Workload (the FMM-alone configuration): 2^20 boundary points on 4096 closed circles with
centres U[0,1]^2 and radii U[0.01,0.05], 256 points each, complex CN(0,1) densities from a
fixed seed; kernel = CADAPDR's GW (DR signs), FP64, evaluated at every point.

One step = one complete FMM evaluation from raw device-resident coordinates: tree build
(binning, sort, leaf table) + P2M, M2M, M2L+L2L, L2P, P2P.  `value` is points per second of
that whole evaluation (Mpts/s).  The apply-only rate (tree kept, as inside one GMRES solve)
is reported beside it, with every stage's time, FP64 fraction and GB/s.  LEVMAX and p are the
FMM autotuner's choice for 1e-10 relative accuracy (2-norm against the O(N^2) sum at 1024
targets) unless given.  The accuracy of the timed configuration is measured against the
tree-free O(N^2) sum on the GPU at 1024 seeded targets, and in the CPU leg against the
long-double sum of the C restatement at the same targets.  The same measurement of the second
FMM workload (2^20 uniform points) and the fixed LEVMAX 8 / p = 32 point are reported beside it.

Prints ONE JSON line.  Single GPU only.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP64 peak of MI355X (vector = matrix) and HBM3E bandwidth
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the dominant kernels, from two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) of this script, scaled by the same counters on a 256 MiB copy (fetch x 1.99968,
# write x 1.000); profiles/fmm_curves_n1048576_L8_p36_hbm_traffic.md has the commands and the
# per-kernel table.  Keyed by (workload, points, levmax, nterms).
PROFILED_TRAFFIC_BYTES = {
    ("curves", 1 << 20, 8, 36): {"m2l": 218904752, "p2p": 390217048},
}


_T0 = time.perf_counter()


def log(msg):
    """Progress on stderr (the one JSON line goes to stdout)."""
    sys.stderr.write(f"bench [{time.perf_counter() - _T0:7.1f} s] {msg}\n")
    sys.stderr.flush()


def leg(name, fn, *a):
    log(f"{name}: start")
    t = time.perf_counter()
    r = fn(*a)
    log(f"{name}: done in {time.perf_counter() - t:.1f} s")
    return r


STAGES = ("charges", "p2m", "m2m", "m2l", "l2l", "l2p", "p2p")


def m2l_flops(split, p):
    """M2L flops: 8 p p_in + 17 (p + p_in) per pair (two series; prescaling per input term,
    post-scaling per output term) with p_in = p for the 12 nearest interaction-list sources and
    p_in = the shorter input length of the 15 farther ones (m2l_split)."""
    pf = split["m2l_far_terms"]
    return ((8.0 * p * p + 34.0 * p) * split["m2l_near"]
            + (8.0 * p * pf + 17.0 * (p + pf)) * split["m2l_far"])


def flops_model(info, n, p):
    """Algorithmic flop convention: complex add 2, complex*real 2, complex*complex 6,
    reciprocal 6, divide 12; per stage counts from the tree.  M2M and L2L count the nonzero
    entries of their triangular Pascal matrices, p (p + 1) / 2 per series."""
    return {
        "charges": 0.0,
        "p2p": 50.0 * info["near_pairs"],
        "m2l": m2l_flops(info, p),
        "p2m": 26.0 * p * n,
        "l2p": 26.0 * p * n,
        "m2m": (4.0 * p * (p + 1) + 30.0 * p) * info["boxes"],
        "l2l": (4.0 * p * (p + 1) + 30.0 * p) * info["boxes"],
    }


def bytes_model(info, n, p):
    """Algorithmic HBM bytes per stage (each datum read or written once): points carry x, y (16 B),
    the weighted tangent (16 B), the density (16 B) and GW (16 B); an expansion of one box is
    2 series x p complex coefficients (32 p B).
      charges: density, weight, permutation in; two charges out            68 B per point
      P2M:     sorted points and charges in; one multipole per leaf out    48 B per point + 32 p per leaf
      M2M:     every child multipole in, every parent out                 64 p per box
      M2L:     every source multipole in, every target local out          64 p per box
      L2L:     parent local in, child local read and written               96 p per box
      L2P:     own and parent locals per leaf, point in, GW read and written 56 B per point + 64 p per leaf
      P2P:     points and charges in, GW out                               64 B per point"""
    leaves, boxes = info["leaves"], info["boxes"]
    return {
        "charges": 68.0 * n,
        "p2m": 48.0 * n + 32.0 * p * leaves,
        "m2m": 64.0 * p * boxes,
        "m2l": 64.0 * p * boxes,
        "l2l": 96.0 * p * boxes,
        "l2p": 56.0 * n + 64.0 * p * leaves,
        "p2p": 64.0 * n,
    }


def stage_table(stage_ms, fl, by):
    """Per stage: time, algorithmic GFLOP and TFLOP/s with the fraction of the FP64 peak, and
    algorithmic GB and GB/s with the fraction of the HBM peak."""
    out = {}
    for k in STAGES:
        t = stage_ms[k] * 1e-3
        out[k] = {"ms": round(stage_ms[k], 4), "gflop": round(fl[k] / 1e9, 4),
                  "tflops": round(fl[k] / t / 1e12, 3) if t > 0 else None,
                  "fp64_frac": round(fl[k] / t / 1e12 / PEAK_FP64_TFLOPS, 4) if t > 0 else None,
                  "gbytes": round(by[k] / 1e9, 4), "gbps": round(by[k] / t / 1e9, 1) if t > 0 else None,
                  "hbm_frac": round(by[k] / t / 1e9 / PEAK_HBM_GBS, 4) if t > 0 else None}
    return out


def direct_check(z, dz, nd, h, dens, gw, ntgt=1024, seed=7):
    """Part of the CPU leg: the relative 2-norm difference between the GPU's GW and the
    long-double O(N^2) sum of the C restatement over all sources, at ntgt random targets."""
    import oracle
    tgt = np.sort(np.random.default_rng(seed).choice(z.size, ntgt, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h, tgt)
    return {"rel_err2": float(np.linalg.norm(gw[tgt] - ref) / np.linalg.norm(ref)), "targets": int(ntgt),
            "sources": int(z.size), "checker": "long-double O(N^2) sum (oracle.direct)",
            "s": round(time.perf_counter() - t0, 2)}


def cpu_baseline(z, dz, nd, h, dens, levmax, p, ncurves_total, budget_s=20.0):
    """Time the reference CADAPDR (compiled in place from its sources into oracle/_ref) or, when
    that library is absent, the C restatement, on a bounded sample of the same workload: the
    first S curves with LEVMAX reduced so that points per leaf box match."""
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    tol = 2.0 ** (-p)
    m = int(nd[0])
    S = 256
    best = None
    while True:
        npts = S * m
        lev = max(2, levmax - int(round(np.log(ncurves_total / S) / np.log(4.0))))
        args = (oracle.DR, lev, nd[:S], z.real[:npts], z.imag[:npts], dz[:npts], dens[:npts], h[:S], tol)
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_cadap(*args)
        else:
            oracle.cadap(*args)
        dt = time.perf_counter() - t0
        best = (npts, lev, dt)
        if dt > budget_s / 4 or S * 4 > ncurves_total:
            break
        S *= 4
    npts, lev, dt = best
    return {"value": round(npts / dt / 1e6, 4), "unit": "Mpts/s", "cores": 1, "kind": kind,
            "sample": f"one CADAPDR call (DR, TOL 2^-{p}, LEVMAX {lev}) on the first {npts} points "
                      f"({npts // m} curves) of the same workload, {dt:.2f} s on 1 host core"}


def stokes_solve_leg(torch, dev, with_cpu):
    """The Stokes solve of one F evaluation (SOLVEI, stokesbubb.f:294-415) on the many-drop
    configuration (64 drops, about 28K points): DGMRES with MAXL 100 and TOL 1e-10 on the GPU
    matvec (FMM with the reference's LEVMAX 7 and p = 43).  The right-hand side is A u for a
    smooth density u (the SETVELF right-hand side is not part of this build)."""
    from electricbubbles_amd import gmres, spectral
    from electricbubbles_amd.workloads import many_drops_cfg
    nd, rku = many_drops_cfg(64, np.random.default_rng(20260506))
    B = spectral.Bodies(nd)
    geo = B.extractdat(torch.tensor(rku, device=dev))
    mv = spectral.MatVec(B, geo["x"], geo["y"], geo["dz"], geo["dsda"], geo["rkappa"], geo["zk"], levmax=7,
                         nterms=43, mode=0)
    th = np.concatenate([2 * np.pi * np.arange(m) / m for m in nd])
    u_true = torch.tensor(np.stack([np.cos(th) + 0.3 * np.sin(2 * th), 0.5 * np.sin(th)], 1).ravel(), device=dev)
    rhs = mv.apply(u_true)
    solver = gmres.GMRES(rhs.numel(), 100)
    gmres.stokes_solve(mv, solver, rhs)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        soln, za, pk, it, err, ierr = gmres.stokes_solve(mv, solver, rhs)
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(10):
        mv.apply(soln)
    torch.cuda.synchronize()
    t_mv = (time.perf_counter() - t0) / 10
    res = float(torch.linalg.norm(mv.apply(soln) - rhs) / torch.linalg.norm(rhs))
    out = {"workload": f"many-drop configuration: 64 drops, {int(nd.sum())} points, {2 * int(nd.sum())} unknowns",
           "levmax": 7, "nterms": 43, "maxl": 100, "tol": 1e-10, "iterations": it, "ierr": ierr,
           "ms_per_solve": round(t_solve * 1e3, 3), "solves_per_s": round(1.0 / t_solve, 3),
           "ms_per_matvec": round(t_mv * 1e3, 4), "rel_residual": res}
    if with_cpu:
        import tempfile
        import oracle
        if oracle.ref_available():
            g = {k: geo[k].cpu().numpy() for k in ("x", "y", "dz", "dsda", "rkappa", "zk")}
            cwd = os.getcwd()
            os.chdir(tempfile.mkdtemp())
            try:
                t0 = time.perf_counter()
                xr, itr, er, ierr_r = oracle.ref_dgmres_stokes(nd, g["x"] + 1j * g["y"], g["dz"], g["dsda"],
                                                               g["rkappa"], g["zk"], rhs.cpu().numpy())
                t_ref = time.perf_counter() - t0
            finally:
                os.chdir(cwd)
            x = soln.cpu().numpy()
            out["cpu_reference"] = {"s_per_solve": round(t_ref, 3), "iterations": itr, "ierr": ierr_r, "cores": 1,
                                    "kind": "reference",
                                    "solution_rel_diff": float(np.linalg.norm(x - xr) / np.linalg.norm(xr))}
    solver.close()
    return out


def spectral_leg(torch, dev, with_cpu):
    """Per-body geometry and stage work of F for the batched-sweep shapes: 512 single drops
    r = 1 + 0.05 cos 2 theta, N log-uniform in 128..2048 (even), all bodies in one launch per
    stage.  Times EXTRACTDAT (one forward FFT and spectral derivatives per body), the F tail
    (EQUIT + frame velocity + FFT, FILTER, inverse FFT), one forward DCFFTF of every body and
    the two RK2 stage updates, each averaged over repeated calls.  Algorithmic bytes per point:
    EXTRACTDAT reads 16 B and writes 104 B (x, y, dz, tha, dsda, rkappa, ckern); the F tail
    reads 48 B and writes 16 B; DCFFTF reads and writes 16 B; an RK2 stage reads 3 and writes
    2 vectors of 2 doubles."""
    from electricbubbles_amd import spectral
    from electricbubbles_amd.workloads import perturbed_circle
    rng = np.random.default_rng(20260707)
    nd = (np.round(np.exp(rng.uniform(np.log(128), np.log(2048), 512)) / 2) * 2).astype(np.int32)
    rku = np.concatenate([perturbed_circle(int(n), [(2, 0.05)])[0] for n in nd]).view(np.float64).copy()
    npts = int(nd.sum())
    B = spectral.Bodies(nd)
    r = torch.tensor(rku, device=dev)

    def timed(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    geo = B.extractdat(r)
    un = torch.tensor(0.05 * rng.normal(size=npts), device=dev)
    c = torch.tensor(rng.normal(size=npts) + 1j * rng.normal(size=npts), device=dev)
    dy = torch.tensor(rng.normal(size=2 * npts), device=dev)
    t_ex = timed(lambda: B.extractdat(r))
    t_tail = timed(lambda: B.ftail(geo["tha"], un, geo["dz"], geo["dsda"]))
    t_fft = timed(lambda: B.dcfft(c, -1))
    phi0, y = spectral.rk2_stage1(1e-3, r, dy)
    t_rk = timed(lambda: spectral.rk2_stage1(1e-3, r, dy))
    gbs = lambda nbytes, t: round(nbytes / t / 1e9, 1)  # noqa: E731
    out = {"workload": f"512 drops, N log-uniform 128..2048, {npts} points",
           "extractdat_ms": round(t_ex * 1e3, 4), "extractdat_GBps": gbs(120 * npts, t_ex),
           "ftail_ms": round(t_tail * 1e3, 4), "ftail_GBps": gbs(64 * npts, t_tail),
           "dcfft_ms": round(t_fft * 1e3, 4), "dcfft_GBps": gbs(32 * npts, t_fft),
           "rk2_stage_ms": round(t_rk * 1e3, 4), "rk2_stage_GBps": gbs(80 * npts, t_rk),
           "hbm_peak_GBps": PEAK_HBM_GBS}
    if with_cpu:
        import oracle
        if oracle.ref_available():
            t0 = time.perf_counter()
            oracle.ref_extractdat(nd, rku)
            t_ref = time.perf_counter() - t0
            out["cpu_reference_extractdat_ms"] = round(t_ref * 1e3, 3)
            out["cpu_reference_cores"] = 1
    return out


def sweep_leg(torch, dev, with_cpu):
    """Batched versus plain Stokes solves at 1e-10 relative (SOLVEI's DGMRES, MAXL 100, TOL
    1e-10, no restart) for the batched-sweep shapes: 512 independent single-drop simulations,
    r = 1 + 0.05 cos 2 theta, N log-uniform in 128..2048 rounded to even.  The right-hand
    side of each is A u for a smooth density u.  Batched: all 512 systems in one batched GMRES
    over the sweep operator (each simulation's O(N^2) layer potential).  Plain: the
    per-simulation GPU path (one MatVec, one GMRES workspace, one solve per simulation, the
    direct sum that EB_MODE_AUTO picks at N <= 2048), timed on every 16th simulation and
    reported per solve.  The reference leg runs SOLVEI's solve of a few simulations on one
    host core.  The "F" part does the same for whole F evaluations (electric.f:968-1109: the
    Laplace solve, the Dirichlet-to-Neumann map, SETVELF, SOLVEI, INTERFACENEW and the tail),
    with E0 = |E|^2, |E| uniform in 0.1..2 and Ca (CAP, the imposed straining flow) log-uniform in
    0.01..0.5 per simulation."""
    from electricbubbles_amd import gmres, spectral
    from electricbubbles_amd.workloads import perturbed_circle
    rng = np.random.default_rng(20260512)
    nsim = 512
    nd = (np.round(np.exp(rng.uniform(np.log(128), np.log(2048), nsim)) / 2) * 2).astype(np.int32)
    rku = np.concatenate([perturbed_circle(int(n), [(2, 0.05)])[0] for n in nd]).view(np.float64).copy()
    B = spectral.Bodies(nd)
    geo = B.extractdat(torch.tensor(rku, device=dev))
    sw = spectral.Sweep(B, np.arange(nsim + 1), geo["x"], geo["y"], geo["dz"], geo["dsda"], geo["rkappa"], geo["zk"])
    th = np.concatenate([2 * np.pi * np.arange(m) / m for m in nd])
    u_true = torch.tensor(np.stack([np.cos(th) + 0.3 * np.sin(2 * th), 0.5 * np.sin(th)], 1).ravel(), device=dev)
    rhs = sw.apply(u_true)
    gb = gmres.GMRESBatch(sw.unknown_offsets, 100)
    gmres.sweep_solve(sw, gb, rhs)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        soln, za, pk, it, err, ierr = gmres.sweep_solve(sw, gb, rhs)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(5):
        sw.apply(soln)
    torch.cuda.synchronize()
    t_apply = (time.perf_counter() - t0) / 5
    res = float(torch.linalg.norm(sw.apply(soln) - rhs) / torch.linalg.norm(rhs))
    pairs = float((nd.astype(np.float64) * (nd - 1)).sum())
    # plain path on every 16th simulation
    off = np.concatenate([[0], np.cumsum(nd)])
    sample = list(range(0, nsim, 16))
    t_plain, it_plain, worst = 0.0, [], 0.0
    for s in sample:
        p0, p1 = off[s], off[s + 1]
        Bs = spectral.Bodies(nd[s:s + 1])
        mv = spectral.MatVec(Bs, geo["x"][p0:p1], geo["y"][p0:p1], geo["dz"][p0:p1], geo["dsda"][p0:p1],
                             geo["rkappa"][p0:p1], geo["zk"][s:s + 1], mode=0)
        solver = gmres.GMRES(2 * int(p1 - p0), 100)
        r = rhs[2 * p0:2 * p1].contiguous()
        gmres.stokes_solve(mv, solver, r)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs, _, _, its, _, _ = gmres.stokes_solve(mv, solver, r)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        it_plain.append(its)
        worst = max(worst, float(torch.linalg.norm(xs - soln[2 * p0:2 * p1]) / torch.linalg.norm(xs)))
        solver.close()
        mv.close()
    per_plain = t_plain / len(sample)
    out = {"workload": f"{nsim} independent single drops, N log-uniform 128..2048 (even), {int(nd.sum())} points",
           "maxl": 100, "tol": 1e-10, "iterations_mean": round(float(it.mean()), 2), "iterations_max": int(it.max()),
           "converged": int((ierr == 0).sum()),
           "batched_ms": round(t_batch * 1e3, 3), "batched_solves_per_s": round(nsim / t_batch, 1),
           "plain_ms_per_solve": round(per_plain * 1e3, 3), "plain_solves_per_s": round(1.0 / per_plain, 1),
           "batched_vs_plain": round((nsim / t_batch) * per_plain, 2),
           "plain_sample": f"every 16th simulation ({len(sample)} solves)",
           "plain_vs_batched_solution_rel_diff": worst,
           "sweep_apply_ms": round(t_apply * 1e3, 4),
           "sweep_apply_TFLOPs": round(50.0 * pairs / t_apply / 1e12, 2), "rel_residual": res}
    # whole F evaluations (electric.f:968-1109) of every simulation: batched and plain
    from electricbubbles_amd import field
    e0 = rng.uniform(0.1, 2.0, nsim) ** 2  # E0 = |E|^2, |E| uniform in 0.1..2
    # Ca: the reference's CAP, the strength of the imposed straining flow (stokesbubb.f:63-67,
    # electric.f:869), log-uniform in 0.01..0.5 per simulation
    cap = np.exp(rng.uniform(np.log(0.01), np.log(0.5), nsim))
    rku_t = torch.tensor(rku, device=dev)
    SF = field.SweepField(B, np.arange(nsim + 1), cap, e0)
    SF.rhs(rku_t)  # warm-up (first geometry)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fb = SF.rhs(rku_t)
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / reps
    lap_it, _, st_it, _ = SF.iterations()
    t_fp, worst_f = 0.0, 0.0
    for s in sample:
        p0, p1 = off[s], off[s + 1]
        Fs = field.Field(spectral.Bodies(nd[s:s + 1]), cap=float(cap[s]), e0=float(e0[s]))
        r = rku_t[2 * p0:2 * p1].contiguous()
        Fs.rhs(r)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fp = Fs.rhs(r)
        torch.cuda.synchronize()
        t_fp += time.perf_counter() - t0
        worst_f = max(worst_f, float(torch.linalg.norm(fp - fb[2 * p0:2 * p1]) / torch.linalg.norm(fp)))
        Fs.close()
    per_fp = t_fp / len(sample)
    out["F"] = {"what": "whole F per simulation: EXTRACTDAT, Laplace solve, Dirichlet-to-Neumann map, SETVELF, "
                        "SOLVEI, INTERFACENEW, tail",
                "laplace_iterations_mean": round(float(lap_it.mean()), 2),
                "stokes_iterations_mean": round(float(st_it.mean()), 2),
                "batched_ms": round(t_fb * 1e3, 3), "batched_F_per_s": round(nsim / t_fb, 1),
                "plain_ms_per_F": round(per_fp * 1e3, 3), "plain_F_per_s": round(1.0 / per_fp, 1),
                "batched_vs_plain": round((nsim / t_fb) * per_fp, 2),
                "plain_vs_batched_rel_diff": worst_f}
    # one Heun step of every simulation: eager (host-controlled solves) and as a replayed
    # hipGraph (device-controlled solves, at most 8 Arnoldi steps each)
    gst = torch.cuda.Stream()
    Bg = spectral.Bodies(nd, stream=gst)
    SFg = field.SweepField(Bg, np.arange(nsim + 1), cap, e0, stream=gst)
    torch.cuda.synchronize()
    dt = 1e-3
    with torch.cuda.stream(gst):
        ye = rku_t.clone()
        yg = rku_t.clone()
        phi0, y = spectral.rk2_stage1(dt, ye, SFg.rhs(ye), stream=gst)  # warm-up
        spectral.rk2_stage2(dt, phi0, SFg.rhs(y), ye, stream=gst)
        SFg.rk2_graph(yg, dt)  # capture
        gst.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            phi0, y = spectral.rk2_stage1(dt, ye, SFg.rhs(ye), stream=gst)
            spectral.rk2_stage2(dt, phi0, SFg.rhs(y), ye, stream=gst)
        gst.synchronize()
        t_eager = (time.perf_counter() - t0) / reps
        redone = 0
        t0 = time.perf_counter()
        for _ in range(reps):
            redone += SFg.rk2_graph(yg, dt)
        gst.synchronize()
        t_graph = (time.perf_counter() - t0) / reps
        diff = float((yg - ye).norm() / ye.norm())
        # the sweep's run: 200 replayed Heun steps of all simulations at dt = 1e-3
        nrun, redone_run = 200, 0
        t0 = time.perf_counter()
        for _ in range(nrun):
            redone_run += SFg.rk2_graph(yg, dt)
        gst.synchronize()
        t_run = time.perf_counter() - t0
        run_finite = bool(torch.isfinite(yg).all())
        # the same run with classic RK4 (4 F per step): 200 replayed steps from the initial state
        yk = rku_t.clone()
        SFg.rk4_graph(yk, dt)  # capture
        gst.synchronize()
        redone_rk4 = 0
        t0 = time.perf_counter()
        for _ in range(nrun):
            redone_rk4 += SFg.rk4_graph(yk, dt)
        gst.synchronize()
        t_rk4 = time.perf_counter() - t0
        rk4_finite = bool(torch.isfinite(yk).all())
        d_rk4 = Bg.deformation(yk).cpu().numpy()
    torch.cuda.synchronize()
    out["rk2_step"] = {"what": "one Heun step (2 F) of all simulations", "dt": dt, "eager_ms": round(t_eager * 1e3, 3),
                       "graph_ms": round(t_graph * 1e3, 3), "graph_budget": 8, "graph_steps_redone": int(redone),
                       "steps_per_s_graph": round(1.0 / t_graph, 1),
                       "simulation_steps_per_s_graph": round(nsim / t_graph, 1),
                       "graph_vs_eager_state_rel_diff": diff,
                       "run": {"steps": nrun, "s": round(t_run, 3), "steps_per_s": round(nrun / t_run, 1),
                               "simulation_steps_per_s": round(nsim * nrun / t_run, 1),
                               "steps_redone": int(redone_run), "state_finite": run_finite}}
    out["rk4_run"] = {"what": "201 replayed classic RK4 steps (4 F each, eb_sweep_field_rk4_graph) of all "
                              "simulations at dt = 1e-3: the capture's step and the 200 timed",
                      "steps": nrun, "s": round(t_rk4, 3), "ms_per_step": round(t_rk4 * 1e3 / nrun, 3),
                      "simulation_steps_per_s": round(nsim * nrun / t_rk4, 1),
                      "F_per_s": round(4 * nsim * nrun / t_rk4, 1), "steps_redone": int(redone_rk4),
                      "state_finite": rk4_finite,
                      "deformation_after": {"min": round(float(d_rk4.min()), 6), "median": round(float(np.median(d_rk4)), 6),
                                            "max": round(float(d_rk4.max()), 6)}}
    SFg.close()
    # the plain path at its best: one simulation per field, its Heun step captured in a hipGraph
    # with device-controlled solves of the one system (eb_field_rk2_graph), same budget
    t_pg, t_pe, red_p, worst_p = 0.0, 0.0, 0, 0.0
    for s in sample:
        p0, p1 = off[s], off[s + 1]
        Fp = field.Field(spectral.Bodies(nd[s:s + 1], stream=gst), cap=float(cap[s]), e0=float(e0[s]), stream=gst)
        with torch.cuda.stream(gst):
            ye = rku_t[2 * p0:2 * p1].clone()
            yg = ye.clone()
            Fp.rk2_graph(yg, dt)  # eager F (plans), capture, first replay
            phi0, y = spectral.rk2_stage1(dt, ye, Fp.rhs(ye), stream=gst)
            spectral.rk2_stage2(dt, phi0, Fp.rhs(y), ye, stream=gst)
            gst.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                phi0, y = spectral.rk2_stage1(dt, ye, Fp.rhs(ye), stream=gst)
                spectral.rk2_stage2(dt, phi0, Fp.rhs(y), ye, stream=gst)
            gst.synchronize()
            t_pe += (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                red_p += Fp.rk2_graph(yg, dt)
            gst.synchronize()
            t_pg += (time.perf_counter() - t0) / reps
            worst_p = max(worst_p, float((yg - ye).norm() / ye.norm()))
        Fp.close()
    torch.cuda.synchronize()
    pg = len(sample) / t_pg  # simulation steps per second of the plain graph path
    out["rk2_step"]["plain"] = {
        "what": "one simulation per field (eb_field_rk2_graph): its Heun step in a hipGraph, device-controlled "
                "solves of the one system, budget 8", "sample": f"every 16th simulation ({len(sample)} fields)",
        "eager_ms_per_step": round(t_pe * 1e3 / len(sample), 3), "graph_ms_per_step": round(t_pg * 1e3 / len(sample), 3),
        "graph_steps_redone": int(red_p), "simulation_steps_per_s_graph": round(pg, 1),
        "graph_vs_eager_state_rel_diff": worst_p,
        "batched_graph_vs_plain_graph": round(out["rk2_step"]["simulation_steps_per_s_graph"] / pg, 2)}
    if with_cpu:
        import tempfile
        import oracle
        if oracle.ref_available():
            cwd = os.getcwd()
            os.chdir(tempfile.mkdtemp())
            sim = oracle.RefSim(oracle.LIBREF)
            t_ref, n_ref = 0.0, 0
            t_rf, n_rf = 0.0, 0
            try:
                for s in sample[:4]:
                    p0, p1 = off[s], off[s + 1]
                    sim.setup(nd[s:s + 1], cap=float(cap[s]), e0=float(e0[s]))
                    t0 = time.perf_counter()
                    sim.rhs(rku[2 * p0:2 * p1])
                    t_rf += time.perf_counter() - t0
                    n_rf += 1
                for s in sample[:8]:
                    p0, p1 = off[s], off[s + 1]
                    g = {k: geo[k][p0:p1].cpu().numpy() for k in ("x", "y", "dz", "dsda", "rkappa")}
                    sim.setup(nd[s:s + 1])
                    t0 = time.perf_counter()
                    oracle.ref_dgmres_stokes(nd[s:s + 1], g["x"] + 1j * g["y"], g["dz"], g["dsda"], g["rkappa"],
                                             geo["zk"][s:s + 1].cpu().numpy(), rhs[2 * p0:2 * p1].cpu().numpy())
                    t_ref += time.perf_counter() - t0
                    n_ref += 1
            finally:
                os.chdir(cwd)
            out["cpu_reference"] = {"solves_per_s": round(n_ref / t_ref, 2), "cores": 1, "kind": "reference",
                                    "sample": f"SOLVEI's solve of {n_ref} of the sampled simulations"}
            out["F"]["cpu_reference"] = {"F_per_s": round(n_rf / t_rf, 2), "cores": 1, "kind": "reference",
                                         "sample": f"the reference's F of {n_rf} of the sampled simulations"}
    gb.close()
    out["segmented_fmm"] = sweep_large(torch, dev)
    return out


def sweep_large(torch, dev):
    """Batched simulations too large for the O(N^2) sums: 16 simulations of eight 2048-point
    drops (16,384 points each, 262,144 in all), whole F of all of them in one call, with every
    layer potential on segmented FMM plans (one segment per simulation; eb_sweep_field_set_mode)
    against the same F with each simulation's O(N^2) sums."""
    from electricbubbles_amd import field, spectral
    from electricbubbles_amd.workloads import perturbed_circle
    rng = np.random.default_rng(20261204)
    nds, zs, sim_off = [], [], [0]
    for s in range(16):
        for b in range(8):
            c = 3.0 * (b % 4) + 3.0j * (b // 4)
            zs.append(perturbed_circle(2048, [(2, 0.05), (3, rng.uniform(-0.02, 0.02))], center=c)[0])
            nds.append(2048)
        sim_off.append(sim_off[-1] + 8)
    nd = np.array(nds, np.int32)
    rku = torch.tensor(np.concatenate(zs).view(np.float64).copy(), device=dev)
    out = {"workload": "16 simulations x 8 drops x 2048 points (16,384 per simulation), E0 = 1, CAP = 0.1"}
    res = {}
    for mode, name in ((1, "fmm"), (2, "direct")):
        SF = field.SweepField(spectral.Bodies(nd), np.array(sim_off), 0.1, 1.0)
        SF.set_mode(mode)
        SF.rhs(rku)  # plans, warm-up
        torch.cuda.synchronize()
        reps = 3 if mode == 1 else 1
        t0 = time.perf_counter()
        for _ in range(reps):
            f = SF.rhs(rku)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        lap_it, _, st_it, _ = SF.iterations()
        res[name] = f
        out[name] = {"ms_per_F_of_all": round(t * 1e3, 2), "F_per_s": round(16 / t, 1),
                     "laplace_iterations_mean": round(float(lap_it.mean()), 2),
                     "stokes_iterations_mean": round(float(st_it.mean()), 2)}
        SF.close()
    out["fmm_vs_direct"] = round(out["direct"]["ms_per_F_of_all"] / out["fmm"]["ms_per_F_of_all"], 2)
    out["rel_diff"] = float((res["fmm"] - res["direct"]).norm() / res["direct"].norm())
    return out


def single_drop_leg(torch, dev, with_cpu):
    """The single drop r = 1 + 0.05 cos 2 theta, N = 256, |E| = 1 (E0 = 1), no straining flow:
    100 classic RK4 steps at dt = 1e-3 as replayed hipGraphs (eb_field_rk4_graph, 4 F per step)
    with the deformation D = (L - B) / (L + B) on the device after every step (eb_deformation).
    The reference leg times the reference's own Heun step (it has no RK4) on one host core."""
    from electricbubbles_amd import field, spectral
    from electricbubbles_amd.workloads import perturbed_circle
    z, _, nd, _ = perturbed_circle(256, [(2, 0.05)])
    y0 = z.view(np.float64).copy()
    st = torch.cuda.Stream()
    B = spectral.Bodies(nd, stream=st)
    F = field.Field(B, cap=0.0, e0=1.0, stream=st)
    nstep, dt = 100, 1e-3
    with torch.cuda.stream(st):
        y = torch.tensor(y0, device=dev)
        F.rk4_graph(y.clone(), dt)  # plans, capture (on a copy)
        st.synchronize()
        y = torch.tensor(y0, device=dev)
        F.rk4_graph(y, dt)  # capture again for this buffer, first step
        D = [B.deformation(torch.tensor(y0, device=dev)), B.deformation(y)]
        st.synchronize()
        redone = 0
        t0 = time.perf_counter()
        for _ in range(nstep - 1):
            redone += F.rk4_graph(y, dt)
            D.append(B.deformation(y))
        st.synchronize()
        t_run = time.perf_counter() - t0
    Ds = torch.cat(D).cpu().numpy()
    out = {"workload": "one drop r = 1 + 0.05 cos 2 theta, N = 256, E0 = 1, CAP = 0, RK4 at dt = 1e-3",
           "steps": nstep, "ms_per_step": round(t_run * 1e3 / (nstep - 1), 3),
           "timing": f"{nstep - 1} replayed steps with D after each (the first step captures the graph)",
           "steps_redone": int(redone), "D_every_10_steps": [round(float(v), 12) for v in Ds[::10]],
           "D_final": float(Ds[-1])}
    F.close()
    if with_cpu:
        import tempfile
        import oracle
        if oracle.ref_available():
            cwd = os.getcwd()
            os.chdir(tempfile.mkdtemp())
            try:
                sim = oracle.RefSim(oracle.LIBREF)
                sim.setup(nd, cap=0.0, e0=1.0)
                t0 = time.perf_counter()
                sim.rk2(y0, dt, 10)
                t_ref = (time.perf_counter() - t0) / 10
            finally:
                os.chdir(cwd)
            out["cpu_reference"] = {"ms_per_heun_step": round(t_ref * 1e3, 2), "cores": 1, "kind": "reference",
                                    "sample": "10 steps of the reference's RK2 (2 F each) on this drop"}
    return out


def field_leg(torch, dev, with_cpu, only_drops=0):
    """F and the RK2 step (dt = 5e-4) of one simulation on the device, for the many-drop
    configuration (64 drops, about 28K points, E = (1.5, 0): E0 = 2.25, no background flow): every layer
    potential on the FMM (LEVMAX 7), the Laplace and Stokes solves in the device GMRES.  The
    reference's F cannot run this size as written (PVINTEV stops above 12,500 points,
    pv_integral.f:131-135), so the reference leg times its F on a 16-drop configuration of
    the same construction, next to the device F on that same configuration."""
    from electricbubbles_amd import field, spectral
    from electricbubbles_amd.workloads import many_drops_cfg
    out = {}
    for tag, ndrops, seed in (("many_drop_64", 64, 20260506), ("many_drop_16", 16, 20260516)):
        if only_drops and ndrops != only_drops:
            continue
        nd, rku = many_drops_cfg(ndrops, np.random.default_rng(seed))
        B = spectral.Bodies(nd)
        F = field.Field(B, cap=0.0, e0=2.25)
        y0 = torch.tensor(rku, device=dev)
        F.rhs(y0)  # warm-up: plans and trees
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            dy = F.rhs(y0)
        torch.cuda.synchronize()
        t_f = (time.perf_counter() - t0) / reps
        lap_it, _, st_it, _ = F.iterations()
        ybar = y0.clone()
        dt = 5e-4  # the many-drop configuration's step
        t0 = time.perf_counter()
        for _ in range(reps):
            phi0, y = spectral.rk2_stage1(dt, ybar, F.rhs(ybar))
            spectral.rk2_stage2(dt, phi0, F.rhs(y), ybar)
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / reps
        o = {"workload": f"{ndrops} drops, {int(nd.sum())} points", "ms_per_F": round(t_f * 1e3, 3),
             "laplace_iterations": lap_it, "stokes_iterations": st_it, "rk2_dt": dt,
             "ms_per_rk2_step": round(t_step * 1e3, 3), "steps_per_s": round(1.0 / t_step, 2),
             "iterations_after_steps": list(F.iterations()[::2])}
        if with_cpu and ndrops == 16:
            import tempfile
            import oracle
            if oracle.ref_available():
                cwd = os.getcwd()
                os.chdir(tempfile.mkdtemp())
                try:
                    sim = oracle.RefSim(oracle.LIBREF)
                    sim.setup(nd, cap=0.0, e0=2.25)
                    t0 = time.perf_counter()
                    ref = sim.rhs(rku)
                    t_ref = time.perf_counter() - t0
                finally:
                    os.chdir(cwd)
                o["cpu_reference"] = {"s_per_F": round(t_ref, 3), "cores": 1, "kind": "reference",
                                      "F_rel_diff": float(np.linalg.norm(dy.cpu().numpy() - ref) / np.linalg.norm(ref))}
        out[tag] = o
        F.close()
    return out


def adjoint_leg(torch, dev, with_cpu, steps=200):
    """Gradient of the summed deformation D = (L - B)/(L + B) of 16 drops after `steps` RK4
    steps at dt = 1e-3 with respect to |E| (E0 = |E|^2 = 1), Ca (the imposed strain's CAP =
    0.1) and the 64 real Fourier modes of each initial shape (1024 modes: cos and sin of
    m = 1..32, amplitudes U[0, 0.02], mode m entering with weight 1/m^2); drops of radius 1 and
    N = 512 on a 4 x 4 grid with spacing 4 (8192 points).  With equal weights the shapes carry
    curvatures of order 100, where the reference's own F gives |dz/dt| of 200 to 2600 and the
    explicit step at dt = 1e-3 diverges at once; with weight 1/m the high-frequency content
    grows from the first steps and the run diverges by step 40; with 1/m^2 it runs the 200
    steps with |dz/dt| near 1.1 and flat GMRES counts.  GMRES
    tolerance 1e-12 in both directions.  Forward alone, then the adjoint (forward with
    checkpoints every 20 steps, reverse sweep with recomputation), and a spot check of the
    gradient along a random direction of all 1026 parameters against Richardson-extrapolated
    central differences of the same forward over the first 20 steps."""
    from electricbubbles_amd import adjoint, field, spectral
    P = adjoint.Problem(16, 512, tol=1e-12)
    centres, modes = adjoint.grid_problem(4, 512, 4.0, 64, 0.02, seed=5, device=dev)
    E, Ca, dt, decay = 1.0, 0.1, 1e-3, 2.0
    # warm-up (FFT plans, workspace)
    adjoint.forward_loss(P, centres, modes, E, Ca, 1, dt, decay=decay)
    torch.cuda.synchronize()
    P.sys.iters = {k: [] for k in P.sys.iters}
    t0 = time.perf_counter()
    lf = adjoint.forward_loss(P, centres, modes, E, Ca, steps, dt, decay=decay)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    it_fwd = {k: float(np.mean([a for a, _ in v])) for k, v in P.sys.iters.items() if v}
    P.sys.iters = {k: [] for k in P.sys.iters}
    log(f"adjoint: forward {steps} steps in {t_fwd:.1f} s")
    loss, gE, gCa, gm, tim = adjoint.gradient(P, centres, modes, E, Ca, steps, dt, checkpoint=20, decay=decay)
    log(f"adjoint: gradient in {tim['total_s']:.1f} s")
    it_grad = {k: float(np.mean([a for a, _ in v])) for k, v in P.sys.iters.items() if v}
    # spot check over 20 steps along a random unit direction (dE, dCa, dmodes)
    ns = min(20, steps)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(2 + modes.numel())
    v /= np.linalg.norm(v)
    vm = torch.tensor(v[2:].reshape(modes.shape), device=dev)
    _, gE10, gCa10, gm10, _ = adjoint.gradient(P, centres, modes, E, Ca, ns, dt, checkpoint=5, decay=decay)
    an = gE10 * v[0] + gCa10 * v[1] + float((gm10 * vm).sum())
    h = 1e-5

    def L(e):
        return adjoint.forward_loss(P, centres, modes + e * vm, E + e * v[0], Ca + e * v[1], ns, dt, decay=decay)
    fd = (4 * (L(h / 2) - L(-h / 2)) / h - (L(h) - L(-h)) / (2 * h)) / 3
    # the library's own F (FMM path, reference tolerances) at this configuration, for scale
    z0 = adjoint.initial_shapes(centres, modes, 512, decay).reshape(-1)
    rku = torch.stack([z0.real, z0.imag], 1).reshape(-1).contiguous()
    B = spectral.Bodies([512] * 16)
    Fl = field.Field(B, cap=Ca, e0=E * E)
    Fl.rhs(rku)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        Fl.rhs(rku)
    torch.cuda.synchronize()
    t_lib = (time.perf_counter() - t0) / 3
    Fl.close()
    out = {"workload": "16 drops x 512 points (4 x 4 grid, spacing 4), 1024 shape modes (weight 1/m^2) "
                       f"+ |E| + Ca, {steps} RK4 steps at dt = 1e-3, checkpoints every 20",
           "s_per_grad": round(tim["total_s"], 3), "forward_s": round(t_fwd, 3),
           "grad_over_forward": round(tim["total_s"] / t_fwd, 3),
           "grad_phases_s": {k: round(val, 3) for k, val in tim.items()},
           "loss": loss, "loss_forward_rel_diff": abs(loss - lf) / abs(lf),
           "dL_dE": gE, "dL_dCa": gCa, "dL_dmodes_norm": float(gm.norm()),
           "gmres_iterations_mean": {"forward": it_fwd, "gradient": it_grad},
           "fd_check": {"steps": ns, "direction": "random unit vector in all 1026 parameters", "adjoint": an,
                        "central_difference": fd, "rel_diff": abs(an - fd) / abs(fd)},
           "library_F_ms": round(t_lib * 1e3, 3),
           "differentiable_F_ms": round(t_fwd / (4 * steps) * 1e3, 3)}
    if with_cpu:
        import tempfile
        import oracle
        if oracle.ref_available():
            cwd = os.getcwd()
            os.chdir(tempfile.mkdtemp())
            try:
                sim = oracle.RefSim(oracle.LIBREF)
                sim.setup(np.full(16, 512, np.int32), cap=Ca, e0=E * E)
                t0 = time.perf_counter()
                sim.rhs(rku.cpu().numpy())
                t_ref = time.perf_counter() - t0
            finally:
                os.chdir(cwd)
            nparam = 2 + modes.numel()
            out["cpu_reference"] = {"s_per_F": round(t_ref, 3), "cores": 1, "kind": "reference",
                                    "sample": "one F of the configuration",
                                    "fd_gradient_s_estimate": round(2 * nparam * 4 * steps * t_ref, 1),
                                    "note": "the reference has no gradient; central differences need 2 runs of "
                                            f"{4 * steps} F per parameter ({nparam} parameters)"}
    P.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["curves", "uniform"], default="curves")
    ap.add_argument("--levmax", type=int, default=0, help="0: the autotuner's choice")
    ap.add_argument("--nterms", type=int, default=0, help="0: the autotuner's choice (1e-10 relative)")
    ap.add_argument("--npts", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the fixed LEVMAX 8 / p = 32 comparison point")
    ap.add_argument("--no-second", action="store_true", help="skip the second FMM workload (uniform points)")
    ap.add_argument("--no-solve", action="store_true", help="skip the Stokes GMRES solve leg")
    ap.add_argument("--no-spectral", action="store_true", help="skip the spectral geometry leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-versus-plain sweep leg")
    ap.add_argument("--no-field", action="store_true", help="skip the F and RK2 step leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient leg")
    ap.add_argument("--no-single", action="store_true", help="skip the single-drop RK4 leg")
    ap.add_argument("--no-autotune", action="store_true", help="skip the FMM autotuner")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="RK4 steps of the adjoint leg")
    ap.add_argument("--only", choices=["solve", "spectral", "sweep", "field", "adjoint", "single"],
                    help="run one secondary leg alone and print its JSON object (for profiling runs)")
    ap.add_argument("--field-drops", type=int, default=0, choices=[0, 16, 64],
                    help="with --only field: run only the 64- or the 16-drop configuration")
    ap.add_argument("--p2p-tile", type=int, default=0,
                    help="near-field tile of the timed plans (0: the autotuner's, else 42)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (the known-size kernel that scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the FMM is measured on one MI355X); "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    # The one JSON line goes to the original stdout; everything else written to fd 1 (the
    # reference Fortran's unit-6 printing in the CPU legs) is discarded.
    json_fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    os.close(devnull)

    import torch
    import electricbubbles_amd as eb

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if args.only:
        legs = {"solve": stokes_solve_leg, "spectral": spectral_leg, "sweep": sweep_leg, "field": lambda t, d, c: field_leg(t, d, c, args.field_drops),
                "single": single_drop_leg,
                "adjoint": lambda t, d, c: adjoint_leg(t, d, c, args.adjoint_steps)}
        res = {"leg": args.only, args.only: legs[args.only](torch, dev, not args.no_cpu)}
        sys.stdout.flush()
        with os.fdopen(json_fd, "w") as jf:
            jf.write(json.dumps(res) + "\n")
        return
    n = args.npts
    stream = torch.cuda.current_stream()
    if args.calibrate_copy:
        src = torch.empty(32 << 20, dtype=torch.float64, device=dev).uniform_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    head, head_host = leg(f"FMM {args.workload}", fmm_workload, torch, eb, dev, args.workload, n, args, stream)
    other = None
    if not args.no_second:
        kind2 = "uniform" if args.workload == "curves" else "curves"
        other, other_host = leg(f"FMM {kind2}", fmm_workload, torch, eb, dev, kind2, n, args, stream)

    solve = None if args.no_solve else leg("stokes solve", stokes_solve_leg, torch, dev, not args.no_cpu)
    spec = None if args.no_spectral else leg("spectral", spectral_leg, torch, dev, not args.no_cpu)
    sweep = None if args.no_sweep else leg("sweep", sweep_leg, torch, dev, not args.no_cpu)
    fld = None if args.no_field else leg("field", field_leg, torch, dev, not args.no_cpu)
    single = None if args.no_single else leg("single drop", single_drop_leg, torch, dev, not args.no_cpu)
    adj = None if args.no_adjoint else leg("adjoint", adjoint_leg, torch, dev, not args.no_cpu, args.adjoint_steps)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: start")
        z, dz, nd, h, dens, _, _ = head_host
        cpu = cpu_baseline(z, dz, nd, h, dens, head["levmax"], head["nterms"],
                           nd.size if args.workload == "curves" else 4096)
        # the GPU results against the long-double O(N^2) sum of the C restatement (the checker)
        cpu["accuracy_check"] = {args.workload: direct_check(*head_host[:6])}
        if other is not None:
            cpu["accuracy_check"][other["kind"]] = direct_check(*other_host[:6])
        log("cpu baseline: done")

    L, p = head["levmax"], head["nterms"]
    fl, stage = head["_flops"], head["stage_ms"]
    dom = max(("p2p", "m2l"), key=lambda k: stage[k])
    achieved = fl[dom] / (stage[dom] * 1e-3) / 1e12
    traffic = PROFILED_TRAFFIC_BYTES.get((args.workload, n, L, p), {}).get(dom)
    line = {
        "metric": "FMM Mpts/s @N=2^20",
        "value": head["value"],
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": head["config"],
        "apply_only": head["apply_only"],
        "stage_ms": {k: round(v, 4) for k, v in stage.items()},
        "stages": head["stages"],
        "fp64_fraction_total": head["fp64_fraction_total"],
        "accuracy": head["accuracy"],
        "adaptive_tree": head["adaptive_tree"],
        "adaptive_fmm": head["adaptive_fmm"],
        "autotune": head["autotune"],
        "fixed_levmax8_p32": head.get("fixed_levmax8_p32"),
        "second_workload": {k: v for k, v in other.items() if not k.startswith("_")} if other else None,
        "stokes_solve": solve,
        "spectral": spec,
        "sweep": sweep,
        "field": fld,
        "single_drop": single,
        "adjoint": adj,
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                     "flops_per_launch": fl[dom], "pairs": head["_pairs"] if dom == "m2l" else None,
                     "timing": "HIP events on the plan's stream, mean of 5 applies after the timed loops"},
        "cpu_baseline": cpu,
    }
    sys.stdout.flush()
    with os.fdopen(json_fd, "w") as jf:
        jf.write(json.dumps(line) + "\n")


def adaptive_fmm(torch, eb, x, y, w, d, p, tgt, ref, args, stream):
    """The same sums on DAPIF2's adaptive tree (AFMMPlan: at most 40 points per childless box,
    the same p): plan creation (tree build, lists on the host, uploads) once, K applies timed,
    per-stage HIP events, accuracy at the same 1024 targets against the same O(N^2) sum."""
    n = int(x.numel())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P = eb.AFMMPlan(x, y, w, p, napb=40, stream=stream)
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    out = torch.empty(n, dtype=torch.complex128, device=x.device)
    for _ in range(args.warmup):
        P.apply(d, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.apply(d, out)
    torch.cuda.synchronize()
    t_apply = (time.perf_counter() - t0) / args.steps
    P.enable_timing(True)
    acc = None
    for _ in range(5):
        P.apply(d, out)
        st = P.stage_ms()
        acc = st if acc is None else {k: acc[k] + st[k] for k in st}
    P.enable_timing(False)
    stage = {k: v / 5 for k, v in acc.items()}
    P.apply(d, out)
    err = float((out[tgt.long()] - ref).norm() / ref.norm())
    info = P.info()
    info.update(P.m2l_split())
    P.close()
    # algorithmic flops: P2L as P2M per (point, box) term, M2P as 36 p per (point, box)
    fl = {"charges": 0.0, "p2p": 50.0 * info["near_pairs"], "p2m": 26.0 * p * n,
          "m2m": (4.0 * p * (p + 1) + 30.0 * p) * info["shifted_boxes"],
          "m2l": m2l_flops(info, p), "p2l": 26.0 * p * info["p2l_terms"],
          "l2l": (4.0 * p * (p + 1) + 30.0 * p) * info["shifted_boxes"],
          "far": 26.0 * p * n + 36.0 * p * info["m2p_evals"]}
    return {"what": "the same sums on DAPIF2's adaptive tree (eb_afmm: NAPB 40, lists U/V/W/X), apply only",
            "value": round(n / t_apply / 1e6, 3), "unit": "Mpts/s", "ms": round(t_apply * 1e3, 4),
            "create_ms": round(t_create * 1e3, 1), "nterms": p, "napb": 40,
            "stage_ms": {k: round(v, 4) for k, v in stage.items()},
            "fp64_fraction_by_stage": {k: round(fl[k] / (stage[k] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)
                                       for k in fl if fl[k] > 0 and stage[k] > 0},
            "fp64_fraction_total": round(sum(fl.values()) / t_apply / 1e12 / PEAK_FP64_TFLOPS, 4),
            "counts": {k: info[k] for k in ("boxes", "leaves", "levels", "near_pairs", "m2l_pairs", "m2l_near", "m2l_far", "m2l_far_terms", "m2p_evals",
                                            "p2l_terms", "shifted_boxes")},
            "accuracy": {"rel_err2_vs_direct": err, "targets": 1024}}


def fmm_workload(torch, eb, dev, kind, n, args, stream):
    """One FMM-alone workload (2^20 points by default): the circles (4096 x 256 points, radii
    U[0.01, 0.05], centres U[0,1]^2) or uniform points in the unit square, CN(0,1) densities,
    CADAPDR's kernel.  The autotuner's depth and length for 1e-10 unless given; K complete
    evaluations (tree build + apply) timed, then K applies alone, then per-stage HIP events;
    accuracy against the tree-free O(N^2) sum at 1024 seeded targets (outside the timed loops).
    Returns the result object and the host copies the CPU leg needs."""
    from electricbubbles_amd.workloads import complex_normal, curves_cfg, uniform_cfg, weights
    rng = np.random.default_rng(20260303)
    if kind == "curves":
        ncurves = n // 256
        z, dz, nd, h = curves_cfg(ncurves, 256, rng)
    else:
        ncurves = 0
        z, dz, nd, h = uniform_cfg(n, rng)
    dens = complex_normal(n, rng)
    x = torch.tensor(z.real.copy(), device=dev)
    y = torch.tensor(z.imag.copy(), device=dev)
    w = torch.tensor(weights(dz, nd, h), device=dev)
    d = torch.tensor(dens, device=dev)
    out = torch.empty(n, dtype=torch.complex128, device=dev)
    p, L = args.nterms, args.levmax
    tuned = None
    if L <= 0 or p <= 0 or not args.no_autotune:
        from electricbubbles_amd import autotune
        t0 = time.perf_counter()
        tr = autotune.tune_fmm(x, y, w, d, target=1e-10)
        log(f"autotune {kind}: LEVMAX {tr['levmax']}, p {tr['nterms']}")
        tuned = {"target_rel": 1e-10, "check": tr["check"], "levmax": tr["levmax"], "nterms": tr["nterms"],
                 "p2p_tile": tr["p2p_tile"], "apply_ms": tr["ms"], "err_vs_direct": tr["err"],
                 "tuning_s": round(time.perf_counter() - t0, 3), "p_scan": tr["p_scan"],
                 "p_scan_levmax": tr["p_scan_levmax"], "tile_scan": tr["tile_scan"],
                 "candidates": [{k: c[k] for k in ("levmax", "points_per_leaf", "ms", "err")} for c in tr["candidates"]]}
        L = tr["levmax"] if L <= 0 else L
        p = tr["nterms"] if p <= 0 else p
    tile = tuned["p2p_tile"] if tuned and (L, p) == (tuned["levmax"], tuned["nterms"]) else 42
    if args.p2p_tile > 0:
        tile = args.p2p_tile
    log(f"FMM {kind}: {n} points, LEVMAX {L}, p {p}, near-field tile {tile}")
    plan = eb.FMMPlan(x, y, w, L, p, eb.VARIANT_DR, eb.MODE_FMM, stream=stream)
    plan.set_p2p_tile(tile)

    def timed(full, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            if full:
                plan.rebuild(x, y, w)
            plan.apply(d, out)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    timed(True, args.warmup)
    t_full = timed(True, args.steps)  # K complete evaluations (tree build + far + near)
    t_apply = timed(False, args.steps)  # apply only (tree kept: the GMRES inner loop)
    # per-stage times with HIP events on the plan's stream (stages run one after another here)
    plan.enable_timing(True)
    acc = None
    for _ in range(5):
        plan.apply(d, out)
        st = plan.stage_ms()
        acc = st if acc is None else {k: acc[k] + st[k] for k in st}
    stage = {k: v / 5 for k, v in acc.items()}
    plan.enable_timing(False)
    plan.apply(d, out)
    info = plan.info()
    info.update(plan.m2l_split())
    fl, by = flops_model(info, n, p), bytes_model(info, n, p)
    # accuracy: the O(N^2) sum over all sources at 1024 seeded targets (tree-free, on the GPU)
    tgt = torch.tensor(np.sort(np.random.default_rng(7).choice(n, 1024, replace=False)).astype(np.int32), device=dev)
    ref = eb.direct_subset(x, y, w, d, tgt)
    err = float((out[tgt.long()] - ref).norm() / ref.norm())
    gw = out.cpu().numpy()
    # the adaptive tree of the same points (DAPIF2's rules, at most 40 points per childless box)
    from electricbubbles_amd.atree import AdaptiveTree
    at = AdaptiveTree(x, y, 40)
    at.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        at = AdaptiveTree(x, y, 40)
        at_info = {"napb": 40, "boxes": at.nbox, "levels": at.nlev, "max_leaf": at.maxp}
        at.close()
    torch.cuda.synchronize()
    at_info["build_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 3)
    adaptive = adaptive_fmm(torch, eb, x, y, w, d, p, tgt, ref, args, stream)
    res = {
        "kind": kind,
        "value": round(n / t_full / 1e6, 3), "ms_per_step": round(t_full * 1e3, 4),
        "apply_only": {"value": round(n / t_apply / 1e6, 3), "unit": "Mpts/s", "ms": round(t_apply * 1e3, 4)},
        "levmax": L, "nterms": p,
        "config": {"workload": f"FMM alone, {n} points on {ncurves} circles x 256 (r U[0.01,0.05], "
                               f"centres U[0,1]^2)" if ncurves else f"FMM alone, {n} uniform points in the unit square",
                   "kernel": "CADAPDR GW (DR signs), sources = targets",
                   "levmax": L, "nterms": p, "p2p_tile": tile, "leaves_nonempty": info["leaves"],
                   "max_leaf": info["max_leaf"]},
        "stage_ms": stage,
        "stages": stage_table(stage, fl, by),
        "fp64_fraction_total": round(sum(fl.values()) / t_apply / 1e12 / PEAK_FP64_TFLOPS, 4),
        "accuracy": {"rel_err2_vs_direct": err, "targets": 1024, "sources": n,
                     "checker": "O(N^2) sum over all sources on the GPU (eb_direct_subset: no tree, lists or expansions)"},
        "adaptive_tree": at_info,
        "adaptive_fmm": adaptive,
        "autotune": tuned,
        "_flops": fl,
        "_pairs": {k: info[k] for k in ("m2l_pairs", "m2l_near", "m2l_far", "m2l_far_terms")},
    }
    plan.close()
    if not args.no_check and kind == "curves" and (L, p) != (8, 32):
        p32 = eb.FMMPlan(x, y, w, 8, 32, eb.VARIANT_DR, eb.MODE_FMM, stream=stream)
        o32 = p32.apply(d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p32.rebuild(x, y, w)
            p32.apply(d, o32)
        torch.cuda.synchronize()
        t32f = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p32.apply(d, o32)
        torch.cuda.synchronize()
        t32a = (time.perf_counter() - t0) / args.steps
        res["fixed_levmax8_p32"] = {"levmax": 8, "nterms": 32, "value": round(n / t32f / 1e6, 3),
                                    "apply_only_value": round(n / t32a / 1e6, 3),
                                    "rel_err2_vs_direct": float((o32[tgt.long()] - ref).norm() / ref.norm())}
        p32.close()
    return res, (z, dz, nd, h, dens, gw, tgt.cpu().numpy())


if __name__ == "__main__":
    main()
