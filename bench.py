"""Benchmark: particle-updates/s of the MI355X PM KDK step.

Workload (default): 512^3 particles from synthetic 1LPT initial conditions
(BBKS spectrum, sigma_8 = 0.8, z = 49), 1024^3 CIC mesh, box 512 Mpc/h,
log-a KDK schedule z = 49 -> 0 in 64 steps.  One "step" is one force refresh
fused with the K2 / K1 kicks and the drift (hy_pm_kdk), i.e. one KDK step of
the reference's step loop.  Binary32 mesh/FFT/forces/momenta, binary64
positions and time integrals.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s

STAGES = ["bucket", "deposit", "fft_z_r2c", "fft_y_fwd", "fft_x_green", "fft_y_bwd",
          "fft_z_c2r", "gather_kdk"]


def stage_bytes(stage: str, n_p: int, N: int, real: int) -> float:
    """Algorithmic HBM bytes per launch of each stage of a timed step (real =
    bytes per real).  Timed steps are asynchronous with reused keys: binary32
    carries 16-B deposit records from the gather to the next deposit, binary64
    carries bucket keys and ranks."""
    M = N ** 3
    NH = N // 2 + 1
    half = N * N * NH * 2 * real  # one half-spectrum array
    pos = 24.0 * n_p
    mom = 3.0 * real * n_p
    fp32 = real == 4
    nb = (N // 16) ** 3 * 8  # deposit buckets: 16^3 tiles x 8 boundary classes
    return {
        # FP32: capacity scan (counts, overflow counts read; offsets, overflow
        # prefix, next capacities written); FP64: keys + ranks and positions
        # read, 32-B deposit records (cell, index, three fractions) written
        "bucket": 20.0 * nb if fp32 else 8.0 * n_p + pos + 32.0 * n_p,
        # records read (FP32 16 B, FP64 32 B); mesh written
        "deposit": (16.0 if fp32 else 32.0) * n_p + real * M,
        "fft_z_r2c": real * M + half,               # real mesh in, half spectrum out
        "fft_y_fwd": 2.0 * half,
        "fft_x_green": 2.0 * half,
        "fft_y_bwd": 2.0 * half,
        "fft_z_c2r": half + real * M,
        # state in/out + one potential read; FP32: deposit records written,
        # FP64: keys + ranks written
        "gather_kdk": 2.0 * (pos + mom) + real * M + (16.0 if fp32 else 8.0) * n_p,
    }[stage]


def step_alg_bytes(n_p: int, N: int, real: int) -> float:
    """Compulsory bytes of one KDK step: the particle state read and written
    once (FP32-mixed 72 B, FP64 96 B per particle), the positions read once
    more by the deposit (24 B), and the mesh passes (FP32-mixed 24 B, FP64
    48 B per cell): FP32-mixed 96*N_p + 24*M, FP64 120*N_p + 48*M."""
    M = N ** 3
    return (96.0 * n_p + 24.0 * M) if real == 4 else (120.0 * n_p + 48.0 * M)


STAGE_KERNEL = {"bucket": "hy::k_cap_apply", "deposit": "void hy::k_deposit_fixed<16>",
                "gather_kdk": "void hy::k_gather_kdk<float, float, true, true, false>"}
FFT_KERNEL = {"fft_z_r2c": "k_fft_z_r2c", "fft_y_fwd": "k_fft_y_fwd", "fft_x_green": "k_fft_x_green",
              "fft_y_bwd": "k_fft_y_bwd", "fft_z_c2r": "k_fft_z_c2r"}


def measured_traffic(stage: str, n_side: int, N: int, prec: str, plan: dict):
    """HBM bytes per launch of the stage's kernel from the committed calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, if present."""
    path = os.path.join(ROOT, "profiles", f"pm_kdk_step_{n_side}p_{N}m_{prec}_traffic.json")
    if not os.path.exists(path) or prec != "fp32":
        return None, None
    with open(path) as f:
        prof = json.load(f)
    if stage in FFT_KERNEL:
        v = plan["variant"][stage[4:]]
        sg = ", 0" if stage in ("fft_x_green", "fft_y_bwd", "fft_z_c2r") else ""  # spectral-gradient flag
        name = f"void hy::{FFT_KERNEL[stage].replace('k_fft_', 'k_tl_')}<float, {N}, {v}{sg}>"
    else:
        name = STAGE_KERNEL.get(stage)
    k = prof.get("kernels", {}).get(name)
    return (k["traffic_bytes"] if k else None), os.path.relpath(path, ROOT)


def run_gpu(args):
    import numpy as np
    import torch

    import hyowon_amd as H
    from hyowon_amd import ic
    from hyowon_amd._lib import lib

    n_side, N, L = args.n_side, args.mesh, args.box
    n = n_side ** 3
    a0, a1, nsteps = 1.0 / (1.0 + args.z_start), 1.0, args.schedule_steps
    stepper = H.TimeStepper(a0, a1, math.log(a1 / a0) / nsteps, [a1])
    b = stepper.boundaries
    dk = H.DriftKickIntegrator(0.3, 0.7)
    K1, K2, D = [], [], []
    for i in range(len(b) - 1):
        ah = H.geometric_mean(b[i], b[i + 1])
        K1.append(dk.kick_factor(b[i], ah))
        K2.append(dk.kick_factor(ah, b[i + 1]))
        D.append(dk.drift_factor(b[i], b[i + 1]))
    m = ic.particle_mass(n_side, L)

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE under
        # rocprofv3 --pmc calibrate the counters against algorithmic bytes
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)  # elementwise kernel: reads and writes calibrate_copy bytes
        torch.cuda.synchronize()
        del src, dst
    pos, mom = ic.lpt_ics(n_side, L, a0, seed=1, device=dev, rng="philox", max_mode=n_side // 2 - 1)
    results = {}
    for prec in args.precisions:
        mdt = torch.float32 if prec == "fp32" else torch.float64
        st = H.ParticleState(pos[0].clone(), pos[1].clone(), pos[2].clone(),
                             mom[0].to(mdt).contiguous(), mom[1].to(mdt).contiguous(),
                             mom[2].to(mdt).contiguous())
        sp = st.empty_like()
        solver = H.PMSolver(L, N, precision=prec, max_particles=n)
        if args.fft_plan:
            var, _, grp = args.fft_plan.partition(":")
            solver.set_tuning([int(v) for v in var.split(",")], plane_group=int(grp or 0))
            plan = solver.tuning()
        else:
            plan = solver.autotune(reps=3) if args.autotune else solver.tuning()
        # gather schedule (lattice rows per column tile) for these particles
        plan["gather"] = solver.autotune_gather(st.x, st.y, st.z, m, reps=2) if args.autotune else None
        real = 4 if prec == "fp32" else 8
        steps = args.steps if prec == args.precisions[0] else max(2, args.fp64_steps)
        warm = args.warmup if prec == args.precisions[0] else 1

        def call(i, asynchronous):
            nonlocal st, sp
            j = i % nsteps
            kp = K2[(i - 1) % nsteps] if i > 0 else None
            solver.kdk(st, sp, m, kp, K1[j], D[j], asynchronous=asynchronous, reuse_keys=True)
            st, sp = sp, st

        # warmup: initial force + kick/drift, then steady steps (checked synchronously)
        for i in range(warm):
            call(i, asynchronous=False)
        torch.cuda.synchronize()
        lib().hy_pm_profile(solver.handle, 1, steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(warm, warm + steps):
            call(i, asynchronous=True)
        solver.check()  # synchronises the context stream, raises on any kernel error
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (C.c_double * 8)()
        cnt = (C.c_uint64 * 8)()
        calls = C.c_uint64()
        lib().hy_pm_profile_read(solver.handle, ms, cnt, C.byref(calls))
        lib().hy_pm_profile(solver.handle, 0, 0)
        stage_ms = {STAGES[k]: ms[k] / max(cnt[k], 1) for k in range(8)}
        step_ms = (t1 - t0) * 1e3 / steps
        results[prec] = dict(step_ms=step_ms, stage_ms=stage_ms, real=real, steps=steps, plan=plan)
        info = solver.info()
        results[prec]["device_bytes"] = int(info.device_bytes)
        solver.close()
        del st, sp
        torch.cuda.empty_cache()
    # comparison and accuracy runs come after the timed loops: their large
    # transient allocations must not shape the timed contexts' memory
    if args.plain_fp64 and "fp32" in args.precisions:
        results["plain_fp64"] = plain_fp64(args, pos, mom, m, K1, K2, D, nsteps)
    if args.rocfft and "fp32" in args.precisions:
        results["rocfft"] = rocfft_compare(N, dev)
    if args.accuracy and "fp32" in args.precisions:
        results["accuracy"] = accuracy(args, pos, mom, m, K1, K2, D, nsteps)
    return n, N, results


def plain_fp64(args, pos, mom, m, K1, K2, D, nsteps, steps=2):
    """The unfused binary64 GPU step that the >=5x target is defined against:
    the reference's stage order (src/time/leapfrog_integrator.cpp:143-279 over
    src/gravity/pm_solver.cpp:649-843) as separate passes -- K1 kick, drift,
    zeroed acceleration, deposit + density source (the library's binary64
    deposit), rocFFT r2c (torch.fft, binary64), Green multiply, rocFFT c2r
    (1/N^3 inside), centered-gradient gather, K2 kick.  Comparison only: not
    a product path.  Per-stage times with torch events; the force of its first
    refresh is compared with the fused binary64 step's."""
    import torch
    import hyowon_amd as H
    n_side, N, L = args.n_side, args.mesh, args.box
    n = n_side ** 3
    dev = pos[0].device
    x, y, z = (p.clone() for p in pos)
    px, py, pz = (q.to(torch.float64).clone() for q in mom)
    g = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    s = H.PMSolver(L, N, precision="fp64", max_particles=n)
    # Green factor 4 pi G / (lam_x + lam_y + lam_z), lam = (-4 (1/dx)^2 s) s (src/mesh/green_function.cpp)
    dx = L / N
    i = torch.arange(N, device=dev, dtype=torch.float64)
    mode = torch.where(i <= N // 2, i, i - N)
    sn = torch.sin(math.pi * mode / N)
    lam = (-4.0 * (1.0 / dx) ** 2 * sn) * sn
    nh = N // 2 + 1
    green = lam.view(N, 1, 1) + lam.view(1, N, 1) + lam[:nh].view(1, 1, nh)
    green = (4.0 * math.pi * H.gravitational_constant()) / green
    green[0, 0, 0] = 0.0
    names = ["kick_K1", "drift", "zero_acc", "deposit_source", "rocfft_r2c", "green", "rocfft_c2r", "gather",
             "kick_K2"]
    tot = {k: 0.0 for k in names}

    def force():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        ev[0].record()
        for a in g:
            a.zero_()
        ev[1].record()
        src = s.density(x, y, z, None, m)
        ev[2].record()
        spec = torch.fft.rfftn(src)
        del src
        ev[3].record()
        spec.mul_(green)
        ev[4].record()
        phi = torch.fft.irfftn(spec, s=(N, N, N))
        del spec
        ev[5].record()
        return ev, phi

    def step(j, timed):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        e[0].record()
        for q, a in zip((px, py, pz), g):
            q.add_(a, alpha=K1[j])
        e[1].record()
        for c, q in zip((x, y, z), (px, py, pz)):
            c.add_(q, alpha=D[j])
            torch.remainder(c, L, out=c)
        e[2].record()
        ev, phi = force()
        f0 = torch.cuda.Event(enable_timing=True)
        f0.record()
        s.gather(phi, x, y, z, *g)
        del phi
        f1 = torch.cuda.Event(enable_timing=True)
        f1.record()
        for q, a in zip((px, py, pz), g):
            q.add_(a, alpha=K2[j])
        e[3].record()
        torch.cuda.synchronize()
        if timed:
            tot["kick_K1"] += e[0].elapsed_time(e[1])
            tot["drift"] += e[1].elapsed_time(e[2])
            tot["zero_acc"] += ev[0].elapsed_time(ev[1])
            tot["deposit_source"] += ev[1].elapsed_time(ev[2])
            tot["rocfft_r2c"] += ev[2].elapsed_time(ev[3])
            tot["green"] += ev[3].elapsed_time(ev[4])
            tot["rocfft_c2r"] += ev[4].elapsed_time(ev[5])
            tot["gather"] += f0.elapsed_time(f1)
            tot["kick_K2"] += f1.elapsed_time(e[3])

    # initial force at the ICs, compared with the fused binary64 force
    ev, phi = force()
    s.gather(phi, x, y, z, *g)
    del phi
    ref = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    s.compute_forces(pos[0], pos[1], pos[2], None, m, *ref)
    d = sum(float(((a - b) ** 2).sum()) for a, b in zip(g, ref))
    gg = sum(float((b ** 2).sum()) for b in ref)
    del ref
    step(0, False)  # warm-up (plans, allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        step(k % nsteps, True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    s.close()
    del x, y, z, px, py, pz, g, green
    torch.cuda.empty_cache()
    return {"ms_per_step": ms, "steps": steps, "stage_ms": {k: round(v / steps, 4) for k, v in tot.items()},
            "force_rms_rel_vs_fused_fp64": math.sqrt(d / gg),
            "definition": "reference stage order as separate passes: K1 kick, drift, zeroed acceleration, "
                          "binary64 deposit + density source (library), rocFFT r2c (torch.fft, binary64), "
                          "Green multiply, rocFFT c2r, centered-gradient gather (library), K2 kick"}


def _time_cuda(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def rocfft_compare(N, dev, reps=5):
    """rocFFT (through torch.fft / hipFFT) r2c + c2r of the same binary32 N^3
    mesh, timed beside the hand-written pencil passes (comparison only)."""
    import torch
    x = torch.randn(N, N, N, dtype=torch.float32, device=dev)
    spec = torch.fft.rfftn(x)
    r2c = _time_cuda(lambda: torch.fft.rfftn(x), reps)
    c2r = _time_cuda(lambda: torch.fft.irfftn(spec, s=(N, N, N)), reps)
    del x, spec
    torch.cuda.empty_cache()
    return {"r2c_ms": r2c, "c2r_ms": c2r, "via": f"torch {torch.__version__} torch.fft (hipFFT/rocFFT)"}


def accuracy(args, pos, mom, m, K1, K2, D, nsteps):
    """FP32-mixed against binary64 on the same ICs: RMS relative force error
    sqrt(sum|g32-g64|^2 / sum|g64|^2), median and p99 of |dg|/|g| per particle,
    then the full log-a run in both precisions (final positions as periodic
    minimum-image distances in cells, and the matter power spectrum below half
    the Nyquist wavenumber from hy_pm_power_spectrum).  The binary64 GPU step
    is itself pinned to the oracle by the parity tests."""
    import torch
    import hyowon_amd as H
    n_side, N, L = args.n_side, args.mesh, args.box
    n = n_side ** 3
    dev = pos[0].device
    acc = {}
    for prec in ("fp32", "fp64"):
        s = H.PMSolver(L, N, precision=prec, max_particles=n)
        g = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
        s.compute_forces(pos[0], pos[1], pos[2], None, m, *g)
        acc[prec] = torch.stack(g)
        s.close()
        del g
    d2 = ((acc["fp32"] - acc["fp64"]) ** 2).sum(0)
    g2 = (acc["fp64"] ** 2).sum(0)
    rms = math.sqrt(float(d2.sum() / g2.sum()))
    rel = torch.sqrt(d2 / g2.clamp_min(1e-300)).float()
    gen = torch.Generator(device=dev).manual_seed(3)
    samp = rel[torch.randint(0, n, (min(n, 1 << 23),), device=dev, generator=gen)]
    q = torch.quantile(samp, torch.tensor([0.5, 0.99], device=dev)).tolist()
    del acc, d2, g2, rel, samp
    torch.cuda.empty_cache()

    final = {}
    for prec in ("fp32", "fp64"):
        mdt = torch.float32 if prec == "fp32" else torch.float64
        s = H.PMSolver(L, N, precision=prec, max_particles=n)
        st = H.ParticleState(pos[0].clone(), pos[1].clone(), pos[2].clone(),
                             mom[0].to(mdt).contiguous(), mom[1].to(mdt).contiguous(),
                             mom[2].to(mdt).contiguous())
        sp = st.empty_like()
        for i in range(nsteps):  # the final K2 kick does not move particles
            s.kdk(st, sp, m, K2[i - 1] if i > 0 else None, K1[i], D[i], asynchronous=True,
                  reuse_keys=True)
            st, sp = sp, st
        s.check()
        # MatterPowerSpectrum of the final positions (binary64 estimator, 32
        # logarithmic shells up to half the Nyquist wavenumber, interlaced)
        s64 = H.PMSolver(L, N, precision="fp64", max_particles=n)
        bins = s64.power_spectrum(st.x, st.y, st.z, 32, False, True, 0.5, uniform_mass=m)
        s64.close()
        final[prec] = (torch.stack([st.x, st.y, st.z]),
                       [(b.k_mean, b.p_corrected, int(b.mode_count)) for b in bins])
        s.close()
        del sp
        torch.cuda.empty_cache()
    dx = L / N
    dpos = final["fp32"][0] - final["fp64"][0]
    dpos = dpos - L * torch.round(dpos / L)
    dist = torch.sqrt((dpos ** 2).sum(0)) / dx
    pos_rms = float(torch.sqrt((dist ** 2).mean()))
    pos_max = float(dist.max())
    del dpos, dist
    p32, p64 = final["fp32"][1], final["fp64"][1]
    rel = [abs(a[1] / b[1] - 1.0) for a, b in zip(p32, p64) if b[2] and b[1] != 0.0]
    return {"force_rms_rel_fp32_vs_fp64": rms, "force_rel_median": q[0], "force_rel_p99": q[1],
            "force_at": f"initial conditions, {n} particles, {N}^3 mesh",
            "final_pos_err_cells_rms": pos_rms, "final_pos_err_cells_max": pos_max,
            "pk_max_rel_diff_k_lt_half_nyquist": max(rel),
            "pk_final_fp32": {"k_mean": [b[0] for b in p32], "p_corrected": [b[1] for b in p32],
                              "mode_count": [b[2] for b in p32],
                              "estimator": "hy_pm_power_spectrum, binary64, 32 log shells to 0.5 k_Nyq, "
                                           "interlaced, no shot-noise subtraction"},
            "run": f"{nsteps} KDK steps z={args.z_start:g}->0 in each precision"}


def run_ensemble(args):
    """Config-3 workload: 192 boxes (64 x 32^3 particles on 64^3, 64 x 64^3 on
    128^3, 64 x 96^3 on 192^3; box side twice the mesh), seeds 1..192,
    z = 49 -> 0 in 32 KDK steps, one graph launch per ensemble step.  Times the
    steady graph steps; also times the same steps as eager per-context calls,
    and reports the FP32-vs-FP64 final-position difference per class."""
    import torch
    import hyowon_amd as H
    from hyowon_amd.ensemble import CONFIG3, EnsembleRun
    steps = 32
    a0 = 1.0 / (1.0 + args.z_start)
    res = {}
    finals = {}
    for prec in ("fp32", "fp64"):
        run = EnsembleRun(CONFIG3, a0, 1.0, steps, precision=prec)
        run.first_step()
        run.ensemble.check()
        torch.cuda.synchronize()
        timed = min(args.steps, steps - 2)
        run.next_step()  # first replay records the graph
        run.ensemble.check()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(timed):
            run.next_step()
        run.ensemble.check()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        graph_ms = (t1 - t0) * 1e3 / timed
        # the same kind of step as eager per-context launches (no graph)
        i = run.done
        t0 = time.perf_counter()
        for s, st, m in zip(run.solvers, run.states, run.masses):
            s.kdk(st, st, m, run.K2[i - 1], run.K1[i], run.D[i], asynchronous=True, reuse_keys=True)
        for s in run.solvers:
            s.check()
        torch.cuda.synchronize()
        eager_ms = (time.perf_counter() - t0) * 1e3
        run.done += 1
        while run.done <= steps:
            run.next_step()
        run.ensemble.check()
        finals[prec] = [torch.stack([st.x, st.y, st.z]) for st in run.states]
        res[prec] = {"graph_step_ms": graph_ms, "eager_step_ms": eager_ms}
        run.close()
        torch.cuda.empty_cache()
    n = sum(c.particles for c in CONFIG3)
    alg = sum(c.boxes * (96.0 * c.n_side ** 3 + 24.0 * c.mesh ** 3) for c in CONFIG3)
    pos_err = {}
    for cls, a, b in zip(CONFIG3, finals["fp32"], finals["fp64"]):
        d = a - b
        d = d - cls.box_size * torch.round(d / cls.box_size)
        dist = torch.sqrt((d ** 2).sum(0)) / (cls.box_size / cls.mesh)
        pos_err[f"{cls.n_side}^3/{cls.mesh}^3"] = {"rms_cells": float(torch.sqrt((dist ** 2).mean())),
                                                 "max_cells": float(dist.max())}
    r = res["fp32"]
    return {
        "metric": "particle-updates/s per ensemble KDK step (192 boxes, config 3)",
        "value": n / (r["graph_step_ms"] * 1e-3), "unit": "particle-updates/s", "n_gpus": 1,
        "steps": min(args.steps, steps - 2), "ms_per_step": r["graph_step_ms"], "higher_is_better": True,
        "dtype": "f32 mesh/FFT/forces/momenta, f64 positions", "data": "synthetic 1LPT ICs, seeds 1..192",
        "config": {"workload": "ensemble KDK step: 64 x (32^3, 64^3 mesh, L=128), 64 x (64^3, 128^3, L=256), "
                               "64 x (96^3, 192^3, L=384), 32 steps z=49->0, one HIP graph launch per step",
                   "particles": n},
        "step_hbm_frac": alg / (r["graph_step_ms"] * 1e-3) / HBM_PEAK,
        "eager_ms_per_step": r["eager_step_ms"],
        "fp64": {"ms_per_step": res["fp64"]["graph_step_ms"], "eager_ms_per_step": res["fp64"]["eager_step_ms"]},
        "final_pos_err_fp32_vs_fp64": pos_err,
    }


def run_pk(args):
    """Power-spectrum workload: the config-2 particle load (512^3 1LPT
    particles at z = 49, seed 1, L = 512 Mpc/h) on a 1024^3 CIC mesh, 32
    logarithmic shells up to 0.5 k_Nyq, interlaced.  Times hy_pm_power_spectrum
    (deposits, r2c transforms, shell sums) per precision and reports the shells,
    the FP32-vs-FP64 shell differences and the ratio to the linear input
    spectrum D(a)^2 P_lin(k)."""
    import numpy as np
    import torch
    import hyowon_amd as H
    from hyowon_amd import ic
    n_side, N, L = args.n_side, args.mesh, args.box
    n = n_side ** 3
    a0 = 1.0 / (1.0 + args.z_start)
    dev = torch.device("cuda:0")
    pos, mom = ic.lpt_ics(n_side, L, a0, seed=1, device=dev, rng="philox", max_mode=n_side // 2 - 1)
    del mom
    m = ic.particle_mass(n_side, L)
    nb, frac = 32, 0.5
    res = {}
    for prec in args.precisions:
        s = H.PMSolver(L, N, precision=prec, max_particles=n)
        reps = args.steps if prec == args.precisions[0] else max(2, args.fp64_steps)
        for _ in range(max(1, args.warmup)):
            s.power_spectrum(*pos, nb, False, True, frac, uniform_mass=m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            bins = s.power_spectrum(*pos, nb, False, True, frac, uniform_mass=m)
        torch.cuda.synchronize()
        res[prec] = {"ms": (time.perf_counter() - t0) * 1e3 / reps, "reps": reps,
                     "bins": [[b.k_mean, b.p_corrected, b.p_raw, int(b.mode_count)] for b in bins]}
        print(f"pk {prec}: {res[prec]['ms']:.2f} ms per estimate", file=sys.stderr, flush=True)
        s.close()
        torch.cuda.empty_cache()
    cosmo = ic.Cosmology()
    main_prec = args.precisions[0]
    rows = res[main_prec]["bins"]
    ratio = [b[1] / float(cosmo.power(np.array([b[0]]), a0)[0]) if b[3] else None for b in rows]
    out = {
        "metric": "matter power-spectrum estimates/s (512^3 particles, 1024^3 mesh, interlaced)",
        "value": 1e3 / res[main_prec]["ms"], "unit": "estimates/s", "n_gpus": 1,
        "steps": res[main_prec]["reps"], "warmup": args.warmup, "ms_per_step": res[main_prec]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 mesh/FFT, f64 shell sums" if main_prec == "fp32" else "f64",
        "data": "synthetic 1LPT ICs (Philox mode-keyed field, BBKS, sigma8=0.8, z=49, seed 1, K=n/2-1)",
        "config": {"workload": "MatterPowerSpectrum::compute_pk on the config-2 particle load: "
                               "512^3 particles, 1024^3 mesh, L=512, 32 shells to 0.5 k_Nyq, interlaced",
                   "particles": n, "mesh": N},
        "shells": {"k_mean": [b[0] for b in rows], "p_corrected": [b[1] for b in rows],
                   "mode_count": [b[3] for b in rows], "ratio_to_linear": ratio},
    }
    if "fp64" in res and main_prec != "fp64":
        r64 = res["fp64"]["bins"]
        rel = [abs(a[1] / b[1] - 1.0) for a, b in zip(rows, r64) if b[3] and b[1] != 0.0]
        out["fp64"] = {"ms_per_estimate": res["fp64"]["ms"], "max_rel_shell_diff_fp32_vs_fp64": max(rel)}
    return out


def run_treepm(args):
    """TreePM long-range PM force (HY_METHOD_TREEPM_LONG_RANGE, r_s = 1.25 cells)
    on the config-2 particle load (512^3 1LPT particles, 1024^3 mesh, L = 512):
    one hy_pm_force per call (deposit, r2c, TreePM Green, three spectral
    derivatives and c2r transforms, CIC interpolation of three force meshes),
    timed beside the pure-PM force refresh of the same particles."""
    import torch
    import hyowon_amd as H
    from hyowon_amd import ic
    n_side, N, L = args.n_side, args.mesh, args.box
    n = n_side ** 3
    a0 = 1.0 / (1.0 + args.z_start)
    dev = torch.device("cuda:0")
    pos, mom = ic.lpt_ics(n_side, L, a0, seed=1, device=dev, rng="philox", max_mode=n_side // 2 - 1)
    del mom
    m = ic.particle_mass(n_side, L)
    rs = 1.25 * L / N
    res = {}
    for prec in args.precisions:
        for method in ("treepm_long_range", "pure_pm"):
            s = H.PMSolver(L, N, precision=prec, max_particles=n, method=method, split_scale=rs)
            acc = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
            reps = args.steps if prec == args.precisions[0] else max(2, args.fp64_steps)
            for _ in range(max(1, args.warmup)):
                s.compute_forces(*pos, None, m, *acc)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                s.compute_forces(*pos, None, m, *acc)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            res[f"{prec}_{method}"] = ms
            print(f"{prec} {method}: {ms:.2f} ms per force", file=sys.stderr, flush=True)
            s.close()
            del acc
            torch.cuda.empty_cache()
    main_prec = args.precisions[0]
    ms = res[f"{main_prec}_treepm_long_range"]
    return {
        "metric": "TreePM long-range PM force evaluations/s (512^3 particles, 1024^3 mesh)",
        "value": n / (ms * 1e-3), "unit": "particle-forces/s", "n_gpus": 1, "ms_per_step": ms,
        "higher_is_better": True, "dtype": "f32 mesh/FFT/forces" if main_prec == "fp32" else "f64",
        "data": "synthetic 1LPT ICs (Philox mode-keyed field, BBKS, sigma8=0.8, z=49, seed 1, K=n/2-1)",
        "config": {"workload": "hy_pm_force with PMForceMethod::treepm_long_range(r_s = 1.25 dx) on the "
                               "config-2 particle load: 512^3 particles, 1024^3 mesh, L=512",
                   "particles": n, "mesh": N, "split_scale": rs},
        "ms_per_force": {k: round(v, 3) for k, v in res.items()},
    }


def run_adjoint(args):
    """Config-4 workload: 128^3 particles from a 128^3 white-noise field (seed
    7), 256^3 mesh, L = 256 Mpc/h, 16 KDK steps z = 19 -> 0; loss = mean
    squared difference between the final CIC density contrast and that of the
    seed-8 field run through the binary32 forward; hand-written adjoint with
    checkpoints every 4 refreshes.  Reports forward / backward times and the
    gradient checks: central differences (eps = 1e-3 as specified, 1e-4 and
    1e-5) on 64 random noise components in binary64, and the binary32
    gradient against the binary64 reference gradient of the same loss.  The
    loss is piecewise smooth (CIC weights kink where a particle crosses a mesh
    plane): the central-difference deviation from the adjoint shrinks with eps
    as fewer particles cross a plane inside the differencing interval."""
    import numpy as np
    import torch
    import hyowon_amd as H
    from hyowon_amd import ic
    from hyowon_amd.adjoint import cic_density, density_loss, pm_run
    from hyowon_amd.ensemble import schedule
    n_side, N, L, steps = 128, 256, 256.0, 16
    a0 = 1.0 / 20.0
    K1, D, K2 = schedule(a0, 1.0, steps)
    dev = torch.device("cuda:0")
    noise = ic.white_noise(n_side, 7, dev).to(torch.float64)
    out = {}
    grads = {}
    m = ic.particle_mass(n_side, L)
    s = H.PMSolver(L, N, precision="fp32", max_particles=n_side ** 3)
    with torch.no_grad():  # target: the seed-8 field through the binary32 forward
        pos, mom = ic.lpt_from_noise(ic.white_noise(n_side, 8, dev).to(torch.float64), L, a0)
        st = pm_run((pos[0], pos[1], pos[2], *(t.to(torch.float32) for t in mom)), s, m, K1, D, K2, 4)
        target = cic_density(st[0], st[1], st[2], s, m).to(torch.float64)
    s.close()
    for prec in ("fp32", "fp64"):  # the binary32 gradient and the binary64 reference gradient of one loss
        s = H.PMSolver(L, N, precision=prec, max_particles=n_side ** 3)
        warm = noise.clone().requires_grad_(True)  # first launches load code objects
        density_loss(warm, target, s, L, a0, K1, D, K2, checkpoint_every=4, precision=prec).backward()
        nz = noise.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = density_loss(nz, target, s, L, a0, K1, D, K2, checkpoint_every=4, precision=prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        grads[prec] = nz.grad.detach().reshape(-1).clone()
        out[prec] = {"loss": float(loss), "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3}
        if prec == "fp64":
            rng = np.random.default_rng(2024)
            idx = rng.choice(noise.numel(), size=64, replace=False)
            flat = noise.reshape(-1)
            for eps in (1e-3, 1e-4, 1e-5):
                errs, fds = [], []
                for i in idx:
                    lp = flat.clone()
                    lp[i] += eps
                    lm = flat.clone()
                    lm[i] -= eps
                    with torch.no_grad():
                        Lp = float(density_loss(lp.view_as(noise), target, s, L, a0, K1, D, K2,
                                                precision="fp64"))
                        Lm = float(density_loss(lm.view_as(noise), target, s, L, a0, K1, D, K2,
                                                precision="fp64"))
                    fds.append((Lp - Lm) / (2 * eps))
                    if len(fds) % 16 == 0:
                        print(f"finite differences eps={eps:g}: {len(fds)}/64", file=sys.stderr, flush=True)
                g = grads["fp64"][torch.as_tensor(idx, device=dev)].cpu().numpy()
                fds = np.array(fds)
                rel = np.abs(g - fds) / np.abs(fds)
                out[f"fd_eps_{eps:g}"] = {"median_rel_err": float(np.median(rel)), "max_rel_err": float(rel.max()),
                                          "l2_rel_err": float(np.linalg.norm(g - fds) / np.linalg.norm(fds)),
                                          "components": 64}
        s.close()
        torch.cuda.empty_cache()
    g32, g64 = grads["fp32"], grads["fp64"]
    out["fp32_vs_fp64_grad_l2_rel"] = float(torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64))
    return {"metric": "differentiable PM run: forward + hand-written adjoint (config 4)",
            "value": out["fp32"]["forward_ms"] + out["fp32"]["backward_ms"], "unit": "ms",
            "higher_is_better": False, "n_gpus": 1,
            "config": {"workload": "128^3 particles from white noise (seed 7), 256^3 mesh, L=256, 16 KDK steps "
                                   "z=19->0, density-MSE loss vs the seed-8 forward, checkpoints every 4"},
            "results": out}


def cpu_baseline(args):
    """The oracle (binary64 C, OpenMP over all host cores) on a bounded sample of
    the same workload: one force refresh + K1/drift + K2 of a 128^3-particle,
    256^3-mesh box at the same particle and mesh spacing as the GPU workload."""
    import numpy as np

    import oracle
    from hyowon_amd import ic
    oracle.build(ref=False) if not os.path.exists(oracle.LIB_PATH) else None
    cores = os.cpu_count() or 1
    oracle.set_threads(cores)
    ns = args.cpu_sample_side
    Ns = ns * args.mesh // args.n_side
    Ls = args.box * ns / args.n_side
    a0 = 1.0 / (1.0 + args.z_start)
    # the same Philox mode-keyed field as the GPU workload, drawn on the GPU and copied out
    pos, mom = ic.lpt_ics(ns, Ls, a0, seed=1, device="cuda", rng="philox", max_mode=ns // 2 - 1)
    pos = [p.cpu().numpy() for p in pos]
    mom = [p.cpu().numpy() for p in mom]
    m = ic.particle_mass(ns, Ls)
    b = np.array([a0, a0 * math.exp(math.log(1.0 / a0) / args.schedule_steps)])
    oracle.stage_reset()
    t0 = time.perf_counter()
    oracle.run(pos[0], pos[1], pos[2], mom[0], mom[1], mom[2], Ls, Ns, b, m)
    t1 = time.perf_counter()
    # the run does 2 force refreshes for 1 step; a steady step has 1 refresh
    step_s = (t1 - t0) / 2.0
    stages = oracle.stage_times()
    stage_ms = {k: 1e3 * (v if k == "kdk" else v / 2.0) for k, v in stages.items()}
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": ns ** 3 / step_s, "unit": "particle-updates/s", "cores": cores,
            "kind": "port", "cpu_model": model, "fft": "the oracle's own binary64 FFT",
            "stage_ms_per_step": stage_ms,
            "sample": f"oracle (binary64 C/OpenMP) one KDK step of {ns}^3 particles on a "
                      f"{Ns}^3 mesh, box {Ls:g} Mpc/h (same spacing as the GPU workload); "
                      f"time of 2 force refreshes + kicks + drift halved; {t1 - t0:.1f} s total"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-side", type=int, default=512)
    ap.add_argument("--mesh", type=int, default=1024)
    ap.add_argument("--box", type=float, default=512.0)
    ap.add_argument("--z-start", type=float, default=49.0)
    ap.add_argument("--schedule-steps", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--precisions", default="fp32,fp64")
    ap.add_argument("--cpu-sample-side", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--accuracy", action=argparse.BooleanOptionalAction, default=True,
                    help="FP32-vs-FP64 force, final-position and P(k) errors (full run)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="measure the FFT launch plan before timing (untimed)")
    ap.add_argument("--workload", default="config2",
                    choices=["config2", "ensemble", "adjoint", "pk", "treepm"],
                    help="config2: the 512^3 / 1024^3 KDK step (default); ensemble: the 192-box "
                         "graph-launched ensemble step; adjoint: the differentiable 128^3 / 256^3 run; "
                         "pk: the power-spectrum estimator on the config-2 particles; treepm: the "
                         "TreePM long-range force on the config-2 particles")
    ap.add_argument("--fft-plan", default="",
                    help="pin the FFT launch plan 'v0,v1,v2,v3,v4[:plane_group]' instead of autotuning")
    ap.add_argument("--plain-fp64", action=argparse.BooleanOptionalAction, default=True,
                    help="time the unfused binary64 pipeline (rocFFT) the FP32 step is compared with")
    ap.add_argument("--rocfft", action=argparse.BooleanOptionalAction, default=True,
                    help="time rocFFT r2c/c2r beside the hand-written FFT passes")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="bytes of a plain device copy run first (counter calibration)")
    args = ap.parse_args()
    args.precisions = [p for p in args.precisions.split(",") if p]
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(a slab decomposition over GPUs is not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    if args.workload == "pk":
        print(json.dumps(run_pk(args)), flush=True)
        return
    if args.workload == "treepm":
        print(json.dumps(run_treepm(args)), flush=True)
        return
    n, N, res = run_gpu(args)
    main_prec = args.precisions[0]
    r = res[main_prec]
    value = n / (r["step_ms"] * 1e-3)
    stage_ms = r["stage_ms"]
    dom = max(stage_ms, key=stage_ms.get)
    ach = stage_bytes(dom, n, N, r["real"]) / (stage_ms[dom] * 1e-3)
    traffic, traffic_src = measured_traffic(dom, args.n_side, N, main_prec, r["plan"])
    roofline = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src, "kernel": dom,
                "kernel_ms": stage_ms[dom],
                "alg_bytes_per_launch": stage_bytes(dom, n, N, r["real"])}
    out = {
        "metric": "particle-updates/s per PM KDK step (512^3 particles, 1024^3 mesh)",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 mesh/FFT/forces/momenta, f64 positions" if main_prec == "fp32" else "f64",
        "data": "synthetic 1LPT ICs (Philox mode-keyed field, BBKS, sigma8=0.8, z=49, seed 1, K=n/2-1)",
        "config": {"workload": "PM KDK step, config 2: 512^3 particles, 1024^3 CIC mesh, "
                               "L=512 Mpc/h, 64 log-a steps z=49->0",
                   "particles": n, "mesh": N, "box": args.box, "precision": main_prec},
        "roofline": roofline,
        "step_hbm_frac": step_alg_bytes(n, N, r["real"]) / (r["step_ms"] * 1e-3) / HBM_PEAK,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "fft_plan": r["plan"],
    }
    if "fp64" in res and main_prec != "fp64":
        r64 = res["fp64"]
        out["fp64"] = {"ms_per_step": r64["step_ms"], "value": n / (r64["step_ms"] * 1e-3),
                       "speedup_fp32_vs_fp64": r64["step_ms"] / r["step_ms"],
                       "step_hbm_frac": step_alg_bytes(n, N, 8) / (r64["step_ms"] * 1e-3) / HBM_PEAK,
                       "stage_ms": {k: round(v, 4) for k, v in r64["stage_ms"].items()}}
    if "plain_fp64" in res:
        pf = res["plain_fp64"]
        out["plain_fp64"] = dict(pf, value=n / (pf["ms_per_step"] * 1e-3),
                                 speedup_fp32_vs_plain_fp64=pf["ms_per_step"] / r["step_ms"])
    if "rocfft" in res:
        fft_ms = sum(v for k, v in stage_ms.items() if k.startswith("fft_"))
        out["fft_compare"] = dict(res["rocfft"], hip_pencil_passes_ms=fft_ms,
                                  note="hip passes include the Green multiply and 1/N^3")
    if "accuracy" in res:
        out["accuracy"] = res["accuracy"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
