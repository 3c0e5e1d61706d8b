"""Benchmark: batched FP32 acoustic FWI gradient (forward + hand-written adjoint) on one MI355X.

Workload (the metric's configuration): 2-D acoustic, 1000 x 1000 interior cells, h = 5 m,
dt = 0.5 ms, 4000 steps, 40-cell PML, random layered velocity model (8-20 layers,
1500-4500 m/s increasing with depth, 3 Gaussian anomalies of +-15%, seed 2), 32 Ricker shots
(8-15 Hz) evenly spaced along the surface, 1000 receivers per shot.  Observed data are simulated
from the true model; the gradient is taken at a smoothed starting model.  One "step" is one
complete FWI gradient of all 32 shots: forward with checkpoints, misfit, checkpoint
recomputation and the adjoint sweep, all in one batched call.

Prints ONE JSON line.  Besides value/ms_per_step it carries:
  roofline      : the dominant stencil kernel's algorithmic HBM bytes per launch / its average
                  launch duration (HIP events on the launch stream, over the timed region);
  cpu_baseline  : the FP64 CPU restatement (oracle/) on the process's threads, with a one-thread
                  run beside it (like the reference's serial CPU op), on a bounded sample of
                  the same workload;
  plain_single_shot : the plain drop-in path (AcousticPropagatorSolver + SimulatedObservation_
                  + autograd, full histories, one shot) on the same grid.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_acoustic_1000x1000_32shots.txt")


# Kernel shape of the acoustic plan on the benchmark workload: forward 8 rows per wavefront (the
# streaming two-level rectangle, one shot per wavefront; 16 shots per workgroup for the frame
# items), adjoint 2 rows per wavefront, 16 shots per workgroup, 1 shot per batch.  The fastest
# of the shapes timed back to back on one MI355X (forward 96-98 us and adjoint 165 us per level
# at 1000 x 1000 x 32 shots; 4, 12 or 16 streaming rows within 1-4%; 32 shots per frame group
# the same within noise, 8 and 4 slower); it is also the plan's default shape, and
# ads_acoustic_plan_autotune picks shapes within a few percent of it.
DEFAULT_SHAPE = "8,16,1,2,16,1"


def measured_traffic(kernel, shape):
    """HBM bytes per launch of a stencil class, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain device copy of known size).
    Lines: '<class> <bytes> <rows,shots,batch>'.  Returns (bytes, profiled shape); bytes is None
    when the summary is missing or was taken at another kernel shape."""
    try:
        with open(TRAFFIC_FILE) as f:
            lines = [ln.split() for ln in f if ln.strip() and not ln.startswith("#")]
    except OSError:
        return None, None
    kv = {ln[0]: ln[1:] for ln in lines if len(ln) >= 3}
    if kernel not in kv:
        return None, None
    prof_shape = tuple(int(x) for x in kv[kernel][1].split(","))
    half = None if shape is None else tuple(shape[:3] if kernel == "ac_fwd_step" else shape[3:])
    if half != prof_shape:
        return None, prof_shape
    return float(kv[kernel][0]), prof_shape


def velocity_model(NX, NY, seed):
    """Random layered model (depth = j), 1500-4500 m/s increasing with depth, + 3 anomalies."""
    rng = np.random.default_rng(seed)
    nl = int(rng.integers(8, 21))
    depths = np.sort(rng.choice(np.arange(1, NY + 1), size=nl - 1, replace=False))
    vels = np.sort(rng.uniform(1500.0, 4500.0, size=nl))
    c = np.empty((NX + 2, NY + 2))
    layer = np.searchsorted(depths, np.arange(NY + 2), side="right")
    c[:] = vels[layer][None, :]
    ii, jj = np.meshgrid(np.arange(NX + 2), np.arange(NY + 2), indexing="ij")
    for _ in range(3):
        ci, cj = rng.uniform(1, NX), rng.uniform(1, NY)
        sg = rng.uniform(20, 80)
        amp = rng.choice([-1.0, 1.0]) * 0.15
        c *= 1.0 + amp * np.exp(-((ii - ci) ** 2 + (jj - cj) ** 2) / (2 * sg * sg))
    return c


def workload(ads, small=False, nstep=0, kernel=1):
    import torch
    from scipy.ndimage import gaussian_filter

    if small:
        NX = NY = 200
        NSTEP, nshots, npml = 600, 4, 20
    else:
        NX = NY = 1000
        NSTEP, nshots, npml = 4000, 32, 40
    if nstep:
        NSTEP = nstep
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=5.0, DELTAY=5.0,
                                     DELTAT=0.5e-3, NPOINTS_PML=npml, vp_ref=3000.0,
                                     PropagatorKernel=kernel)
    c_true = velocity_model(NX, NY, seed=2)
    c0 = gaussian_filter(c_true, 10.0, mode="reflect")
    ads.check_cfl(p, float(c_true.max()))
    rng = np.random.default_rng(2)
    xs = np.round(np.linspace(npml + 2, NX + 1 - npml, nshots)).astype(np.int64)
    f0s = rng.uniform(8.0, 15.0, nshots)
    srcs, rcvs = [], []
    for s in range(nshots):
        w = ads.ricker_freq(f0s[s], p.DELTAT, NSTEP, 1e4)
        srcs.append(ads.AcousticSource([int(xs[s])], [3], w.reshape(-1, 1)))
        rcvs.append(ads.AcousticReceiver(np.arange(2, NX + 2), np.full(NX, 3)))
    return p, c_true, c0, srcs, rcvs


def workload_uneven(ads, seed=4, nshots=64, NX=2000, nstep_range=(3000, 6000)):
    """2-D acoustic FWI gradient workload with uneven shots: NX x NX interior cells, h = 5 m,
    dt = 0.5 ms (CFL 0.5 at 5000 m/s), 40-cell PML, smooth random velocity 1500-5000 m/s
    (Gaussian-filtered noise, sigma 40 cells), `nshots` Ricker shots (8-12 Hz) at depth 3 evenly
    spaced along the surface, record lengths uniform in nstep_range, spreads of 400-2000
    receivers centred on each source.  Starting model: sigma = 30 cell smoothing.
    Returns (params, c_true, c0, sources, receivers, record lengths)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    NY = NX
    dt, h = 5e-4, 5.0
    nst = rng.integers(nstep_range[0], nstep_range[1] + 1, nshots)
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=int(nst.max()), DELTAX=h, DELTAY=h,
                                     DELTAT=dt, NPOINTS_PML=40, vp_ref=3000.0, PropagatorKernel=1)
    noise = gaussian_filter(rng.standard_normal((NX + 2, NY + 2)), 40.0)
    noise = (noise - noise.min()) / (noise.max() - noise.min())
    c_true = 1500.0 + 3500.0 * noise
    ads.check_cfl(p, float(c_true.max()))
    c0 = gaussian_filter(c_true, 30.0)
    xs = np.round(np.linspace(41, NX - 40, nshots)).astype(int)
    srcs, rcvs = [], []
    for s in range(nshots):
        w = ads.ricker_freq(8.0 + 4.0 * rng.random(), dt, int(nst[s]), 1e4)
        srcs.append(ads.AcousticSource([int(xs[s])], [3], w.reshape(-1, 1)))
        nr = int(rng.integers(400, 2001))
        lo = int(np.clip(xs[s] - nr // 2, 1, NX + 2 - nr))
        rcvs.append(ads.AcousticReceiver(np.arange(lo, lo + nr), np.full(nr, 3)))
    return p, c_true, c0, srcs, rcvs, nst


def _cpu_gradient(O, p, c0, src, nsteps, threads):
    """Oracle forward + adjoint of one shot over the first nsteps steps on `threads` threads
    (1: the serial checker itself, scatter-form step VJPs; more: the gather-form sweep)."""
    O.set_num_threads(threads)
    NX, NY = p.NX, p.NY
    Sx, Sy = O.acoustic_pml(NX, NY, p.DELTAX, p.DELTAY, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    srcv = np.asarray(src.srcv)[:nsteps]
    rcvi, rcvj = np.arange(2, NX + 2), np.full(NX, 3)
    t0 = time.perf_counter()
    u, ph, ps = O.acoustic_forward(1, NX, NY, nsteps, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy,
                                   c0 ** 2, src.srci, src.srcj, srcv)
    tr = O.acoustic_receive(u, rcvi, rcvj, NY)
    gu = np.zeros((nsteps + 1, (NX + 2) * (NY + 2)))
    gu[:, O.flat_index(rcvi, rcvj, NY)] = 2.0 * tr
    if threads == 1:
        O.acoustic_adjoint(1, NX, NY, nsteps, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy, c0 ** 2, u,
                           ph, ps, gu)
    else:
        O.acoustic_adjoint_par(1, NX, NY, nsteps, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy, c0 ** 2, u, gu)
    t = time.perf_counter() - t0
    return NX * NY * (nsteps - 1) / t / 1e9, t


def cpu_baseline(p, c0, src, nsteps):
    """Oracle (FP64) forward + adjoint of one shot over the first nsteps steps, on all the
    threads this process may use (OMP_NUM_THREADS caps the affinity set) and on one thread,
    like the reference's serial custom op."""
    import oracle as O

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = max(1, min(threads, int(os.environ["OMP_NUM_THREADS"])))
    sample = (f"1 shot, {p.NX}x{p.NY} grid, first {nsteps} of {p.NSTEP} steps, FP64 "
              "forward+adjoint (oracle/ads_oracle.c, PropagatorKernel 1)")
    v1, t1 = _cpu_gradient(O, p, c0, src, nsteps, 1)
    out = {"value": v1, "unit": "Gcell-steps/s", "cores": 1, "kind": "port",
           "sample": f"{sample}, one thread, {t1:.1f} s"}
    if threads > 1:
        vm, tm = _cpu_gradient(O, p, c0, src, nsteps, threads)
        out = {"value": vm, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{sample}, {threads} threads (OpenMP rows, gather-form adjoint), {tm:.1f} s",
               "single_thread": out}
    O.set_num_threads(1)
    out["host_cpu"] = platform.processor() or platform.machine()
    out["host_threads_available"] = len(os.sched_getaffinity(0))
    return out


def elastic_workload(ads, nstep=0):
    """Elastic FP32 workload: 600 x 600, h = 10 m, dt = 0.8 ms, 3000 steps, 30-cell CPML.
    Layered Vp 2000-4500 m/s (seed 3) with a 5% smooth random perturbation, Vs = Vp/1.73,
    rho = 310 Vp^0.25 (Gardner), lambda = rho (Vp^2 - 2 Vs^2), mu = rho Vs^2.  16 explosive
    sources (sxx and syy, the same 12 Hz Ricker) evenly spaced below the top CPML, 600 receivers
    recording vx and vy.  The gradient is taken at a sigma = 10 cell smoothing of the model."""
    from scipy.ndimage import gaussian_filter

    NX = NY = 600
    NSTEP = nstep or 3000
    npml, nshots = 30, 16
    p = ads.ElasticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=10.0, DELTAY=10.0,
                                    DELTAT=0.8e-3, f0=12.0, vp_ref=4500.0, NPOINTS_PML=npml)
    rng = np.random.default_rng(3)
    nl = 10
    depths = np.sort(rng.choice(np.arange(1, NY + 1), size=nl - 1, replace=False))
    vels = np.sort(rng.uniform(2000.0, 4500.0, size=nl))
    vp = np.repeat(vels[np.searchsorted(depths, np.arange(NY + 2), side="right")][None, :], NX + 2, 0)
    noise = gaussian_filter(rng.standard_normal((NX + 2, NY + 2)), 8.0)
    vp = vp * (1.0 + 0.05 * np.clip(noise / noise.std(), -1.0, 1.0))

    def lame(v):
        vs = v / 1.73
        rho = 310.0 * v ** 0.25
        return rho, rho * (v * v - 2.0 * vs * vs), rho * vs * vs

    true = lame(vp)
    start = lame(gaussian_filter(vp, 10.0, mode="reflect"))
    w = ads.ricker_freq(12.0, p.DELTAT, NSTEP, 1e6)
    xs = np.round(np.linspace(npml + 2, NX + 1 - npml, nshots)).astype(np.int64)
    srcs, rcvs = [], []
    for s in range(nshots):
        srcs.append(ads.ElasticSource([int(xs[s])] * 2, [npml + 3] * 2, [2, 3],
                                      np.stack([w, w], 1)))
        ri = np.concatenate([np.arange(2, NX + 2)] * 2)
        rcvs.append(ads.ElasticReceiver(ri, np.full(2 * NX, npml + 3), np.repeat([0, 1], NX)))
    return p, true, start, srcs, rcvs


def elastic_leg(ads, L, lib, steps, nstep=0):
    """Batched FP32 elastic FWI gradient (d/d rho, lambda, mu) throughput and kernel roofline."""
    import ctypes as C
    import torch

    p, true, start, srcs, rcvs = elastic_workload(ads, nstep)
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float32)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    K = plan.choose_interval(int(torch.cuda.mem_get_info()[0] * 0.80))
    tt = [torch.as_tensor(x, dtype=torch.float32, device="cuda") for x in true]
    st = [torch.as_tensor(x, dtype=torch.float32, device="cuda") for x in start]
    obs = plan.forward(*tt, srcv, K).clone()
    plan.gradient(*st, srcv, obs, K)
    torch.cuda.synchronize()
    L.check(lib.ads_profile_begin())
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, grads, _ = plan.gradient(*st, srcv, obs, K)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    L.check(lib.ads_profile_end())
    cells = float(p.NX * p.NY * np.sum(plan.nstep))
    kern = {}
    for name, cls in (("el_fwd_stress", 2), ("el_fwd_vel", 3), ("el_adj_vel", 4), ("el_adj_stress", 5)):
        tm, n, b = C.c_double(), C.c_int64(), C.c_double()
        L.check(lib.ads_profile_query(cls, C.byref(tm), C.byref(n), C.byref(b)))
        kern[name] = {"total_ms": round(tm.value, 2), "launches": n.value,
                      "avg_us": round(tm.value / n.value * 1e3, 2) if n.value else None,
                      "alg_GBps": round(b.value / (tm.value * 1e-3) / 1e9, 1) if tm.value else None}
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    ach = kern[dom]["alg_GBps"] or 0.0
    out = {"workload": (f"elastic {p.NX}x{p.NY}, h=10 m, dt=0.8 ms, {p.NSTEP} steps, "
                        f"{p.NPOINTS_PML}-cell CPML, {plan.nshots} explosive shots x {p.NX} "
                        "two-component receivers, FP32 forward + hand-written adjoint for "
                        "(rho, lambda, mu), one batched call per gradient"),
           "value": round(cells / t / 1e9, 4), "unit": "Gcell-steps/s",
           "ms_per_gradient": round(t * 1e3, 2), "checkpoint_interval": K,
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBPS,
                        "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBPS, 4)},
           "kernels": kern, "loss": float(loss.item())}
    plan._ws = None
    torch.cuda.empty_cache()
    return out


def plain_single_shot(ads, p, c0, src, rcv, obs_shot):
    """The plain drop-in: history solver + SimulatedObservation_ + torch autograd, one shot."""
    import torch

    ct = torch.as_tensor(c0, dtype=torch.float32, device="cuda")
    obs = torch.as_tensor(obs_shot, dtype=torch.float32, device="cuda")

    def once():
        cc = ct.clone().requires_grad_(True)
        ap = ads.AcousticPropagatorSolver(p, src, cc)
        r = ads.SimulatedObservation_(ap, rcv)
        loss = torch.sum((r - obs) ** 2)
        loss.backward()
        return cc.grad

    once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    once()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    cells = p.NX * p.NY * p.NSTEP
    return {"value": cells / t / 1e9, "unit": "Gcell-steps/s", "ms": t * 1e3,
            "path": "AcousticPropagatorSolver + SimulatedObservation_ + autograd (full history, "
                    "one shot, FP32)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="reduced grid for quick checks")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--interval", type=int, default=0, help="checkpoint interval (0 = auto)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the kernel shape with the plan's autotuner instead of --shape")
    ap.add_argument("--shape", default=DEFAULT_SHAPE,
                    help="kernel shape 'fr,fs,fb,ar,as,ab' (rows per wavefront, shots per "
                         "workgroup, shots per batch of the forward and adjoint kernels); the "
                         "default is the autotuner's pick for this workload, fixed so that the "
                         "profiled HBM traffic refers to the same launches")
    ap.add_argument("--nstep", type=int, default=0,
                    help="override the record length (profiling runs only; not the metric)")
    ap.add_argument("--no-elastic", action="store_true", help="skip the elastic gradient leg")
    ap.add_argument("--propagator-kernel", type=int, default=1, choices=(0, 1),
                    help="PropagatorKernel of the acoustic workload (0: the reference's default "
                         "Core.jl form, phi/psi from the new u; 1: the custom-op form)")
    ap.add_argument("--elastic-only", action="store_true",
                    help="run only the elastic gradient leg (profiling runs; no bench line)")
    ap.add_argument("--elastic-nstep", type=int, default=0,
                    help="override the elastic record length (profiling runs only)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); shots shard across "
              "GPUs with one gradient all-reduce per iteration, which is not implemented here",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as g

    ads = g.load_package()
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    lib = ads.load_library()
    torch.cuda.init()
    if args.elastic_only:
        print(json.dumps({"elastic": elastic_leg(ads, L, lib, args.steps, args.elastic_nstep)}),
              flush=True)
        return
    p, c_true, c0, srcs, rcvs = workload(ads, small=args.small, nstep=args.nstep,
                                         kernel=args.propagator_kernel)
    if args.calib_copy:
        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    K = args.interval or plan.choose_interval(int(torch.cuda.mem_get_info()[0] * 0.80))
    ct = torch.as_tensor(c_true, dtype=torch.float32, device="cuda")
    c0t = torch.as_tensor(c0, dtype=torch.float32, device="cuda")
    obs = plan.forward(ct, srcv, K).clone()
    torch.cuda.synchronize()
    if args.autotune:
        shape = plan.autotune(c0t, srcv, K, reps=3)
    else:
        shape = tuple(int(x) for x in args.shape.split(","))
        plan.set_shape(shape)
    for _ in range(args.warmup):
        plan.gradient(c0t, srcv, obs, K)
    torch.cuda.synchronize()
    L.check(lib.ads_profile_begin())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, grad, _ = plan.gradient(c0t, srcv, obs, K)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    L.check(lib.ads_profile_end())
    ms = t / args.steps * 1e3
    cells = float(np.sum(p.NX * p.NY * plan.nstep))
    value = cells / (ms * 1e-3) / 1e9
    if not math.isfinite(float(loss.item())):
        raise RuntimeError("non-finite misfit")

    prof = {}
    for name, cls in (("ac_fwd_step", 0), ("ac_adj_step", 1)):
        tm, n, b = C.c_double(), C.c_int64(), C.c_double()
        L.check(lib.ads_profile_query(cls, C.byref(tm), C.byref(n), C.byref(b)))
        prof[name] = (tm.value, n.value, b.value)
    dom = max(prof, key=lambda k: prof[k][0])
    tm, n, b = prof[dom]
    achieved = (b / n) / (tm / n * 1e-3) / 1e9 if n else 0.0
    # the committed PMC summary was taken with PropagatorKernel 1
    traffic, traffic_shape = measured_traffic(dom, shape) if p.PropagatorKernel == 1 else (None, None)
    out = {
        "metric": "FWI gradient Gcell-steps/s (fwd+adj), batched",
        "value": round(value, 4), "unit": "Gcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"acoustic {p.NX}x{p.NY}, h=5 m, dt=0.5 ms, {p.NSTEP} steps, "
                                f"{p.NPOINTS_PML}-cell PML, {plan.nshots} shots x {p.NX} "
                                "receivers, FP32 forward + hand-written adjoint, one batched "
                                "call per gradient"),
                   "PropagatorKernel": p.PropagatorKernel, "checkpoint_interval": K,
                   "kernel_shape": {"fwd_rows": shape[0], "fwd_shots": shape[1],
                                    "fwd_batch": shape[2], "adj_rows": shape[3],
                                    "adj_shots": shape[4], "adj_batch": shape[5]} if shape else None,
                   "shots": plan.nshots, "cells": p.NX * p.NY, "nstep": p.NSTEP},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": traffic, "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)
                     if traffic else (f"profiled at shape {traffic_shape}, not this run's"
                                      if traffic_shape else None),
                     "avg_launch_us": round(tm / n * 1e3, 2) if n else None,
                     "launches": n, "alg_bytes_per_launch": round(b / n) if n else None},
        "kernels": {k: {"total_ms": round(v[0], 2), "launches": v[1],
                        "avg_us": round(v[0] / v[1] * 1e3, 2) if v[1] else None,
                        "alg_GBps": round(v[2] / (v[0] * 1e-3) / 1e9, 1) if v[0] else None}
                    for k, v in prof.items()},
        "loss": float(loss.item()),
    }
    tr_true = plan.unpack_traces(obs)[0].cpu().numpy()
    plan._ws = None  # release the batched workspace before the full-history run
    del grad
    torch.cuda.empty_cache()
    if not args.no_elastic:
        out["elastic"] = elastic_leg(ads, L, lib, 1, args.elastic_nstep)
    if not args.no_plain:
        out["plain_single_shot"] = plain_single_shot(ads, p, c0, srcs[0], rcvs[0], tr_true)
        out["plain_single_shot"]["batched_over_plain"] = round(value / out["plain_single_shot"]["value"], 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, c0, srcs[0], min(args.cpu_steps, p.NSTEP))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
