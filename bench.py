"""Benchmark: batched FP32 acoustic FWI gradient (forward + hand-written adjoint) on one MI355X.

Headline workload: 2-D acoustic FWI gradient with checkpointing, 2000 x 2000 interior cells,
h = 5 m, dt = 0.5 ms, 40-cell PML, smooth random velocity 1500-5000 m/s, 64 shots of uneven
record length (uniform in 3000-6000 steps) and uneven receiver spreads (400-2000 receivers), all
shots in one plan: one forward sweep keeping checkpoints (u^a, u^{a-1} in full, phi/psi outside
the damping-free box), then the reversal schedule that rebuilds each segment from its checkpoint
before its adjoint sweep.  One "step" is one complete FWI gradient of all 64 shots (forward,
misfit, rebuilds and adjoint) in one batched call.  Observed data are simulated from the true
model; the gradient is taken at a smoothed starting model.

Prints ONE JSON line.  Besides value/ms_per_step it carries:
  roofline      : the dominant stencil class's algorithmic HBM bytes per level / its average
                  time per level (HIP events on the launch stream around every 7th launch of
                  each class, over the timed region), with the rocprofv3-measured HBM bytes per
                  level;
  cpu_baseline  : the FP64 CPU restatement (oracle/) on the process's threads, with a one-thread
                  run beside it (like the reference's serial CPU op), on a bounded sample of
                  the headline workload;
  acoustic_1000x1000_32shots : the 1000 x 1000, 4000-step, 32-shot batched gradient, with
                  PropagatorKernel 1 and (the reference's default) 0;
  plain_single_shot : the plain drop-in path (AcousticPropagatorSolver + SimulatedObservation_
                  + autograd, full histories, one shot) on the 1000 x 1000 grid, and the batched
                  path's throughput over it;
  elastic       : the 600 x 600, 3000-step, 16-shot elastic gradient (d/d rho, lambda, mu).
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
# Kernel-class timing brackets launches 0, 7, 14, ... of each class over the timed region with
# HIP events (ads_profile_begin_sampled): events on every launch cost 1-5% of the throughput.
PROF_EVERY = 7
_T_START = time.perf_counter()


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line): a long run reports each
    finished gradient, so a watcher can tell it from a hung one."""
    print(f"bench.py [{time.perf_counter() - _T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)
TRAFFIC_FILES = {
    "uneven": os.path.join(ROOT, "profiles", "traffic_acoustic_2000x2000_64shots.txt"),
    "layered": os.path.join(ROOT, "profiles", "traffic_acoustic_1000x1000_32shots.txt"),
}


# Kernel shape of the acoustic plan on the benchmark workload: forward 8 rows per wavefront (the
# streaming two-level rectangle, one shot per wavefront; 16 shots per workgroup for the frame
# items), adjoint 2 rows per wavefront, 16 shots per workgroup, 1 shot per batch.  The fastest
# of the shapes timed back to back on one MI355X (forward 96-98 us and adjoint 165 us per level
# at 1000 x 1000 x 32 shots; 4, 12 or 16 streaming rows within 1-4%; 32 shots per frame group
# the same within noise, 8 and 4 slower); it is also the plan's default shape, and
# ads_acoustic_plan_autotune picks shapes within a few percent of it.
DEFAULT_SHAPE = "8,16,1,2,16,1"


def measured_traffic(kernel, shape, path):
    """Measured HBM bytes over algorithmic bytes of a stencil class, from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain device copy
    of known size).  Lines: '<class> <measured bytes per level> <rows,shots,batch> <algorithmic
    bytes per level of the profiled run>'.  Returns (ratio, profiled shape); ratio is None when
    the summary is missing or was taken at another kernel shape."""
    try:
        with open(path) as f:
            lines = [ln.split() for ln in f if ln.strip() and not ln.startswith("#")]
    except OSError:
        return None, None
    kv = {ln[0]: ln[1:] for ln in lines if len(ln) >= 4}
    if kernel not in kv:
        return None, None
    prof_shape = tuple(int(x) for x in kv[kernel][1].split(","))
    half = None if shape is None else tuple(shape[:3] if kernel == "ac_fwd_step" else shape[3:])
    if half != prof_shape:
        return None, prof_shape
    return float(kv[kernel][0]) / float(kv[kernel][2]), prof_shape


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


def _cpu_gradient(O, p, c0, src, rcv, nsteps, threads):
    """Oracle forward + adjoint of one shot over the first nsteps steps on `threads` threads
    (1: the serial checker itself, scatter-form step VJPs; more: the gather-form sweep)."""
    O.set_num_threads(threads)
    NX, NY = p.NX, p.NY
    Sx, Sy = O.acoustic_pml(NX, NY, p.DELTAX, p.DELTAY, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    srcv = np.asarray(src.srcv)[:nsteps]
    rcvi, rcvj = np.asarray(rcv.rcvi), np.asarray(rcv.rcvj)
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(p, c0, src, rcv, nsteps, what):
    """Oracle (FP64) forward + adjoint of one shot over the first nsteps steps, on all the
    threads this process may use (OMP_NUM_THREADS caps the affinity set) and on one thread,
    like the reference's serial custom op."""
    import oracle as O

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = max(1, min(threads, int(os.environ["OMP_NUM_THREADS"])))
    sample = (f"1 shot of {what}, {p.NX}x{p.NY} grid, first {nsteps} steps, "
              f"{len(rcv.rcvi)} receivers, FP64 forward+adjoint (oracle/ads_oracle.c, "
              "PropagatorKernel 1)")
    v1, t1 = _cpu_gradient(O, p, c0, src, rcv, nsteps, 1)
    out = {"value": v1, "unit": "Gcell-steps/s", "cores": 1, "kind": "port",
           "sample": f"{sample}, one thread, {t1:.1f} s"}
    if threads > 1:
        vm, tm = _cpu_gradient(O, p, c0, src, rcv, nsteps, threads)
        out = {"value": vm, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{sample}, {threads} threads (OpenMP rows, gather-form adjoint), {tm:.1f} s",
               "single_thread": out}
    O.set_num_threads(1)
    out["host_cpu"] = cpu_model()
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


def elastic_leg(ads, L, lib, steps, nstep=0, adjoint_kernels=None, forward_kernels=None):
    """Batched FP32 elastic FWI gradient (d/d rho, lambda, mu) throughput and kernel roofline.
    adjoint_kernels: (streaming, rows, shots) for ads_elastic_plan_set_adjoint_kernels;
    forward_kernels: (fused, rows) for ads_elastic_plan_set_forward_kernels."""
    import ctypes as C
    import torch

    p, true, start, srcs, rcvs = elastic_workload(ads, nstep)
    torch.cuda.empty_cache()  # memory the allocator still caches from earlier legs counts as free
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float32)
    if adjoint_kernels:
        plan.set_adjoint_kernels(*adjoint_kernels)
    if forward_kernels:
        plan.set_forward_kernels(*forward_kernels)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    K, S = plan.choose_schedule(int(torch.cuda.mem_get_info()[0] * 0.90))
    tt = [torch.as_tensor(x, dtype=torch.float32, device="cuda") for x in true]
    st = [torch.as_tensor(x, dtype=torch.float32, device="cuda") for x in start]
    obs = plan.forward(*tt, srcv, K).clone()
    plan.gradient(*st, srcv, obs, K)
    torch.cuda.synchronize()
    L.check(lib.ads_profile_begin_sampled(PROF_EVERY))
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
        if cls == 3 and n.value == 0:  # the fused forward step: both phases in the stress class
            kern["el_fwd_step"] = kern.pop("el_fwd_stress")
            continue
        kern[name] = {"sampled_ms": round(tm.value, 2), "sampled_launches": n.value,
                      "avg_us": round(tm.value / n.value * 1e3, 2) if n.value else None,
                      "alg_GBps": round(b.value / (tm.value * 1e-3) / 1e9, 1) if tm.value else None}
    dom = max(kern, key=lambda k: kern[k]["sampled_ms"])
    ach = kern[dom]["alg_GBps"] or 0.0
    out = {"workload": (f"elastic {p.NX}x{p.NY}, h=10 m, dt=0.8 ms, {p.NSTEP} steps, "
                        f"{p.NPOINTS_PML}-cell CPML, {plan.nshots} explosive shots x {p.NX} "
                        "two-component receivers, FP32 forward + hand-written adjoint for "
                        "(rho, lambda, mu), one batched call per gradient"),
           "value": round(cells / t / 1e9, 4), "unit": "Gcell-steps/s",
           "ms_per_gradient": round(t * 1e3, 2), "checkpoint_interval": K, "snapshots": S,
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBPS,
                        "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBPS, 4),
                        "event_sample_every": PROF_EVERY},
           "kernels": kern, "loss": float(loss.item())}
    # whole gradient: every cell-step once forward, w (13 + 8 f_x + 8 f_y), and once adjoint,
    # w (24 + 8 f_x + 8 f_y) (f_x, f_y: interior fractions in the x / y CPML strips); rebuilt
    # levels are overhead.  Over the gradient's wall time.
    ax, _, _, ay, _, _ = ads.compute_PML_Params_elastic(p)
    f_x = float(np.mean((ax[0] != 0) | (ax[1] != 0)))
    f_y = float(np.mean((ay[0] != 0) | (ay[1] != 0)))
    bpc = 4.0 * (37.0 + 16.0 * (f_x + f_y))
    out["whole_gradient"] = {"alg_bytes_per_cell_step": round(bpc, 2),
                             "alg_GBps": round(cells * bpc / t / 1e9, 1),
                             "frac": round(cells * bpc / t / 1e9 / PEAK_HBM_GBPS, 4)}
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


def acoustic_leg(ads, L, lib, args, p, c_true, c0, srcs, rcvs, nstep, steps, warmup, traffic):
    """Batched FP32 acoustic gradient: one plan over every shot, the fastest checkpoint schedule
    that fits 90% of the free device memory (or --interval/--snapshots), `warmup` untimed and
    `steps` timed gradients.  Returns (result dict, observed traces of shot 0)."""
    import torch

    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    if args.interval:
        plan.set_snapshots(args.snapshots)
        K, S = args.interval, args.snapshots
    else:
        K, S = plan.choose_schedule(int(torch.cuda.mem_get_info()[0] * 0.90))
    ws_bytes = plan.workspace_size(K)
    M = -(-(int(plan.nstep.max()) - 1) // K)
    ct = torch.as_tensor(c_true, dtype=torch.float32, device="cuda")
    c0t = torch.as_tensor(c0, dtype=torch.float32, device="cuda")
    what = f"acoustic {p.NX}x{p.NY}, {len(srcs)} shots, kernel {p.PropagatorKernel}"
    progress(f"{what}: simulating the observed data (K = {K}, {M} segments)")
    obs = plan.forward(ct, srcv, K).clone()
    torch.cuda.synchronize()
    if args.autotune:
        shape = plan.autotune(c0t, srcv, K, reps=3)
    else:
        shape = tuple(int(x) for x in args.shape.split(","))
        plan.set_shape(shape)
    for k in range(warmup):
        plan.gradient(c0t, srcv, obs, K)
        torch.cuda.synchronize()
        progress(f"{what}: warmup gradient {k + 1}/{warmup} done")
    torch.cuda.synchronize()
    L.check(lib.ads_profile_begin_sampled(PROF_EVERY))
    t0 = time.perf_counter()
    for k in range(steps):
        loss, grad, _ = plan.gradient(c0t, srcv, obs, K)
        torch.cuda.synchronize()  # (one wait per gradient of seconds: no measurable cost)
        progress(f"{what}: timed gradient {k + 1}/{steps} done")
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    L.check(lib.ads_profile_end())
    ms = t / steps * 1e3
    cells = float(np.sum(p.NX * p.NY * np.asarray(plan.nstep, dtype=np.float64)))
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
    ratio, prof_shape = measured_traffic(dom, shape, traffic) if (traffic and p.PropagatorKernel == 1) \
        else (None, None)
    # HBM bytes per level at this run's mix of active shots: the profiled measured/algorithmic
    # ratio times this run's algorithmic bytes per level
    hbm = round(ratio * b / n) if (ratio and n) else None
    out = {
        "value": round(cells / (ms * 1e-3) / 1e9, 4), "unit": "Gcell-steps/s",
        "ms_per_gradient": round(ms, 3), "cell_steps": cells,
        "schedule": {"interval": K, "segments": M, "snapshots": S,
                     "kind": "one-level" if S >= M - 1 else "binomial",
                     "workspace_GB": round(ws_bytes / 1e9, 2)},
        "kernel_shape": {"fwd_rows": shape[0], "fwd_shots": shape[1], "fwd_batch": shape[2],
                         "adj_rows": shape[3], "adj_shots": shape[4], "adj_batch": shape[5]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": hbm,
                     "traffic_source": os.path.relpath(traffic, ROOT) if hbm else
                     (f"profiled at shape {prof_shape}, not this run's" if prof_shape else None),
                     "avg_level_us": round(tm / n * 1e3, 2) if n else None,
                     "sampled_levels": n, "event_sample_every": PROF_EVERY,
                     "alg_bytes_per_level": round(b / n) if n else None},
        "kernels": {k: {"sampled_ms": round(v[0], 2), "sampled_levels": v[1],
                        "avg_us": round(v[0] / v[1] * 1e3, 2) if v[1] else None,
                        "alg_GBps": round(v[2] / (v[0] * 1e-3) / 1e9, 1) if v[0] else None}
                    for k, v in prof.items()},
        "loss": float(loss.item()),
    }
    # algorithmic bytes of the whole gradient: every cell-step once forward, w (4 + 4 f), and
    # once adjoint, w (7 + 4 f) (f: interior fraction in the PML strips); rebuilt levels are
    # overhead.  Over the gradient's wall time.
    sx = np.zeros(p.NX + 2)
    ty = np.zeros(p.NY + 2)
    L.check(lib.ads_acoustic_pml_profiles(C.byref(p.to_c()), sx.ctypes.data, ty.ctypes.data))
    f = float(np.mean((sx[1:-1, None] != 0) | (ty[None, 1:-1] != 0)))
    alg = cells * 4.0 * (11.0 + 8.0 * f)
    out["whole_gradient"] = {"alg_bytes_per_cell_step": round(4.0 * (11.0 + 8.0 * f), 2),
                             "alg_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
                             "frac": round(alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)}
    obs0 = plan.unpack_traces(obs)[0].cpu().numpy()
    plan._ws = None
    del grad, obs
    torch.cuda.empty_cache()
    return out, obs0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline leg only (no 1000x1000, plain, elastic or CPU legs)")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--interval", type=int, default=0,
                    help="checkpoint interval (0 = the schedule chooser's)")
    ap.add_argument("--snapshots", type=int, default=-1,
                    help="snapshot slots with --interval (-1: one per segment start)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the kernel shape with the plan's autotuner instead of --shape")
    ap.add_argument("--shape", default=DEFAULT_SHAPE,
                    help="kernel shape 'fr,fs,fb,ar,as,ab' (rows per wavefront, shots per "
                         "workgroup, shots per batch of the forward and adjoint kernels), fixed "
                         "so that the profiled HBM traffic refers to the same launches")
    ap.add_argument("--max-nstep", type=int, default=0,
                    help="clip every record to this many steps (profiling runs only; not the "
                         "metric)")
    ap.add_argument("--workload", choices=("uneven", "layered"), default="uneven",
                    help="headline workload: 2000x2000 x 64 uneven shots, or the 1000x1000 x 32 "
                         "layered one (profiling runs)")
    ap.add_argument("--no-elastic", action="store_true", help="skip the elastic gradient leg")
    ap.add_argument("--propagator-kernel", type=int, default=1, choices=(0, 1),
                    help="PropagatorKernel of the headline workload (0: the reference's default "
                         "Core.jl form, phi/psi from the new u; 1: the custom-op form)")
    ap.add_argument("--elastic-only", action="store_true",
                    help="run only the elastic gradient leg (profiling runs; no bench line)")
    ap.add_argument("--elastic-nstep", type=int, default=0,
                    help="override the elastic record length (profiling runs only)")
    ap.add_argument("--elastic-adjoint", default="",
                    help="elastic adjoint kernels 'streaming,rows,shots' (default: the plan's, "
                         "2,8,4)")
    ap.add_argument("--elastic-forward", default="",
                    help="elastic forward kernels 'fused,rows' (default: the plan's, 1,0)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); shots shard across "
              "GPUs with one gradient all-reduce per iteration, which is not implemented here",
              file=sys.stderr)
        sys.exit(2)

    progress("loading torch and the HIP library")
    import torch
    import __graft_entry__ as g

    ads = g.load_package()
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    lib = ads.load_library()
    torch.cuda.init()
    if args.calib_copy:
        n = args.calib_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.elastic_only:
        ek = tuple(int(x) for x in args.elastic_adjoint.split(",")) if args.elastic_adjoint else None
        fk = tuple(int(x) for x in args.elastic_forward.split(",")) if args.elastic_forward else None
        print(json.dumps({"elastic": elastic_leg(ads, L, lib, args.steps, args.elastic_nstep, ek, fk)}),
              flush=True)
        return
    if args.workload == "uneven":
        p, c_true, c0, srcs, rcvs, nst = workload_uneven(ads)
        p.PropagatorKernel = args.propagator_kernel
        what = (f"acoustic {p.NX}x{p.NY}, h=5 m, dt=0.5 ms, 40-cell PML, smooth random model "
                f"1500-5000 m/s, {len(srcs)} shots of uneven record length "
                f"({int(nst.min())}-{int(nst.max())} steps) and uneven receiver spreads "
                f"({min(len(r.rcvi) for r in rcvs)}-{max(len(r.rcvi) for r in rcvs)} receivers), "
                "FP32 forward + hand-written adjoint, all shots in one checkpointed batched call")
    else:
        p, c_true, c0, srcs, rcvs = workload(ads, kernel=args.propagator_kernel)
        nst = np.full(len(srcs), p.NSTEP)
        what = (f"acoustic {p.NX}x{p.NY}, h=5 m, dt=0.5 ms, {p.NSTEP} steps, 40-cell PML, "
                f"{len(srcs)} shots x {p.NX} receivers, FP32 forward + hand-written adjoint, "
                "one batched call per gradient")
    if args.max_nstep:
        nst = np.minimum(nst, args.max_nstep)
        srcs = [ads.AcousticSource(s.srci, s.srcj, np.asarray(s.srcv)[:int(k)])
                for s, k in zip(srcs, nst)]
        p.NSTEP = int(nst.max())
    head, _ = acoustic_leg(ads, L, lib, args, p, c_true, c0, srcs, rcvs, nst, args.steps,
                           args.warmup, TRAFFIC_FILES[args.workload])
    out = {
        "metric": "FWI gradient Gcell-steps/s (fwd+adj), batched",
        "value": head.pop("value"), "unit": head.pop("unit"), "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head.pop("ms_per_gradient"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": what, "PropagatorKernel": p.PropagatorKernel, "shots": len(srcs),
                   "cells": p.NX * p.NY, "nstep_min": int(nst.min()), "nstep_max": int(nst.max()),
                   "cell_steps": head.pop("cell_steps"), "schedule": head.pop("schedule"),
                   "kernel_shape": head.pop("kernel_shape")},
        "roofline": head.pop("roofline"),
    }
    out.update(head)
    if not args.no_extra and args.workload == "uneven" and not args.max_nstep:
        layered = {}
        pl, ct_l, c0_l, srcs_l, rcvs_l = workload(ads, kernel=1)
        nst_l = np.full(len(srcs_l), pl.NSTEP)
        r1, obs0 = acoustic_leg(ads, L, lib, args, pl, ct_l, c0_l, srcs_l, rcvs_l, nst_l, 2, 1,
                                TRAFFIC_FILES["layered"])
        layered["PropagatorKernel_1"] = r1
        pl0, ct0, c00, srcs0, rcvs0 = workload(ads, kernel=0)
        r0, _ = acoustic_leg(ads, L, lib, args, pl0, ct0, c00, srcs0, rcvs0, nst_l, 2, 1, None)
        layered["PropagatorKernel_0"] = r0
        layered["workload"] = (f"acoustic {pl.NX}x{pl.NY}, h=5 m, dt=0.5 ms, {pl.NSTEP} steps, "
                               f"{pl.NPOINTS_PML}-cell PML, random layered model, {len(srcs_l)} "
                               f"shots x {pl.NX} receivers, FP32")
        out["acoustic_1000x1000_32shots"] = layered
        if not args.no_plain:
            progress("plain single-shot path")
            pp = plain_single_shot(ads, pl, c0_l, srcs_l[0], rcvs_l[0], obs0)
            pp["batched_over_plain"] = round(r1["value"] / pp["value"], 2)
            pp["batched_path"] = "acoustic_1000x1000_32shots.PropagatorKernel_1"
            out["plain_single_shot"] = pp
        if not args.no_elastic:
            progress("elastic 600x600, 16 shots")
            out["elastic"] = elastic_leg(ads, L, lib, 1, args.elastic_nstep)
        if not args.no_cpu:
            progress("CPU baseline")
            out["cpu_baseline"] = cpu_baseline(p, c0, srcs[0], rcvs[0],
                                               min(args.cpu_steps, int(nst.min())),
                                               "the headline workload")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
