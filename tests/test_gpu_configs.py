"""The plain FP64 acoustic configuration run end to end against the oracle.

200 x 200 grid, h = 10 m, dt = 1 ms, 1000 steps, 20-cell PML, vp_ref 1000 (the parameter
default), PropagatorKernel 0 (the default).  Velocity: three equal horizontal layers at
1500/2500/3500 m/s plus a +10% Gaussian anomaly (sigma = 15 cells) at the centre.  One 10 Hz
Ricker (frequency-domain form, amplitude 1) at depth 5 cells, mid-x; 200 receivers on the first
interior row.  Observed data come from the true model; the gradient is taken at a sigma = 20
cell Gaussian smoothing of it.  Bar: traces 1e-12, misfit 1e-10, gradient 1e-10 relative (FP64).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def layered_model(NX, NY):
    c = np.empty((NX + 2, NY + 2))
    layer = np.minimum(np.arange(NY + 2) * 3 // (NY + 2), 2)
    c[:] = np.array([1500.0, 2500.0, 3500.0])[layer][None, :]
    ii, jj = np.meshgrid(np.arange(NX + 2), np.arange(NY + 2), indexing="ij")
    ci, cj = (NX + 1) / 2.0, (NY + 1) / 2.0
    return c * (1.0 + 0.10 * np.exp(-((ii - ci) ** 2 + (jj - cj) ** 2) / (2 * 15.0 ** 2)))


def test_plain_fp64_configuration(orc, ads, cuda):
    from scipy.ndimage import gaussian_filter

    NX = NY = 200
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=1000, DELTAX=10.0, DELTAY=10.0,
                                     DELTAT=1e-3, NPOINTS_PML=20)
    assert p.vp_ref == 1000.0 and p.PropagatorKernel == 0
    c_true = layered_model(NX, NY)
    c0 = gaussian_filter(c_true, 20.0, mode="reflect")
    ads.check_cfl(p, float(c_true.max()))
    w = ads.ricker_freq(10.0, p.DELTAT, p.NSTEP, 1.0)
    src = ads.AcousticSource([NX // 2 + 1], [5], w.reshape(-1, 1))
    rcv = ads.AcousticReceiver(np.arange(2, NX + 2), np.full(NX, 2))
    Sx, Sy = orc.acoustic_pml(NX, NY, p.DELTAX, p.DELTAY, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    u, _, _ = orc.acoustic_forward(0, NX, NY, p.NSTEP, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy,
                                   c_true ** 2, src.srci, src.srcj, src.srcv)
    obs = orc.acoustic_receive(u, rcv.rcvi, rcv.rcvj, NY)
    del u
    L_ref, g_ref, tr_ref = orc.acoustic_misfit_gradient(0, NX, NY, p.NSTEP, p.DELTAT, p.DELTAX,
                                                        p.DELTAY, Sx, Sy, c0, src.srci, src.srcj,
                                                        src.srcv, rcv.rcvi, rcv.rcvj, obs)
    plan = ads.AcousticPlan(p, [src], [rcv], dtype=torch.float64)
    srcv = plan.pack_srcv([src.srcv])
    K = plan.choose_interval()
    ct = torch.as_tensor(c0, dtype=torch.float64, device="cuda")
    loss, g, tr = plan.gradient(ct, srcv, plan.pack_traces([obs]), K, want_traces=True)
    assert rel(plan.unpack_traces(tr)[0].cpu(), tr_ref) < 1e-12
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    assert rel(g.cpu(), g_ref) < 1e-10, rel(g.cpu(), g_ref)
    # the same through the plain drop-in (full histories + torch autograd)
    cg = ct.clone().requires_grad_(True)
    ap = ads.AcousticPropagatorSolver(p, src, cg)
    rr = ads.SimulatedObservation_(ap, rcv)
    torch.sum((rr - torch.as_tensor(obs, device="cuda")) ** 2).backward()
    assert rel(cg.grad.cpu(), g_ref) < 1e-10


@pytest.mark.timeout(900)
def test_large_uneven_batch_configuration(ads, cuda):
    """The 2000 x 2000, 64-shot configuration with uneven records (NSTEP_s ~ U{3000..6000},
    spreads of 400..2000 receivers centred on each source), FP32, dt = 0.5 ms (chosen: CFL
    0.5 at 5000 m/s), 40-cell PML, smooth random model 1500..5000 m/s (seed 4).  Its
    checkpointed workspace does not fit one plan (about 300 GB for 64 shots at the best
    interval), so the shots go through plans of 16 and of 8 (acoustic_gradient_in_batches).
    The oracle is out of reach at this size; checked instead through size-independent
    properties: (1) the two batchings (different batch sizes, shot groupings and checkpoint
    intervals) give the same misfit and gradient to FP32 reassociation; (2) a central finite
    difference of the misfit along a perturbation of 0.5% of the model in the direction of the
    (smoothed) gradient matches <gradient, perturbation> within 2%."""
    import time
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(4)
    NX = NY = 2000
    nshots = 64
    dt, h = 5e-4, 5.0
    nst = rng.integers(3000, 6001, nshots)
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
    T = lambda x: torch.as_tensor(x, dtype=torch.float32, device="cuda")
    # observed data from the true model, 16 shots per plan
    obs = ads.acoustic_traces_in_batches(p, srcs, rcvs, T(c_true), 16, nstep=nst)
    cells = NX * NY * int(nst.sum())
    res = {}
    for ms in (16, 8):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res[ms] = ads.acoustic_gradient_in_batches(p, srcs, rcvs, obs, T(c0), ms, nstep=nst)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        print(f"{ms} shots per plan: {el:.2f} s, {cells / el / 1e9:.1f} Gcell-steps/s", flush=True)
    (L16, g16), (L8, g8) = res[16], res[8]
    assert L16.item() == pytest.approx(L8.item(), rel=1e-6)
    assert rel(g8.cpu(), g16.cpu()) < 1e-5
    # directional derivative along the (smoothed) gradient itself, scaled to a perturbation of
    # at most 0.5% of the model: the largest first-order signal against FP32 rounding
    gs = gaussian_filter(g16.double().cpu().numpy(), 5.0)
    d = gs * (0.005 * c0.min() / np.abs(gs).max())
    def misfit(cm):
        tr = ads.acoustic_traces_in_batches(p, srcs, rcvs, T(cm), 16, nstep=nst)
        return sum(float(torch.sum((a.double() - b.double()) ** 2)) for a, b in zip(tr, obs))

    fd = (misfit(c0 + d) - misfit(c0 - d)) / 2.0
    an = float(torch.sum(g16.double() * torch.as_tensor(d, device="cuda")).item())
    print(f"directional derivative: finite difference {fd:.6e}, gradient {an:.6e}, "
          f"misfit {L16.item():.6e}", flush=True)
    assert abs(fd - an) <= 0.02 * abs(an)
