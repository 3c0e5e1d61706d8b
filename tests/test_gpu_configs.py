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
    0.5 at 5000 m/s), 40-cell PML, smooth random model 1500..5000 m/s -- bench.py's headline
    workload -- with all 64 shots in ONE plan under the checkpoint schedule chosen for 90% of
    the free device memory.  The oracle is out of reach at this size; checked instead through
    size-independent properties: (1) plans of 16 shots (other shot groupings and schedules)
    give the same misfit and gradient to FP32 reassociation; (2) the binomial schedule with 4
    snapshot slots gives the chosen schedule's gradient bit for bit; (3) a central finite
    difference of the misfit along a perturbation of 0.2% of the model in the direction of the
    smoothed gradient, restricted to the imaging region below the acquisition PML band, matches
    <gradient, perturbation> within 2%."""
    import sys
    import time
    from scipy.ndimage import gaussian_filter
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    p, c_true, c0, srcs, rcvs, nst = bench.workload_uneven(ads)
    NX, NY = p.NX, p.NY
    T = lambda x: torch.as_tensor(x, dtype=torch.float32, device="cuda")
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nst)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    K, S = plan.choose_schedule(int(torch.cuda.mem_get_info()[0] * 0.9))
    M = -(-(p.NSTEP - 1) // K)
    print(f"one plan of 64 shots: K {K}, S {S} (M {M}), workspace "
          f"{plan.workspace_size(K) / 1e9:.1f} GB", flush=True)
    obs = plan.forward(T(c_true), srcv, K).clone()
    cells = NX * NY * int(nst.sum())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L1, g1, _ = plan.gradient(T(c0), srcv, obs, K)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(f"one plan: {el:.2f} s, {cells / el / 1e9:.1f} Gcell-steps/s, misfit {L1.item():.6e}",
          flush=True)
    L1, g1 = L1.item(), g1.cpu().numpy()
    # (2) binomial schedule with 4 slots: the same arithmetic, more rebuilds
    plan.set_snapshots(4)
    L4, g4, _ = plan.gradient(T(c0), srcv, obs, K)
    assert L4.item() == L1
    np.testing.assert_array_equal(g4.cpu().numpy(), g1)
    plan.set_snapshots(S)
    # (3) directional derivative in the imaging region
    m = np.zeros_like(c0)
    w = p.NPOINTS_PML + 2
    m[w:NX + 2 - w, p.NPOINTS_PML + 20:NY + 2 - w] = 1.0
    m = gaussian_filter(m, 10.0) * m
    gs = gaussian_filter(g1 * m, 5.0) * m
    d = gs * (0.002 * c0.min() / np.abs(gs).max())

    def misfit(cm):
        tr = plan.forward(T(cm), srcv, K)
        return float(torch.sum((tr.double() - obs.double()) ** 2).item())

    fd = (misfit(c0 + d) - misfit(c0 - d)) / 2.0
    an = float(np.sum(g1.astype(np.float64) * d))
    obs_per = [t.clone() for t in plan.unpack_traces(obs)]
    plan._ws = None
    del obs
    torch.cuda.empty_cache()
    print(f"directional derivative: finite difference {fd:.6e}, gradient {an:.6e}", flush=True)
    # (1) plans of 16 shots
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L16, g16 = ads.acoustic_gradient_in_batches(p, srcs, rcvs, obs_per, T(c0), 16, nstep=nst)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(f"16 shots per plan: {el:.2f} s, {cells / el / 1e9:.1f} Gcell-steps/s", flush=True)
    assert L16.item() == pytest.approx(L1, rel=1e-6)
    assert rel(g16.cpu(), g1) < 1e-5
    assert abs(fd - an) <= 0.02 * abs(an)
