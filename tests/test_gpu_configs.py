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
