"""GPU parity tests of the elastic HIP path against the FP64 oracle.

Tolerances: FP64 <= 1e-10 relative L2 (traces, histories, gradients); FP32 traces <= 1e-4 and
gradients <= 1e-3 relative L2 against the FP64 oracle.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def T(x, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")


def make_case(orc, ads, NX=34, NY=28, NSTEP=60, npml=6, nshots=2, seed=0, uneven=False, dt=1e-3):
    rng = np.random.default_rng(seed)
    p = ads.ElasticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=10.0, DELTAY=10.0, DELTAT=dt,
                                    NPOINTS_PML=npml, vp_ref=2500.0, f0=8.0)
    vp = 2300.0 + 400.0 * rng.random((NX + 2, NY + 2))
    vs = vp / 1.73
    rho = 310 * vp ** 0.25
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    nstep = [NSTEP - (5 * s if uneven else 0) for s in range(nshots)]
    srcs, rcvs = [], []
    for s in range(nshots):
        w = orc.ricker(nstep[s], 5.0, 10.0 + 2 * s, 1e8)
        i0 = 6 + 7 * s
        if s % 2 == 0:  # explosive (sxx + syy), plus a vy force on the ring
            srcs.append(ads.ElasticSource([i0, i0, 1], [5, 5, 4], [2, 3, 1], np.stack([w, w, 0.2 * w], 1)))
        else:
            srcs.append(ads.ElasticSource([i0, i0], [6, 6], [0, 4], np.stack([w, -w], 1)))
        ri = np.concatenate([np.arange(2, NX + 1, 2), np.arange(3, NX + 1, 3), [NX // 2, NX // 2]])
        rj = np.full(len(ri), 3 + s % 2)
        rt = np.concatenate([np.zeros(len(np.arange(2, NX + 1, 2))), np.full(len(np.arange(3, NX + 1, 3)), 1),
                             [2, 2]]).astype(np.int64)
        rcvs.append(ads.ElasticReceiver(ri, rj, rt))
    return p, (rho, lam, mu), srcs, rcvs, nstep


def oracle(orc, p, models, src, rcv, nstep, scale=None, obs=None):
    rho, lam, mu = models
    pml = orc.elastic_pml(p.NX, p.NY, p.DELTAX, p.DELTAY, p.DELTAT, p.NPOINTS_PML, p.NPOWER, p.K_MAX_PML,
                          p.ALPHA_MAX_PML, p.Rcoef, p.vp_ref)
    if scale is not None:
        rho, lam, mu = rho * scale[0], lam * scale[1], mu * scale[2]
    if obs is None:
        hist, _ = orc.elastic_forward(p.NX, p.NY, nstep, p.DELTAX, p.DELTAY, p.DELTAT, rho, lam, mu,
                                      pml, src.srci, src.srcj, src.srctype, src.srcv, keep_mem=False)
        return orc.elastic_receive(hist, rcv.rcvi, rcv.rcvj, rcv.rcvtype, p.NX, p.NY), hist
    return orc.elastic_misfit_gradient(p.NX, p.NY, nstep, p.DELTAX, p.DELTAY, p.DELTAT, rho, lam, mu,
                                       pml, src.srci, src.srcj, src.srctype, src.srcv, rcv.rcvi,
                                       rcv.rcvj, rcv.rcvtype, obs)


@pytest.mark.parametrize("K", [9, 1000])
def test_elastic_plan_forward_and_gradient_f64(orc, ads, cuda, K):
    p, models, srcs, rcvs, nstep = make_case(orc, ads, nshots=3, uneven=True)
    obs, L_ref, g_ref = [], 0.0, [0.0, 0.0, 0.0]
    for s in range(3):
        o, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.02, 0.97, 1.04))
        L, g, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        L_ref += L
        g_ref = [a + b for a, b in zip(g_ref, g)]
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m) for m in models]
    tr = plan.unpack_traces(plan.forward(*mt, srcv, K))
    for s in range(3):
        ref, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s])
        assert rel(tr[s].cpu(), ref) < 1e-12, (s, rel(tr[s].cpu(), ref))
    loss, g, _ = plan.gradient(*mt, srcv, plan.pack_traces(obs), K)
    torch.cuda.synchronize()
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    for a, b, name in zip(g, g_ref, ("rho", "lambda", "mu")):
        assert rel(a.cpu(), b) < 1e-10, (name, rel(a.cpu(), b))


def test_elastic_plan_f32(orc, ads, cuda):
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NX=50, NY=44, NSTEP=200, npml=8, nshots=2,
                                             seed=3, dt=8e-4)
    obs, trs, g_ref = [], [], [0.0, 0.0, 0.0]
    for s in range(2):
        o, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.01, 0.98, 1.02))
        L, g, tr = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        trs.append(tr)
        g_ref = [a + b for a, b in zip(g_ref, g)]
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float32)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    loss, g, tr = plan.gradient(*[T(m, torch.float32) for m in models], srcv, plan.pack_traces(obs),
                                40, want_traces=True)
    per = plan.unpack_traces(tr)
    for s in range(2):
        assert rel(per[s].cpu(), trs[s]) < 1e-4, rel(per[s].cpu(), trs[s])
    for a, b, name in zip(g, g_ref, ("rho", "lambda", "mu")):
        assert rel(a.cpu(), b) < 1e-3, (name, rel(a.cpu(), b))


def test_elastic_history_dropin_and_autograd(orc, ads, cuda):
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NX=26, NY=22, NSTEP=40, nshots=1, seed=4)
    src, rcv = srcs[0], rcvs[0]
    _, hist_ref = oracle(orc, p, models, src, rcv, p.NSTEP)
    mt = [T(m).requires_grad_(True) for m in models]
    ep = ads.ElasticPropagatorSolver(p, src, *mt)
    for k, f in enumerate((ep.vx, ep.vy, ep.sigmaxx, ep.sigmayy, ep.sigmaxy)):
        assert rel(f.detach().cpu(), hist_ref[k]) < 1e-12, k
    obs, _ = oracle(orc, p, models, src, rcv, p.NSTEP, scale=(1.03, 1.0, 0.97))
    rr = ads.SimulatedObservation_elastic(ep, rcv)
    assert tuple(rr.shape) == (len(rcv.rcvi), p.NSTEP + 1)
    loss = torch.sum((rr - T(obs)) ** 2)
    loss.backward()
    L_ref, g_ref, _ = oracle(orc, p, models, src, rcv, p.NSTEP, obs=obs)
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    for t, b, name in zip(mt, g_ref, ("rho", "lambda", "mu")):
        assert rel(t.grad.cpu(), b) < 1e-10, (name, rel(t.grad.cpu(), b))
