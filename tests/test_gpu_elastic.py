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
                          p.ALPHA_MAX_PML, p.Rcoef, p.vp_ref,
                          use=(p.USE_PML_XMIN, p.USE_PML_XMAX, p.USE_PML_YMIN, p.USE_PML_YMAX))
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


# interior rectangle / CPML frame splits: partial CPML sides, odd sizes, no CPML (the rectangle
# is the whole update range), and a CPML that leaves no interior rectangle
EL_SPLITS = [
    dict(NX=29, NY=33, npml=5, use=(False, True, True, False)),
    dict(NX=26, NY=21, npml=4, use=(True, True, True, True)),
    dict(NX=17, NY=23, npml=4, use=(False, False, False, False)),
    dict(NX=14, NY=14, npml=7, use=(True, True, True, True)),
]


@pytest.mark.parametrize("split", range(len(EL_SPLITS)))
def test_elastic_grid_splits_f64(orc, ads, cuda, split):
    cfg = EL_SPLITS[split]
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NX=cfg["NX"], NY=cfg["NY"], NSTEP=40,
                                             npml=cfg["npml"], nshots=2, seed=30 + split,
                                             uneven=True)
    p.USE_PML_XMIN, p.USE_PML_XMAX, p.USE_PML_YMIN, p.USE_PML_YMAX = cfg["use"]
    obs, L_ref, g_ref = [], 0.0, [0.0, 0.0, 0.0]
    for s in range(2):
        o, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.02, 0.97, 1.04))
        L, g, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        L_ref += L
        g_ref = [a + b for a, b in zip(g_ref, g)]
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m) for m in models]
    loss, g, tr = plan.gradient(*mt, srcv, plan.pack_traces(obs), 7, want_traces=True)
    per = plan.unpack_traces(tr)
    for s in range(2):
        ref, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s])
        assert rel(per[s].cpu(), ref) < 1e-12, (s, rel(per[s].cpu(), ref))
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    for a, b, name in zip(g, g_ref, ("rho", "lambda", "mu")):
        assert rel(a.cpu(), b) < 1e-10, (name, rel(a.cpu(), b))
    # the full-history drop-in (dense layout) on the same split
    ep = ads.ElasticPropagatorSolver(p, srcs[0], *mt)
    _, hist_ref = oracle(orc, p, models, srcs[0], rcvs[0], p.NSTEP)
    for k, f in enumerate((ep.vx, ep.vy, ep.sigmaxx, ep.sigmayy, ep.sigmaxy)):
        assert rel(f.detach().cpu(), hist_ref[k]) < 1e-12, k


# ------------------------------------------------------------------------------------------
# one MPIElastic block (MPIElastic.jl), stepped with zero halos
# ------------------------------------------------------------------------------------------

def _mpi_case(ads, orc, n=16, NSTEP=40, seed=41):
    """Global 2n x 2n grid split into 2 x 2 blocks; rank 1 steps block (II, JJ) = (1, 2).
    DELTAX != DELTAY exercises the reference's y origin (JJ-1) n DELTAX."""
    p = ads.MPIElasticPropagatorParams(NX=2 * n, NY=2 * n, n=n, NSTEP=NSTEP, DELTAX=10.0,
                                       DELTAY=8.0, DELTAT=8e-4, NPOINTS_PML=5, vp_ref=2500.0,
                                       f0=8.0)
    ads.compute_PML_Params_mpi(p, rank=1, size=4)
    assert (p.II, p.JJ) == (1, 2)
    rng = np.random.default_rng(seed)
    vp = 2300.0 + 400.0 * rng.random((n + 4, n + 4))
    vs = vp / 1.73
    rho = 310 * vp ** 0.25
    models = (rho, rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2)
    w = orc.ricker(NSTEP, 5.0, 10.0, 1e8)
    # global indices; the first two lie in block (1, 2), the third in another block
    src = ads.MPIElasticSource(p, [6, 6, 30], [n + 7, n + 7, 3], [2, 3, 0],
                               np.stack([w, w, w], 1))
    assert list(src.srci) == [8, 8] and list(src.srcj) == [9, 9]
    gi = np.arange(2, n, 2)
    rcv = ads.MPIElasticReceiver(p, np.concatenate([gi, gi]), np.full(2 * len(gi), n + 3),
                                 np.concatenate([np.zeros(len(gi)), np.ones(len(gi))]).astype(np.int64))
    assert np.all(rcv.rcvj == 3)
    # the oracle's block profiles: this block's origin on the global axes
    ax, bx, kx = orc.elastic_pml_axis(n, p.DELTAX, 0.0, p.NX, p.NPOINTS_PML, p.NPOWER,
                                      p.K_MAX_PML, p.ALPHA_MAX_PML, p.Rcoef, p.vp_ref, p.DELTAT)
    ay, by, ky = orc.elastic_pml_axis(n, p.DELTAY, (p.JJ - 1) * n * p.DELTAX, p.NY,
                                      p.NPOINTS_PML, p.NPOWER, p.K_MAX_PML, p.ALPHA_MAX_PML,
                                      p.Rcoef, p.vp_ref, p.DELTAT)
    pml = (ax, bx, kx, ay, by, ky)
    return p, models, src, rcv, pml


def test_mpi_elastic_block_plan_and_dropin_f64(orc, ads, cuda):
    n = 16
    p, models, src, rcv, pml = _mpi_case(ads, orc, n=n)
    NX = n + 2   # the oracle's padded-interior size of a block
    rr_i, rr_j = rcv.rcvi + 2, rcv.rcvj + 2  # padded 1-based for the oracle
    scaled = (models[0] * 1.02, models[1] * 0.97, models[2] * 1.04)
    hist_o, _ = orc.elastic_forward(NX, NX, p.NSTEP, p.DELTAX, p.DELTAY, p.DELTAT, *scaled, pml,
                                    src.srci, src.srcj, src.srctype, src.srcv, keep_mem=False,
                                    sub=True)
    obs = orc.elastic_receive(hist_o, rr_i, rr_j, rcv.rcvtype, NX, NX)
    L_ref, g_ref, tr_ref = orc.elastic_misfit_gradient(NX, NX, p.NSTEP, p.DELTAX, p.DELTAY,
                                                       p.DELTAT, *models, pml, src.srci, src.srcj,
                                                       src.srctype, src.srcv, rr_i, rr_j,
                                                       rcv.rcvtype, obs, sub=True)
    # batched plan
    plan = ads.ElasticPlan(p, [ads.ElasticSource(src.srci, src.srcj, src.srctype, src.srcv)],
                           [ads.ElasticReceiver(rcv.rcvi, rcv.rcvj, rcv.rcvtype)],
                           dtype=torch.float64)
    srcv = plan.pack_srcv([src.srcv])
    mt = [T(m) for m in models]
    loss, g, tr = plan.gradient(*mt, srcv, plan.pack_traces([obs]), 9, want_traces=True)
    assert rel(plan.unpack_traces(tr)[0].cpu(), tr_ref) < 1e-12
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    for a, b, name in zip(g, g_ref, ("rho", "lambda", "mu")):
        assert rel(a.cpu(), b) < 1e-10, (name, rel(a.cpu(), b))
    # the drop-in: block histories (NSTEP+1, n, n), MPISimulatedObservation_, autograd
    mg = [T(m).requires_grad_(True) for m in models]
    ep = ads.MPIElasticPropagatorSolver(p, src, *mg)
    assert tuple(ep.vx.shape) == (p.NSTEP + 1, n, n)
    h_ref, _ = orc.elastic_forward(NX, NX, p.NSTEP, p.DELTAX, p.DELTAY, p.DELTAT, *models, pml,
                                   src.srci, src.srcj, src.srctype, src.srcv, keep_mem=False,
                                   sub=True)
    for k, f in enumerate((ep.vx, ep.vy, ep.sigmaxx, ep.sigmayy, ep.sigmaxy)):
        assert rel(f.detach().cpu(), h_ref[k][:, 2:n + 2, 2:n + 2]) < 1e-12, k
    out = ads.MPISimulatedObservation_(ep, rcv)
    assert tuple(out.shape) == (len(rcv.rcvi), p.NSTEP + 1)
    torch.sum((out - T(obs)) ** 2).backward()
    for t, b, name in zip(mg, g_ref, ("rho", "lambda", "mu")):
        assert rel(t.grad.cpu(), b) < 1e-10, (name, rel(t.grad.cpu(), b))


@pytest.mark.parametrize("kernels", [(2, 8, 4), (2, 4, 8), (2, 3, 2), (2, 32, 1), (2, 16, 4),
                                     (1, 8, 4), (1, 4, 8), (1, 3, 2), (1, 32, 1), (0, 16, 4)])
def test_elastic_adjoint_rectangle_kernels_f64(orc, ads, cuda, kernels):
    """Every adjoint kernel variant gives the oracle's gradient, and a second gradient on the
    same plan is bit-identical to the first (no order-dependent sums): the full-grid streaming
    kernels (2), the streaming rectangle kernels with frame_b in the same launch (4 shots per
    group) or launched apart (8, 2, 1 shots per group; 5 uneven shots, so the last group is
    partial and, near the end, fewer shots are active than a group holds), chunk heights that
    do not divide the grid or the rectangle, and the LDS-tiled kernels (0).  (2, 16, 4) in FP64
    needs more LDS than the streaming kernels may take with 4 shots: the plan then takes the
    tiled and, with fewer active shots, the streaming rectangle kernels, never the full-grid
    ones, whose memory-variable cotangents are kept apart."""
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NX=41, NY=37, NSTEP=50, nshots=5,
                                             uneven=True, seed=3)
    obs, g_ref = [], [0.0, 0.0, 0.0]
    for s in range(5):
        o, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.02, 0.97, 1.04))
        _, g, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        g_ref = [a + b for a, b in zip(g_ref, g)]
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    plan.set_adjoint_kernels(*kernels)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m) for m in models]
    L, g, _ = plan.gradient(*mt, srcv, plan.pack_traces(obs), 11)
    errs = [rel(a.cpu(), b) for a, b in zip(g, g_ref)]
    assert max(errs) < 1e-10, (kernels, errs)
    L2, g2, _ = plan.gradient(*mt, srcv, plan.pack_traces(obs), 11)
    assert L2.item() == L.item()
    for a, b in zip(g2, g):
        assert torch.equal(a, b), kernels


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_elastic_fused_forward_matches_two_phase(orc, ads, cuda, dtype):
    """The fused one-launch forward step agrees with the two-phase step (stress launch, then
    velocity launch) on traces, misfit and gradients, for chunk heights from 1 row to more than
    the grid, on a grid wide enough for plain wavefronts inside the interior rectangle and
    general ones over the CPML frame, with uneven shots (fewer active shots than a workgroup
    holds near the end).  The two differ only where the compiler contracts products into
    fused multiply-adds differently, so the bar is rounding level; the FP64 gradient also
    meets the oracle's 1e-10."""
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NX=150, NY=141, NSTEP=60, npml=8, nshots=5,
                                             uneven=True, seed=12)
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=dtype, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m, dtype) for m in models]
    tm = [T(m * f, dtype) for m, f in zip(models, (1.02, 0.97, 1.04))]
    plan.set_forward_kernels(0)
    obs = plan.forward(*tm, srcv, 13).clone()
    L0, g0, t0 = plan.gradient(*mt, srcv, obs, 13, want_traces=True)
    g0 = [x.clone() for x in g0]
    t0 = t0.clone()
    bar_t, bar_g = (1e-12, 1e-11) if dtype == torch.float64 else (1e-5, 1e-4)
    for rows in (1, 5, 8, 32, 200, 0):
        plan.set_forward_kernels(1, rows)
        ob = plan.forward(*tm, srcv, 13)
        assert rel(ob.cpu(), obs.cpu()) < bar_t, rows
        L1, g1, t1 = plan.gradient(*mt, srcv, obs, 13, want_traces=True)
        assert rel(t1.cpu(), t0.cpu()) < bar_t, rows
        assert abs(L1.item() - L0.item()) <= bar_t * abs(L0.item()), rows
        for a, b in zip(g1, g0):
            assert rel(a.cpu(), b.cpu()) < bar_g, rows
    if dtype == torch.float64:
        o, g_ref = [], [0.0, 0.0, 0.0]
        for s in range(len(srcs)):
            oo, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.02, 0.97, 1.04))
            _, g, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=oo)
            o.append(oo)
            g_ref = [a + b for a, b in zip(g_ref, g)]
        plan.set_forward_kernels(1, 32)
        _, g1, _ = plan.gradient(*mt, srcv, plan.pack_traces(o), 13)
        assert max(rel(a.cpu(), b) for a, b in zip(g1, g_ref)) < 1e-10
    with pytest.raises(Exception, match="fused"):
        plan.set_forward_kernels(2, 8)


def test_mpi_elastic_fused_forward_matches_two_phase(orc, ads, cuda):
    """The fused forward step in the subdomain form (no parameter averaging, the block's own
    update range and profile index) agrees with the two-phase step on an 80 x 80 block, wide
    enough for plain wavefronts inside its interior rectangle."""
    p, models, src, rcv, pml = _mpi_case(ads, orc, n=80, NSTEP=40, seed=43)
    plan = ads.ElasticPlan(p, [ads.ElasticSource(src.srci, src.srcj, src.srctype, src.srcv)],
                           [ads.ElasticReceiver(rcv.rcvi, rcv.rcvj, rcv.rcvtype)],
                           dtype=torch.float64)
    srcv = plan.pack_srcv([src.srcv])
    mt = [T(m) for m in models]
    tm = [T(m * f) for m, f in zip(models, (1.02, 0.97, 1.04))]
    plan.set_forward_kernels(0)
    obs = plan.forward(*tm, srcv, 9).clone()
    L0, g0, t0 = plan.gradient(*mt, srcv, obs, 9, want_traces=True)
    g0 = [x.clone() for x in g0]
    t0 = t0.clone()
    for rows in (8, 3, 0):
        plan.set_forward_kernels(1, rows)
        assert rel(plan.forward(*tm, srcv, 9).cpu(), obs.cpu()) < 1e-12, rows
        L1, g1, t1 = plan.gradient(*mt, srcv, obs, 9, want_traces=True)
        assert rel(t1.cpu(), t0.cpu()) < 1e-12, rows
        assert abs(L1.item() - L0.item()) <= 1e-12 * abs(L0.item()), rows
        for a, b in zip(g1, g0):
            assert rel(a.cpu(), b.cpu()) < 1e-11, rows


def _min_rebuilds(n, s, memo={}):
    """Least segment advances to reverse n segments with s free slots (exhaustive search)."""
    key = (n, s)
    if key not in memo:
        if n == 1:
            memo[key] = 1
        elif s == 0:
            memo[key] = n * (n + 1) // 2
        else:
            memo[key] = min(j + _min_rebuilds(n - j, s - 1) + _min_rebuilds(j, s) for j in range(1, n))
    return memo[key]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_elastic_binomial_schedules_match_one_level(orc, ads, cuda, dtype):
    """Every snapshot count gives the one-level schedule's misfit and gradients bit for bit,
    the FP64 gradient matches the oracle, and the forward levels run equal the initial pass plus
    K times the fewest segment rebuilds an exhaustive search finds."""
    import ctypes as C
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NSTEP=60, nshots=3, seed=4)
    obs, g_ref = [], [0.0, 0.0, 0.0]
    for s in range(3):
        o, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], scale=(1.02, 0.97, 1.04))
        _, g, _ = oracle(orc, p, models, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        g_ref = [a + b for a, b in zip(g_ref, g)]
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=dtype, nstep=nstep)
    lib = plan.lib
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m, dtype) for m in models]
    ob = plan.pack_traces(obs)
    for K in (6, 5):                       # 10 and 12 segments of exactly K levels
        M = p.NSTEP // K
        plan.set_snapshots(-1)
        L1, g1, _ = plan.gradient(*mt, srcv, ob, K)
        g1 = [x.cpu().numpy() for x in g1]
        bar = 1e-10 if dtype == torch.float64 else 1e-3
        assert max(rel(a, b) for a, b in zip(g1, g_ref)) < bar
        for S in (0, 1, 2, 4, M - 2, M - 1):
            plan.set_snapshots(S)
            L.check(lib.ads_profile_begin())
            LS, gS, _ = plan.gradient(*mt, srcv, ob, K)
            torch.cuda.synchronize()
            L.check(lib.ads_profile_end())
            tm, n, b = C.c_double(), C.c_int64(), C.c_double()
            L.check(lib.ads_profile_query(2, C.byref(tm), C.byref(n), C.byref(b)))
            assert LS.item() == L1.item(), (K, S)
            for a, bb in zip(gS, g1):
                np.testing.assert_array_equal(a.cpu().numpy(), bb, err_msg=f"K {K}, S {S}")
            want = p.NSTEP + K * (_min_rebuilds(M, min(S, M - 1)) - M)
            assert n.value == want, (K, S, n.value, want)
    plan.set_snapshots(-1)
    budget = min(plan.workspace_size(k) for k in range(1, p.NSTEP + 1)) - 1  # no one-level fits
    K, S = plan.choose_schedule(budget)
    assert plan.snapshots == S and S < -(-p.NSTEP // K) - 1 and plan.workspace_size(K) <= budget
    LS, gS, _ = plan.gradient(*mt, srcv, ob, K)
    assert max(rel(a.cpu(), b) for a, b in zip(gS, g_ref)) < (1e-10 if dtype == torch.float64 else 1e-3)


def test_profile_sampling_brackets_every_kth_launch(orc, ads, cuda):
    """ads_profile_begin_sampled(k) times launches 0, k, 2k, ... of each class: an elastic
    gradient (one level per launch) then reports ceil(N / k) of the N launches ads_profile_begin
    counts, with the same algorithmic bytes per launch; a stride below 1 is refused."""
    import ctypes as C
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    p, models, srcs, rcvs, nstep = make_case(orc, ads, NSTEP=40, nshots=2, seed=6)
    plan = ads.ElasticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nstep)
    lib = plan.lib
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    mt = [T(m, torch.float32) for m in models]
    ob = plan.forward(*[T(m * 1.02, torch.float32) for m in models], srcv, 7).clone()

    def counts(begin):
        L.check(begin())
        plan.gradient(*mt, srcv, ob, 7)
        torch.cuda.synchronize()
        L.check(lib.ads_profile_end())
        out = []
        for cls in (2, 3, 4, 5):
            tm, n, b = C.c_double(), C.c_int64(), C.c_double()
            L.check(lib.ads_profile_query(cls, C.byref(tm), C.byref(n), C.byref(b)))
            out.append((tm.value, n.value, b.value))
        return out

    full = counts(lib.ads_profile_begin)
    # the fused forward step is timed in the stress class; the velocity class then runs nothing
    assert full[0][1] > 0 and full[2][1] > 0 and full[3][1] > 0
    for k in (1, 3, 4):
        part = counts(lambda: lib.ads_profile_begin_sampled(k))
        for (tf, nf, bf), (tp, np_, bp) in zip(full, part):
            assert np_ == -(-nf // k), (k, nf, np_)
            if nf:
                assert tp > 0.0
                assert abs(bp / np_ - bf / nf) <= 1e-9 * (bf / nf)
    assert lib.ads_profile_begin_sampled(0) != 0
    assert "stride" in lib.ads_last_error().decode()
