"""GPU parity tests of the acoustic HIP path against the FP64 oracle and the golden vectors.

Tolerances: FP64 paths match the oracle to 1e-10 relative (1e-12 for single steps; the
drop-in per-step op is bit-exact); FP32 paths meet the stated bar of <= 1e-4 relative L2 on
traces/wavefields and <= 1e-3 on gradients against the FP64 oracle.
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


# ------------------------------------------------------------------------------------------
# drop-in per-step op
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["10x10", "7x12"])
def test_one_step_op_bit_exact_f64(ads, golden, cuda, name):
    g = golden("acoustic_one_step.npz")
    NX, NY = (int(x) for x in g[f"{name}/dims"])
    dt, hx, hy = (float(x) for x in g[f"{name}/scalars"])
    ins = [T(g[f"{name}/{k}"]) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")]
    for t in ins:
        t.requires_grad_(True)
    u, po, so = ads.acoustic_one_step(*ins, dt, hx, hy, NX, NY)
    np.testing.assert_array_equal(u.detach().cpu().numpy(), g[f"{name}/u"])
    np.testing.assert_array_equal(po.detach().cpu().numpy(), g[f"{name}/phiout"])
    np.testing.assert_array_equal(so.detach().cpu().numpy(), g[f"{name}/psiout"])
    gu, gp, gs = (T(g[f"{name}/{k}"]) for k in ("grad_u", "grad_phiout", "grad_psiout"))
    grads = torch.autograd.grad((u, po, so), [ins[0], ins[1], ins[2], ins[3], ins[6]], (gu, gp, gs))
    for gr, k in zip(grads, ("grad_w", "grad_wold", "grad_phi", "grad_psi", "grad_c")):
        np.testing.assert_array_equal(gr.cpu().numpy(), g[f"{name}/{k}"], err_msg=k)


def test_one_step_op_f32(ads, golden, cuda):
    g = golden("acoustic_one_step.npz")
    name = "10x10"
    NX, NY = (int(x) for x in g[f"{name}/dims"])
    dt, hx, hy = (float(x) for x in g[f"{name}/scalars"])
    ins = [T(g[f"{name}/{k}"], torch.float32) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")]
    u, po, so = ads.acoustic_one_step(*ins, dt, hx, hy, NX, NY)
    assert rel(u.cpu(), g[f"{name}/u"]) < 1e-6
    assert rel(po.cpu(), g[f"{name}/phiout"]) < 1e-6


def test_mpi_one_step_op_bit_exact(ads, golden, cuda):
    import ctypes as C
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    g = golden("mpi_acoustic_one_step.npz")
    n = int(g["n"])
    dt, hx, hy = (float(x) for x in g["scalars"])
    ins = [T(g[k]) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")]
    outs = [torch.empty(n * n, dtype=torch.float64, device="cuda") for _ in range(3)]
    lib = ads.load_library()
    L.check(lib.ads_mpi_acoustic_one_step_f64(*[C.c_void_p(t.data_ptr()) for t in ins], dt, hx, hy, n,
                                              *[C.c_void_p(t.data_ptr()) for t in outs], None))
    torch.cuda.synchronize()
    for o, k in zip(outs, ("u", "phiout", "psiout")):
        np.testing.assert_array_equal(o.cpu().numpy(), g[k], err_msg=k)


def test_index_ops(ads, golden, cuda):
    import ctypes as C
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    g = golden("index_ops.npz")
    lib = ads.load_library()
    v, upd = T(g["v"]), T(g["upd"])
    ii = torch.as_tensor(g["ii"], device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())
    out = torch.empty(len(g["ii"]), dtype=torch.float64, device="cuda")
    L.check(lib.ads_gather_f64(P(v), 50, P(ii), len(g["ii"]), P(out), None))
    sa = torch.empty(50, dtype=torch.float64, device="cuda")
    L.check(lib.ads_scatter_add_f64(P(v), 50, P(ii), P(upd), len(g["ii"]), P(sa), None))
    snd = torch.empty(50, dtype=torch.float64, device="cuda")
    L.check(lib.ads_scatter_nd_f64(P(ii), P(upd), len(g["ii"]), 50, P(snd), None))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), g["gather"])
    np.testing.assert_array_equal(sa.cpu().numpy(), g["scatter_add"])
    np.testing.assert_array_equal(snd.cpu().numpy(), g["scatter_nd"])


# ------------------------------------------------------------------------------------------
# solver level
# ------------------------------------------------------------------------------------------

def make_case(orc, ads, NX=30, NY=26, NSTEP=80, npml=6, nshots=1, seed=0, uneven=False,
              kernel=0, h=10.0, dt=1e-3):
    rng = np.random.default_rng(seed)
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=h, DELTAY=h, DELTAT=dt,
                                     NPOINTS_PML=npml, vp_ref=2000.0, PropagatorKernel=kernel)
    c = 1800.0 + 600.0 * rng.random((NX + 2, NY + 2))
    nstep = [NSTEP - (7 * s if uneven else 0) for s in range(nshots)]
    srcs, rcvs = [], []
    for s in range(nshots):
        i0 = 2 + (s * 5) % (NX - 2)
        w = orc.ricker(nstep[s], 6.0, 12.0 + s, 1e6)
        if s % 3 == 2:  # two sources, one on the ring, one duplicated
            srcs.append(ads.AcousticSource([i0, 1, i0], [3, 5, 3],
                                           np.stack([w, 0.5 * w, -0.25 * w], 1)))
        else:
            srcs.append(ads.AcousticSource([i0], [3], w[:, None]))
        ri = np.arange(1, NX + 3, 2 + s % 2)
        rj = np.full(len(ri), 2 + s % 2)
        if s % 2 == 1:  # duplicated receivers (on the ring, inside) and one in the PML corner
            ri = np.concatenate([ri, [ri[0], ri[1], 2]])
            rj = np.concatenate([rj, [rj[0], rj[1], NY]])
        rcvs.append(ads.AcousticReceiver(ri, rj))
    return p, c, srcs, rcvs, nstep


def oracle_shot(orc, p, c, src, rcv, nstep, cmod=None, obs=None):
    """Oracle traces (and gradient if obs is given) for one shot of record length nstep."""
    NX, NY = p.NX, p.NY
    Sx, Sy = orc.acoustic_pml(NX, NY, p.DELTAX, p.DELTAY, p.NPOINTS_PML, p.Rcoef, p.vp_ref,
                              use=(p.USE_PML_XMIN, p.USE_PML_XMAX, p.USE_PML_YMIN, p.USE_PML_YMAX))
    k = 0 if p.PropagatorKernel == 0 else 1
    cc = c if cmod is None else cmod
    u, ph, ps = orc.acoustic_forward(k, NX, NY, nstep, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy, cc ** 2,
                                     src.srci, src.srcj, src.srcv)
    tr = orc.acoustic_receive(u, rcv.rcvi, rcv.rcvj, NY)
    if obs is None:
        return tr
    L, g, _ = orc.acoustic_misfit_gradient(k, NX, NY, nstep, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy,
                                           c, src.srci, src.srcj, src.srcv, rcv.rcvi, rcv.rcvj, obs)
    return tr, L, g


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("K", [7, 1000])
def test_plan_forward_and_gradient_f64(orc, ads, cuda, kernel, K):
    p, c, srcs, rcvs, nstep = make_case(orc, ads, nshots=3, kernel=kernel, uneven=True)
    obs, L_ref, g_ref = [], 0.0, np.zeros_like(c)
    for s in range(3):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03)
        tr, L, g = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        L_ref += L
        g_ref += g
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct = T(c)
    tr = plan.forward(ct, srcv, K)
    per = plan.unpack_traces(tr)
    for s in range(3):
        ref = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s])
        assert rel(per[s].cpu(), ref) < 1e-12, (s, rel(per[s].cpu(), ref))
    loss, g, tr2 = plan.gradient(ct, srcv, plan.pack_traces(obs), K, want_traces=True)
    torch.cuda.synchronize()
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    assert rel(g.cpu(), g_ref) < 1e-10, rel(g.cpu(), g_ref)
    np.testing.assert_array_equal(tr2.cpu().numpy(), tr.cpu().numpy())
    # backward with an explicit trace cotangent equals gradient()'s residual path
    tr3 = plan.forward(ct, srcv, K)
    gb = plan.backward(ct, srcv, 2.0 * (tr3 - plan.pack_traces(obs)), K)
    assert rel(gb.cpu(), g_ref) < 1e-10


@pytest.mark.parametrize("kernel", [0, 1])
def test_shots_without_sources(orc, ads, cuda, kernel):
    """A shot with no source records zeros and adds nothing to the gradient, beside shots with
    sources in the same plan (the frame work's source table has no entry for it)."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, nshots=3, kernel=kernel, seed=17)
    srcs[1] = ads.AcousticSource(np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros((p.NSTEP, 0)))
    obs, g_ref = [], np.zeros_like(c)
    for s in (0, 2):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03)
        g_ref += oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)[2]
        obs.append(o)
    obs.insert(1, np.zeros((p.NSTEP + 1, len(rcvs[1].rcvi))))
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    per = plan.unpack_traces(plan.forward(T(c), srcv, 9))
    assert not per[1].abs().max().item()
    for s in (0, 2):
        assert rel(per[s].cpu(), oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s])) < 1e-12, s
    _, g, _ = plan.gradient(T(c), srcv, plan.pack_traces(obs), 9)
    assert rel(g.cpu(), g_ref) < 1e-10


# interior rectangle / PML frame splits: partial PML sides, odd sizes, an interior narrower than
# one vector, no PML at all, and a PML that covers the whole grid (empty rectangle)
SPLITS = [
    dict(NX=29, NY=33, npml=5, use=(False, True, True, False)),
    dict(NX=21, NY=19, npml=4, use=(True, True, False, False)),
    dict(NX=18, NY=15, npml=6, use=(True, True, True, True)),
    dict(NX=17, NY=23, npml=4, use=(False, False, False, False)),
    dict(NX=12, NY=12, npml=7, use=(True, True, True, True)),
]


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("split", range(len(SPLITS)))
def test_plan_grid_splits_f64(orc, ads, cuda, kernel, split):
    cfg = SPLITS[split]
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=cfg["NX"], NY=cfg["NY"], NSTEP=60,
                                        npml=cfg["npml"], nshots=3, kernel=kernel, seed=20 + split,
                                        uneven=True)
    p.USE_PML_XMIN, p.USE_PML_XMAX, p.USE_PML_YMIN, p.USE_PML_YMAX = cfg["use"]
    obs, L_ref, g_ref = [], 0.0, np.zeros_like(c)
    for s in range(3):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03)
        tr, L, g = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        L_ref += L
        g_ref += g
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    for shape in [(1, 2, 1, 1, 2, 1), (2, 3, 2, 4, 1, 1)]:
        plan.set_shape(shape)
        loss, g, tr = plan.gradient(T(c), srcv, plan.pack_traces(obs), 8, want_traces=True)
        per = plan.unpack_traces(tr)
        for s in range(3):
            ref = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s])
            assert rel(per[s].cpu(), ref) < 1e-12, (shape, s)
        assert loss.item() == pytest.approx(L_ref, rel=1e-10)
        assert rel(g.cpu(), g_ref) < 1e-10, (shape, rel(g.cpu(), g_ref))
    # the full-history (dense layout) drop-in on the same split
    ct = T(c).requires_grad_(True)
    ap = ads.AcousticPropagatorSolver(p, srcs[0], ct)
    rr = ads.SimulatedObservation_(ap, rcvs[0])
    o0 = oracle_shot(orc, p, c, srcs[0], rcvs[0], p.NSTEP, cmod=c * 1.03)
    torch.sum((rr - T(o0)) ** 2).backward()
    _, _, g0 = oracle_shot(orc, p, c, srcs[0], rcvs[0], p.NSTEP, obs=o0)
    assert rel(ct.grad.cpu(), g0) < 1e-10


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("kernel", [0, 1])
def test_two_level_blocking_is_bit_identical(orc, ads, cuda, kernel, dtype):
    """The forward's temporal blocking (two levels per interior-rectangle pass, the default)
    gives the single-level traces, loss and gradient bit for bit: over the grid splits plus a
    grid wide and tall enough for several column segments and row strips, five kernel shapes
    (the last two with the streaming rectangle, 8 and 16 rows per wavefront),
    two checkpoint intervals, an odd record length and uneven records.  A plan with a source
    inside the rectangle (the last case) steps one level at a time."""
    cases = SPLITS + [dict(NX=70, NY=300, npml=8, use=(True, True, True, True)),
                      dict(NX=40, NY=44, npml=5, use=(True, True, True, True), centre=True)]
    seen = set()
    for split, cfg in enumerate(cases):
        p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=cfg["NX"], NY=cfg["NY"], NSTEP=41,
                                            npml=cfg["npml"], nshots=5, kernel=kernel,
                                            seed=40 + split, uneven=True)
        p.USE_PML_XMIN, p.USE_PML_XMAX, p.USE_PML_YMIN, p.USE_PML_YMAX = cfg["use"]
        if cfg.get("centre"):
            srcs[1] = ads.AcousticSource([cfg["NX"] // 2], [cfg["NY"] // 2], srcs[1].srcv)
        plan = ads.AcousticPlan(p, srcs, rcvs, dtype=dtype, nstep=nstep)
        seen.add(plan.blocking)
        if cfg.get("centre"):
            assert plan.blocking == 1
        if split == len(SPLITS):
            assert plan.blocking == 2
        srcv = plan.pack_srcv([s.srcv for s in srcs])
        ct = T(c, dtype)
        obs = plan.forward(ct * 1.02, srcv, 9).clone()
        for shape in [(1, 2, 1, 1, 2, 1), (2, 3, 1, 2, 3, 1), (4, 5, 1, 1, 5, 1),
                      (8, 3, 1, 2, 3, 1), (16, 5, 1, 1, 5, 1)]:
            plan.set_shape(shape)
            for K in (5, 1000):
                res = []
                for levels in (1, 2):
                    plan.set_blocking(levels)
                    loss, g, tr = plan.gradient(ct, srcv, obs, K, want_traces=True)
                    res.append((loss.clone(), g.clone(), tr.clone()))
                (l1, g1, t1), (l2, g2, t2) = res
                assert torch.equal(t1, t2), (split, shape, K)
                assert torch.equal(g1, g2), (split, shape, K)
                assert torch.equal(l1, l2), (split, shape, K)
        plan.set_blocking(2)
    assert seen == {1, 2}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_two_level_frame_heights_bit_identical(orc, ads, cuda, dtype):
    """The two-level frame work takes items of 4 .. 64 rows, forced per plan or chosen per
    launch.  66 shots with records of 500 - 7 s steps let the number of active shots fall from
    66 to 1 within one gradient; with every forced height and with the per-launch choice,
    single- and two-level stepping must agree bit for bit, with the streaming and the
    register-blocked rectangle."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=70, NY=300, NSTEP=500, npml=8, nshots=66,
                                        kernel=1, seed=77, uneven=True)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=dtype, nstep=nstep)
    assert plan.blocking == 2
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct = T(c, dtype)
    obs = plan.forward(ct * 1.02, srcv, 1000).clone()
    for shape in [(8, 16, 1, 2, 16, 1), (2, 16, 1, 2, 16, 1)]:
        plan.set_shape(shape)
        plan.set_blocking(1)
        l1, g1, t1 = plan.gradient(ct, srcv, obs, 1000, want_traces=True)
        l1, g1, t1 = l1.clone(), g1.clone(), t1.clone()
        plan.set_blocking(2)
        for rows in (4, 8, 16, 32, 64, 0):
            plan.set_frame_rows(rows)
            l2, g2, t2 = plan.gradient(ct, srcv, obs, 1000, want_traces=True)
            assert torch.equal(t1, t2), (shape, rows)
            assert torch.equal(g1, g2), (shape, rows)
            assert torch.equal(l1, l2), (shape, rows)
    with pytest.raises(Exception, match="frame rows"):
        plan.set_frame_rows(12)
    plan.set_frame_rows(0)


@pytest.mark.parametrize("kernel", [0, 1])
def test_plan_f32_against_f64_oracle(orc, ads, cuda, kernel):
    """FP32 fields vs the FP64 oracle: traces <= 1e-4 rel L2, gradient <= 1e-3 rel L2."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=60, NY=50, NSTEP=300, npml=10, nshots=2,
                                        kernel=kernel, seed=3)
    obs, g_ref, trs = [], np.zeros_like(c), []
    for s in range(2):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.02)
        tr, L, g = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)
        obs.append(o)
        trs.append(tr)
        g_ref += g
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    loss, g, tr = plan.gradient(T(c, torch.float32), srcv, plan.pack_traces(obs), 40,
                                want_traces=True)
    per = plan.unpack_traces(tr)
    for s in range(2):
        assert rel(per[s].cpu(), trs[s]) < 1e-4, rel(per[s].cpu(), trs[s])
    assert rel(g.cpu(), g_ref) < 1e-3, rel(g.cpu(), g_ref)


@pytest.mark.parametrize("kernel", [0, 1])
def test_history_dropin_and_autograd(orc, ads, cuda, kernel):
    """AcousticPropagatorSolver + SimulatedObservation_ + torch autograd (the plain drop-in)."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=24, NY=20, NSTEP=60, npml=5, kernel=kernel, seed=5)
    src, rcv = srcs[0], rcvs[0]
    Sx, Sy = orc.acoustic_pml(p.NX, p.NY, p.DELTAX, p.DELTAY, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    k = 0 if kernel == 0 else 1
    u_ref, ph_ref, ps_ref = orc.acoustic_forward(k, p.NX, p.NY, p.NSTEP, p.DELTAT, p.DELTAX, p.DELTAY,
                                                 Sx, Sy, c ** 2, src.srci, src.srcj, src.srcv)
    ct = T(c).requires_grad_(True)
    ap = ads.AcousticPropagatorSolver(p, src, ct)
    assert tuple(ap.u.shape) == (p.NSTEP + 1, p.NX + 2, p.NY + 2)
    assert rel(ap.u.detach().cpu(), u_ref) < 1e-12
    assert rel(ap.φ.cpu(), ph_ref) < 1e-12
    assert rel(ap.ψ.cpu(), ps_ref) < 1e-12
    obs = orc.acoustic_receive(orc.acoustic_forward(k, p.NX, p.NY, p.NSTEP, p.DELTAT, p.DELTAX,
                                                    p.DELTAY, Sx, Sy, (1.04 * c) ** 2, src.srci,
                                                    src.srcj, src.srcv)[0], rcv.rcvi, rcv.rcvj, p.NY)
    rr = ads.SimulatedObservation_(ap, rcv)
    loss = torch.sum((rr - T(obs)) ** 2)
    loss.backward()
    _, L_ref, g_ref = oracle_shot(orc, p, c, src, rcv, p.NSTEP, obs=obs)
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    assert rel(ct.grad.cpu(), g_ref) < 1e-10


def test_traces_autograd_function(orc, ads, cuda):
    p, c, srcs, rcvs, nstep = make_case(orc, ads, nshots=2, kernel=0, seed=7)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct = T(c).requires_grad_(True)
    tr = ads.acoustic_traces(plan, ct, srcv, 9)
    w = T(np.random.default_rng(1).standard_normal(tr.numel()))
    (tr * w).sum().backward()
    # dot-product check: <d tr/dc v, w> vs <v, grad>
    v = np.random.default_rng(2).standard_normal(c.shape)
    eps = 1e-3
    f = lambda cc: float((plan.forward(T(cc), srcv, 9) * w).sum().item())
    fd = (f(c + eps * v) - f(c - eps * v)) / (2 * eps)
    assert np.sum(ct.grad.cpu().numpy() * v) == pytest.approx(fd, rel=1e-6)


@pytest.mark.parametrize("kernel", [0, 1])
def test_traces_autograd_interleaved_calls(orc, ads, cuda, kernel):
    """Two acoustic_traces calls with different models on one plan in one autograd graph, plus
    a plan.gradient call between the forwards and the backward: each pullback must still see
    its own forward's checkpoints (the workspace token makes it re-run the forward when another
    call used the workspace).  Both gradients equal the oracle's."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, nshots=2, kernel=kernel, seed=8)
    c2 = c * (1.0 + 0.05 * np.random.default_rng(9).random(c.shape))
    obs = [oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03) for s in range(2)]
    g1_ref = sum(oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=obs[s])[2] for s in range(2))
    g2_ref = sum(oracle_shot(orc, p, c2, srcs[s], rcvs[s], nstep[s], obs=obs[s])[2] for s in range(2))
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ob = plan.pack_traces(obs)
    K = 9  # several segments: the pullback needs the checkpoints
    a = T(c).requires_grad_(True)
    b = T(c2).requires_grad_(True)
    t1 = ads.acoustic_traces(plan, a, srcv, K)
    t2 = ads.acoustic_traces(plan, b, srcv, K)
    plan.gradient(T(c * 0.97), srcv, ob, K)          # overwrites the workspace once more
    loss = torch.sum((t1 - ob) ** 2) + torch.sum((t2 - ob) ** 2)
    loss.backward()
    assert rel(a.grad.cpu(), g1_ref) < 1e-10, rel(a.grad.cpu(), g1_ref)
    assert rel(b.grad.cpu(), g2_ref) < 1e-10, rel(b.grad.cpu(), g2_ref)


def test_workspace_too_small_is_an_error(orc, ads, cuda):
    import ctypes as C
    p, c, srcs, rcvs, nstep = make_case(orc, ads)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ws = torch.empty(1024, dtype=torch.uint8, device="cuda")
    tr = torch.empty(plan.trace_elems, device="cuda")
    rc = plan.lib.ads_acoustic_plan_forward(plan.h, C.c_void_p(T(c, torch.float32).data_ptr()),
                                            C.c_void_p(srcv.data_ptr()), C.c_void_p(tr.data_ptr()),
                                            C.c_void_p(ws.data_ptr()), 1024, 5, None)
    assert rc == 6


def test_choose_interval_minimises_recomputation(orc, ads, cuda):
    """For a budget, the chosen K needs the fewest segments that fit and, among the K with that
    segment count, recomputes the fewest levels; checked by brute force over every K."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NSTEP=80)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32)
    L = p.NSTEP - 1
    sizes = {K: plan.workspace_size(K) for K in range(1, L + 1)}
    for budget in (sizes[L], sizes[L] - 1, sizes[40], sizes[27] + 7, sizes[9], sizes[1]):
        K = plan.choose_interval(budget)
        assert sizes[K] <= budget
        segs = lambda k: -(-L // k)
        fit = [k for k in sizes if sizes[k] <= budget]
        best_m = min(segs(k) for k in fit)
        assert segs(K) == best_m, (budget, K)
        assert K == min(k for k in fit if segs(k) == best_m), (budget, K)
    assert plan.choose_interval(sizes[L]) == L  # everything fits: no recomputation


@pytest.mark.parametrize("kernel", [0, 1])
def test_kernel_shapes_and_autotune_agree(orc, ads, cuda, kernel):
    """Every (rows per wavefront, shots per workgroup, shots per batch) shape, the streaming
    forward rectangle (8 and 16 rows) included, and the autotuned one, gives the oracle's
    gradient (shot groups change only the order of the dL/dC sum)."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=37, NY=70, NSTEP=50, nshots=5, kernel=kernel,
                                        seed=11, uneven=True)
    obs, g_ref = [], np.zeros_like(c)
    for s in range(5):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 0.97)
        g_ref += oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)[2]
        obs.append(o)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float64, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct, ob = T(c), plan.pack_traces(obs)
    for shape in [(2, 1, 1, 2, 1, 1), (4, 2, 1, 1, 3, 2), (1, 5, 2, 4, 2, 1), (2, 32, 2, 2, 32, 2),
                  (1, 3, 2, 1, 5, 2), (8, 2, 1, 2, 2, 1), (16, 4, 1, 1, 4, 1)]:
        plan.set_shape(shape)
        _, g, _ = plan.gradient(ct, srcv, ob, 6)
        assert rel(g.cpu(), g_ref) < 1e-10, (shape, rel(g.cpu(), g_ref))
    shape = plan.autotune(ct, srcv, 6, reps=1)
    assert len(shape) == 6 and shape[0] in (1, 2, 4, 8) and shape[3] in (1, 2, 4)
    assert shape[2] in (1, 2) and shape[5] in (1, 2)
    _, g, _ = plan.gradient(ct, srcv, ob, 6)
    assert rel(g.cpu(), g_ref) < 1e-10


@pytest.mark.parametrize("kernel", [0, 1])
def test_graph_replay_matches_stream_launches(orc, ads, cuda, kernel):
    """Checkpoint segments captured into hipGraphs and replayed give the stream-launched
    results bit for bit, on every replay, with eager work on the same buffers in between."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, nshots=3, kernel=kernel, uneven=True, seed=13)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct = T(c, torch.float32)
    obs = plan.forward(ct * 1.02, srcv, 9).clone()
    loss0, g0, tr0 = plan.gradient(ct, srcv, obs, 9, want_traces=True)
    g0, tr0, loss0 = g0.clone(), tr0.clone(), loss0.item()
    plan.set_graph(True)
    for rep in range(3):
        loss, g, tr = plan.gradient(ct, srcv, obs, 9, want_traces=True)
        torch.cuda.synchronize()
        assert loss.item() == loss0, rep
        assert torch.equal(g, g0), rep
        assert torch.equal(tr, tr0), rep
        # eager work on the plan's buffers between replays
        plan.set_graph(False)
        _ = plan.forward(ct * 0.99, srcv, 9)
        plan.set_graph(True)


# ------------------------------------------------------------------------------------------
# checkpoint schedules: binomial reversal with S snapshot slots
# ------------------------------------------------------------------------------------------

def _min_rebuilds(n, s, memo={}):
    """Least segment advances to reverse n segments from a stored state with s free slots, by
    exhaustive search over every first split (each segment is rebuilt once before its adjoint;
    the initial forward has already made the first descent's advances)."""
    key = (n, s)
    if key not in memo:
        if n == 1:
            memo[key] = 1
        elif s == 0:
            memo[key] = n * (n + 1) // 2
        else:
            memo[key] = min(j + _min_rebuilds(n - j, s - 1) + _min_rebuilds(j, s)
                            for j in range(1, n))
    return memo[key]


def _fwd_levels(lib, L):
    import ctypes as C
    tm, n, b = C.c_double(), C.c_int64(), C.c_double()
    L.check(lib.ads_profile_query(0, C.byref(tm), C.byref(n), C.byref(b)))
    return n.value


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_binomial_schedules_match_one_level(orc, ads, cuda, kernel, dtype):
    """Every snapshot count S gives the one-level schedule's misfit and gradient bit for bit
    (the rebuilt levels are the same arithmetic), the FP64 gradient matches the oracle, and the
    forward levels run equal the initial pass plus K times the fewest segment rebuilds an
    exhaustive search finds for (M segments, S slots) -- the schedule is optimal."""
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=33, NY=29, NSTEP=121, nshots=3,
                                        kernel=kernel, seed=5)
    obs, g_ref = [], np.zeros_like(c)
    for s in range(3):
        o = oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03)
        g_ref += oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], obs=o)[2]
        obs.append(o)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=dtype, nstep=nstep)
    lib = plan.lib
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    ct, ob = T(c, dtype), plan.pack_traces(obs)
    levels = p.NSTEP - 1
    for K in (8, 10):                     # 15 and 12 segments of exactly K levels
        M = levels // K
        plan.set_snapshots(-1)
        L1, g1, tr1 = plan.gradient(ct, srcv, ob, K, want_traces=True)
        g1 = g1.cpu().numpy()
        if dtype == torch.float64:
            assert rel(g1, g_ref) < 1e-10
        else:
            assert rel(g1, g_ref) < 1e-3
        sizes = []
        for S in (0, 1, 2, 3, 5, M - 2, M - 1, M + 4):
            plan.set_snapshots(S)
            sizes.append(plan.workspace_size(K))
            L.check(lib.ads_profile_begin())
            LS, gS, trS = plan.gradient(ct, srcv, ob, K, want_traces=True)
            torch.cuda.synchronize()
            L.check(lib.ads_profile_end())
            ran = _fwd_levels(lib, L)
            assert LS.item() == L1.item(), (K, S)
            assert torch.equal(trS, tr1), (K, S)
            np.testing.assert_array_equal(gS.cpu().numpy(), g1, err_msg=f"K {K}, S {S}")
            s_eff = min(S, M - 1)
            want = levels + K * (_min_rebuilds(M, s_eff) - M)
            assert ran == want, (K, S, ran, want)
        assert sizes == sorted(sizes) and sizes[0] < sizes[-1]
        assert sizes[-2] == sizes[-1]   # S >= M-1 is the one-level layout


def test_choose_schedule_fits_and_prefers_fewer_rebuilds(orc, ads, cuda):
    """choose_schedule returns a (K, S) whose workspace fits the budget; with room for the
    whole record it rebuilds nothing, and a budget below every one-level layout still gets a
    binomial schedule whose gradient equals the one-level gradient."""
    p, c, srcs, rcvs, nstep = make_case(orc, ads, NX=33, NY=29, NSTEP=121, nshots=2, kernel=1,
                                        seed=6, uneven=True)
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=torch.float32, nstep=nstep)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    obs = [oracle_shot(orc, p, c, srcs[s], rcvs[s], nstep[s], cmod=c * 1.03) for s in range(2)]
    ct, ob = T(c, torch.float32), plan.pack_traces(obs)
    levels = p.NSTEP - 1
    plan.set_snapshots(-1)
    one = {K: plan.workspace_size(K) for K in range(1, levels + 1)}
    K, S = plan.choose_schedule(one[levels])
    assert (K, S) == (levels, 0) and plan.workspace_size(K) <= one[levels]
    L1, g1, _ = plan.gradient(ct, srcv, ob, K)
    budget = min(one.values()) - 1                     # no one-level layout fits
    K, S = plan.choose_schedule(budget)
    assert plan.snapshots == S and S < -(-levels // K) - 1
    assert plan.workspace_size(K) <= budget
    LS, gS, _ = plan.gradient(ct, srcv, ob, K)
    assert LS.item() == pytest.approx(L1.item(), rel=1e-6)
    assert rel(gS.cpu().numpy(), g1.cpu().numpy()) < 1e-6
