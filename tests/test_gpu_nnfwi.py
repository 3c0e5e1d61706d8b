"""NN-reparametrised FWI (examples/nn_fwi/NNFWI.jl:60-75): velocity = add_initial_model(
Generator(z) * scale, v0); the misfit's gradient w.r.t. the network weights flows through
acoustic_traces, whose pullback is the HIP adjoint.  Checked against central finite
differences of the misfit on 20 random weights (FP64), and run once at the 400 x 800,
24-shot, 2500-step scale in FP32."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def make(ads, NX, NY, NSTEP, nshots, dtype, h=10.0, dt=1e-3, npml=10, seed=5):
    rng = np.random.default_rng(seed)
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=h, DELTAY=h, DELTAT=dt,
                                     NPOINTS_PML=npml, vp_ref=2000.0)
    w = ads.ricker_freq(10.0, dt, NSTEP, 1e4)
    xs = np.round(np.linspace(npml + 2, NX - npml, nshots)).astype(int)
    srcs = [ads.AcousticSource([int(x)], [3], w.reshape(-1, 1)) for x in xs]
    rcvs = [ads.AcousticReceiver(np.arange(2, NX + 2), np.full(NX, 3)) for _ in xs]
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=dtype)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    zz = np.linspace(0, 1, NY + 2)[None, :]
    v0 = 1800.0 + 1500.0 * zz + 0.0 * np.arange(NX + 2)[:, None]
    vtrue = v0 + 200.0 * np.exp(-((np.arange(NX + 2)[:, None] - NX / 2) ** 2 +
                                  (np.arange(NY + 2)[None, :] - NY / 2) ** 2) / (0.02 * NX * NY))
    return p, plan, srcv, torch.as_tensor(v0, dtype=dtype, device="cuda"), \
        torch.as_tensor(vtrue, dtype=dtype, device="cuda")


def test_nnfwi_gradient_matches_finite_differences(ads, cuda):
    torch.manual_seed(1)
    dtype = torch.float64
    NX, NY = 40, 36
    p, plan, srcv, v0, vtrue = make(ads, NX, NY, 150, 3, dtype)
    K = plan.choose_interval()
    obs = plan.forward(vtrue, srcv, K).clone()
    G = ads.Generator(num_layer=2, w0=math.ceil((NX + 2) / 4) + 1, h0=math.ceil((NY + 2) / 4),
                      vmin=-1.5, vmax=1.5).to(device="cuda", dtype=dtype)
    z = torch.rand(1, 8, dtype=dtype, device="cuda")

    def loss_of():
        vp = ads.add_initial_model(G(z) * 100.0, v0)
        return torch.sum((ads.acoustic_traces(plan, vp.contiguous(), srcv, K) - obs) ** 2)

    L = loss_of()
    G.zero_grad()
    L.backward()
    params = [q for q in G.parameters()]
    rng = np.random.default_rng(7)
    # 20 random weights among those whose derivative stands out of the finite-difference
    # rounding noise (the top tenth by |dL/dw| of every layer)
    pool = []
    for qi, q in enumerate(params):
        a = q.grad.abs().reshape(-1)
        thr = torch.quantile(a, 0.9)
        pool += [(qi, int(k)) for k in torch.nonzero(a >= thr).reshape(-1).tolist()]
    checked = 0
    for pick in rng.choice(len(pool), 20, replace=False):
        qi, flat = pool[pick]
        q = params[qi]
        idx = tuple(int(x) for x in np.unravel_index(flat, tuple(q.shape)))
        g = float(q.grad[idx])
        eps = 1e-6 * max(1.0, abs(float(q.data[idx])))
        with torch.no_grad():
            q.data[idx] += eps
            Lp = float(loss_of())
            q.data[idx] -= 2 * eps
            Lm = float(loss_of())
            q.data[idx] += eps
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - g) <= 1e-5 * max(abs(fd), abs(g)), (qi, idx, fd, g)
        checked += 1
    assert checked == 20


def test_nnfwi_step_at_400x800_24_shots(ads, cuda):
    """One NN-FWI gradient at the 400 x 800 grid, 24 shots, 2500 steps (FP32), through the
    reference generator of NNFWI.jl (num_layer 4, vmin/vmax +-1.5, scale 1000 m/s), followed
    by an Adam step that lowers the misfit."""
    import time
    torch.manual_seed(2)
    dtype = torch.float32
    NX, NY = 400, 800
    p, plan, srcv, v0, vtrue = make(ads, NX, NY, 2500, 24, dtype, npml=20)
    K = plan.choose_interval()
    obs = plan.forward(vtrue, srcv, K).clone()
    G = ads.Generator(num_layer=4, w0=math.ceil((NX + 2) / 16), h0=math.ceil((NY + 2) / 16),
                      vmin=-1.5, vmax=1.5).to(device="cuda", dtype=dtype)
    z = torch.rand(1, 8, dtype=dtype, device="cuda")
    opt = torch.optim.Adam(G.parameters(), lr=1e-3)
    losses = []
    for it in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vp = ads.add_initial_model(G(z) * 1000.0 * 0.05, v0)
        L = torch.sum((ads.acoustic_traces(plan, vp.contiguous(), srcv, K) - obs) ** 2)
        opt.zero_grad()
        L.backward()
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - t0
        gn = sum(float(q.grad.norm()) ** 2 for q in G.parameters()) ** 0.5
        print(f"NN-FWI 400x800x24x2500: loss {float(L):.6e} grad-norm {gn:.3e} "
              f"{dt_ * 1e3:.1f} ms ({24 * NX * NY * 2500 / dt_ / 1e9:.1f} Gcell-steps/s)", flush=True)
        assert math.isfinite(float(L)) and math.isfinite(gn) and gn > 0
        losses.append(float(L))
        opt.step()
    assert min(losses[1:]) < losses[0]


def _c5(ads, dtype):
    """The network-parametrised configuration: 400 x 800 grid, h = 10 m, dt = 1 ms, 2500
    steps, 20-cell PML (chosen), 24 shots of a 10 Hz Ricker at depth 3 evenly spaced along the
    surface, 400 receivers each; velocity = 1500 + 3000 sigmoid(G(z)) with G the 4-layer
    upsampling CNN (latent 8 x 16 x 16 from a standard normal, seed 5; 3x3 convolutions with
    64/32/16/1 channels, LeakyReLU 0.1).  Observed data come from a layered model with a
    Gaussian anomaly."""
    NX, NY, NSTEP, nshots, npml = 400, 800, 2500, 24, 20
    p = ads.AcousticPropagatorParams(NX=NX, NY=NY, NSTEP=NSTEP, DELTAX=10.0, DELTAY=10.0,
                                     DELTAT=1e-3, NPOINTS_PML=npml, vp_ref=3000.0)
    ads.check_cfl(p, 4500.0)
    w = ads.ricker_freq(10.0, p.DELTAT, NSTEP, 1e4)
    xs = np.round(np.linspace(npml + 2, NX + 1 - npml, nshots)).astype(int)
    srcs = [ads.AcousticSource([int(x)], [3], w.reshape(-1, 1)) for x in xs]
    rcvs = [ads.AcousticReceiver(np.arange(2, NX + 2), np.full(NX, 3)) for _ in xs]
    plan = ads.AcousticPlan(p, srcs, rcvs, dtype=dtype)
    srcv = plan.pack_srcv([s.srcv for s in srcs])
    jj = np.arange(NY + 2)[None, :]
    ii = np.arange(NX + 2)[:, None]
    vtrue = 1800.0 + 2200.0 * (jj / (NY + 1)) + 0.0 * ii
    vtrue = vtrue + 300.0 * np.exp(-((ii - NX / 2) ** 2 + (jj - NY / 2) ** 2) / (2 * 40.0 ** 2))
    G = ads.UpsamplingGenerator((NX + 2, NY + 2), dtype=dtype)
    return p, plan, srcv, torch.as_tensor(vtrue, dtype=dtype, device="cuda"), G


@pytest.mark.timeout(600)
def test_upsampling_generator_configuration(ads, cuda):
    """FP64: dL/dweights through the HIP adjoint against central finite differences of the
    misfit on 20 weights drawn uniformly at random over all the network's weights (relative
    1e-4, or 1e-6 of the largest of the 20 where a weight's gradient is tiny).  FP32: one
    NN-FWI iteration (generator, 24-shot forward, misfit, adjoint, backpropagation into the
    weights) timed at full size."""
    import time
    from adseismic_amd.acoustic import acoustic_traces

    # FP64 finite differences
    p, plan, srcv, vtrue, G = _c5(ads, torch.float64)
    K = plan.choose_interval()
    obs = plan.forward(vtrue, srcv, K).clone()

    def misfit():
        return torch.sum((acoustic_traces(plan, G(), srcv, K) - obs) ** 2)

    G.zero_grad()
    L = misfit()
    L.backward()
    params = [q for q in G.parameters()]
    sizes = [q.numel() for q in params]
    rng = np.random.default_rng(20)
    picks = sorted(rng.choice(sum(sizes), size=20, replace=False))
    offs = np.cumsum([0] + sizes)
    an, fd = [], []
    with torch.no_grad():
        for k in picks:
            t = int(np.searchsorted(offs, k, side="right") - 1)
            q, idx = params[t].view(-1), int(k - offs[t])
            an.append(float(params[t].grad.view(-1)[idx]))
            w0 = float(q[idx])
            eps = 1e-4 * max(abs(w0), 1e-2)
            q[idx] = w0 + eps
            lp = float(misfit())
            q[idx] = w0 - eps
            lm = float(misfit())
            q[idx] = w0
            fd.append((lp - lm) / (2 * eps))
    an, fd = np.array(an), np.array(fd)
    print(f"C5 FP64: misfit {L.item():.6e}; 20 random weights: max |fd - g| / max|g| "
          f"{np.max(np.abs(fd - an)) / np.max(np.abs(an)):.2e}", flush=True)
    np.testing.assert_allclose(fd, an, rtol=1e-4, atol=1e-6 * np.max(np.abs(an)))
    plan._ws = None
    torch.cuda.empty_cache()
    # FP32 iteration at full size
    p, plan, srcv, vtrue, G = _c5(ads, torch.float32)
    K = plan.choose_interval()
    obs = plan.forward(vtrue, srcv, K).clone()
    for it in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        G.zero_grad()
        L = torch.sum((acoustic_traces(plan, G(), srcv, K) - obs) ** 2)
        L.backward()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    cells = p.NX * p.NY * p.NSTEP * 24
    gn = float(sum(torch.sum(q.grad.double() ** 2) for q in G.parameters()) ** 0.5)
    print(f"C5 FP32 NN-FWI iteration: {el * 1e3:.1f} ms ({cells / el / 1e9:.1f} Gcell-steps/s "
          f"incl. the network), misfit {L.item():.6e}, |dL/dw| {gn:.3e}", flush=True)
    assert math.isfinite(L.item()) and gn > 0
