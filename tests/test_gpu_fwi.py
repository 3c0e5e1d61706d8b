"""FWI driver on the Marmousi2 example (examples/nn_fwi/FWI_inversion.jl): MAT loaders ->
batched plan -> L-BFGS on the masked velocity, plus the drop-in driver path
(compute_loss_and_grads_GPU through the plain solver) against the batched gradient."""
import numpy as np
import pytest
import torch

from matfixture import marmousi_fixture, write_mat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def marm(tmp_path_factory):
    fx = marmousi_fixture()
    d = tmp_path_factory.mktemp("marm")
    return fx, write_mat(str(d / "true.mat"), fx, fx["vp_true"]), \
        write_mat(str(d / "smooth.mat"), fx, fx["vp_smooth"], mask=fx["mask"])


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def observed(ads, p, src, rcv, vp, dtype):
    plan = ads.AcousticPlan(p, src, rcv, dtype=dtype)
    tr = plan.forward(torch.as_tensor(vp, dtype=dtype, device="cuda"),
                      plan.pack_srcv([s.srcv for s in src]), plan.choose_interval())
    return [t.cpu().numpy() for t in plan.unpack_traces(tr)]


def test_marmousi_lbfgs_reduces_misfit(ads, cuda, marm):
    fx, ftrue, fsmooth = marm
    p = ads.load_params(fsmooth, vp_ref=1e3)            # FWI_inversion.jl:34-36
    src = ads.load_acoustic_source(fsmooth)
    rcv = ads.load_acoustic_receiver(fsmooth)
    vp0 = ads.matread(fsmooth)["vp"]
    mask = ads.matread(fsmooth)["mask"]
    ads.check_cfl(p, float(fx["vp_true"].max()))
    obs = observed(ads, p, src, rcv, ads.matread(ftrue)["vp"], torch.float32)
    prob = ads.AcousticFWI(p, src, rcv, obs, dtype=torch.float32, mask=mask, c0=vp0)
    f0, g0 = prob(vp0.reshape(-1))
    assert np.isfinite(f0) and f0 > 0
    g0 = g0.reshape(vp0.shape)
    assert np.all(g0[mask == 0] == 0) and np.abs(g0[mask == 1]).max() > 0  # masked update
    seen = []
    losses = ads.LBFGS_(prob, vp0.reshape(-1), maxiter=6,
                        callback=lambda x, it, loss: seen.append((it, loss)))
    print("L-BFGS losses", f0, losses, "evaluations", prob.evaluations, flush=True)
    assert [it for it, _ in seen] == list(range(1, len(losses) + 1))
    assert len(losses) >= 3
    assert all(b <= a for a, b in zip([f0] + losses, losses))  # accepted steps never increase L
    assert losses[-1] < 0.8 * f0
    x = ads.LBFGS_.x.reshape(vp0.shape)
    assert np.array_equal(x[mask == 0], vp0[mask == 0]) or np.allclose(x[mask == 0], vp0[mask == 0])


def test_adam_runs_and_descends(ads, cuda, marm):
    fx, ftrue, fsmooth = marm
    p = ads.load_params(fsmooth, vp_ref=1e3)
    src = ads.load_acoustic_source(fsmooth)[:2]
    rcv = ads.load_acoustic_receiver(fsmooth)[:2]
    vp0 = ads.matread(fsmooth)["vp"]
    obs = observed(ads, p, src, rcv, ads.matread(ftrue)["vp"], torch.float32)
    prob = ads.AcousticFWI(p, src, rcv, obs, dtype=torch.float32)
    x, losses = ads.Adam_(prob, vp0.reshape(-1), lr=5.0, maxiter=4)
    print("Adam losses", losses, flush=True)
    assert len(losses) == 4 and all(np.isfinite(losses)) and min(losses[1:]) < losses[0]


def test_dropin_driver_matches_batched_gradient(ads, cuda, marm):
    """compute_loss_and_grads (per-shot plain solver + autograd, FP64) against the batched
    plan's gradient (FP64), two shots, PropagatorKernel 0 (the default)."""
    fx, ftrue, fsmooth = marm
    ap = ads.load_acoustic_model(fsmooth, inv_vp=True, vp_ref=1e3)
    p = ap.param
    src = ads.load_acoustic_source(fsmooth)[:2]
    rcv = ads.load_acoustic_receiver(fsmooth)[:2]
    obs = observed(ads, p, src, rcv, ads.matread(ftrue)["vp"], torch.float64)
    loss, g = ads.compute_loss_and_grads(ap, src, rcv, obs, ap.variables[0])
    prob = ads.AcousticFWI(p, src, rcv, obs, dtype=torch.float64)
    vp0 = ads.matread(fsmooth)["vp"]
    f, gb = prob(vp0.reshape(-1))
    assert loss.item() == pytest.approx(f, rel=1e-11)
    assert rel(g.cpu().numpy().reshape(-1), gb) < 1e-10
