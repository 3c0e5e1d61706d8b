"""FWI driver: the misfit and gradient over all shots, and the optimiser loops around it.

Mirrors ``compute_loss_and_grads_GPU`` (src/Utils.jl:300-332), ``LBFGS!`` (src/Optim.jl:135-193,
Optim.jl's L-BFGS behind ADCME) and the (commented-out) Adam alternative of
examples/nn_fwi/FWI_inversion.jl:97-113.

Two objective paths:

- :class:`AcousticFWI` -- the batched path: every shot in one :class:`AcousticPlan`, the misfit
  ``L = sum_s sum (traces_s - observed_s)^2`` (Utils.jl:308, no 1/2) and its gradient from one
  ``ads_acoustic_plan_gradient`` call (checkpointed forward, HIP adjoint, gradient reduction on
  the device);
- :func:`compute_loss_and_grads` -- the drop-in path of the reference's own driver: a propagator
  function per shot (e.g. from :func:`load_acoustic_model`), ``SimulatedObservation_`` and
  torch autograd through the plain solver.

The optimisers see a flat float64 host vector, as Optim.jl does; the model goes to the device
once per evaluation.
"""
from __future__ import annotations

from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .acoustic import AcousticPlan, SimulatedObservation_
from .params import AcousticPropagatorParams, AcousticReceiver, AcousticSource


class AcousticFWI:
    """Batched acoustic FWI objective over shots ``srcs`` / ``rcvs`` with observed traces
    ``observed[s]`` of shape (NSTEP+1, nrcv_s).

    ``mask`` (NX+2, NY+2) freezes cells as in FWI_inversion.jl:38-41: the model is
    ``mask * x + (1 - mask) * c0`` and the gradient w.r.t. ``x`` is ``mask * dL/dc``.
    ``K`` is the checkpoint interval (default: chosen from 60% of free HBM)."""

    def __init__(self, param: AcousticPropagatorParams, srcs: Sequence[AcousticSource],
                 rcvs: Sequence[AcousticReceiver], observed, dtype=torch.float32, mask=None,
                 c0=None, K: Optional[int] = None):
        self.param = param
        self.dtype = dtype
        self.plan = AcousticPlan(param, list(srcs), list(rcvs), dtype=dtype)
        self.srcv = self.plan.pack_srcv([s.srcv for s in srcs])
        self.observed = self.plan.pack_traces(list(observed))
        self.K = int(K) if K else self.plan.choose_interval()
        self.shape = (param.NX + 2, param.NY + 2)
        self.mask = None if mask is None else torch.as_tensor(np.asarray(mask, np.float64), dtype=dtype,
                                                              device="cuda")
        self.c0 = None if c0 is None else torch.as_tensor(np.asarray(c0, np.float64), dtype=dtype,
                                                          device="cuda")
        if (self.mask is None) != (self.c0 is None):
            raise ValueError("mask and c0 go together")
        self.evaluations = 0

    def model(self, x: torch.Tensor) -> torch.Tensor:
        if self.mask is None:
            return x
        return self.mask * x + (1 - self.mask) * self.c0

    def loss_and_grad_device(self, x: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """(loss (float64, device), dL/dx) for a device model x of shape (NX+2, NY+2)."""
        loss, g, _ = self.plan.gradient(self.model(x).contiguous(), self.srcv, self.observed, self.K)
        if self.mask is not None:
            g = g * self.mask
        self.evaluations += 1
        return loss, g

    def __call__(self, x: np.ndarray) -> Tuple[float, np.ndarray]:
        """The optimiser interface: flat float64 host vector -> (loss, flat float64 gradient)."""
        xd = torch.as_tensor(np.asarray(x, np.float64).reshape(self.shape), dtype=self.dtype,
                             device="cuda")
        loss, g = self.loss_and_grad_device(xd)
        return float(loss.item()), g.double().cpu().numpy().reshape(-1)


def _plans_in_batches(param, srcs, rcvs, dtype, max_shots: int, nstep, budget_fraction: float):
    """Plans of at most ``max_shots`` consecutive shots, one after another, sharing one device
    workspace (grown when a batch needs more), so that released workspaces do not pin torch's
    cache.  Yields (shot indices, plan, K)."""
    n = len(srcs)
    ws = None
    for b0 in range(0, n, int(max_shots)):
        idx = list(range(b0, min(n, b0 + int(max_shots))))
        plan = AcousticPlan(param, [srcs[i] for i in idx], [rcvs[i] for i in idx], dtype=dtype,
                            nstep=None if nstep is None else [int(nstep[i]) for i in idx])
        held = 0 if ws is None else ws.numel()
        K = plan.choose_interval(int((_lib.available_device_bytes() + held) * budget_fraction))
        need = plan.workspace_size(K)
        if ws is None or ws.numel() < need:
            ws = None
            torch.cuda.empty_cache()
            ws = _lib.device_bytes(need)
        plan._ws = ws
        yield idx, plan, K
        plan._ws = None
        del plan


def acoustic_traces_in_batches(param: AcousticPropagatorParams, srcs: Sequence[AcousticSource],
                               rcvs: Sequence[AcousticReceiver], c: torch.Tensor, max_shots: int,
                               nstep=None, budget_fraction: float = 0.6) -> List[torch.Tensor]:
    """Traces of every shot, (NSTEP_s+1, nrcv_s) each, through plans of at most ``max_shots``
    shots (see :func:`acoustic_gradient_in_batches`)."""
    out: List[torch.Tensor] = []
    for idx, plan, K in _plans_in_batches(param, srcs, rcvs, c.dtype, max_shots, nstep, budget_fraction):
        tr = plan.forward(c, plan.pack_srcv([srcs[i].srcv for i in idx]), K)
        out += [t.clone() for t in plan.unpack_traces(tr)]
    return out


def acoustic_gradient_in_batches(param: AcousticPropagatorParams, srcs: Sequence[AcousticSource],
                                 rcvs: Sequence[AcousticReceiver], observed, c: torch.Tensor,
                                 max_shots: int, nstep=None, budget_fraction: float = 0.6):
    """Misfit and gradient over a shot set too large for one plan's workspace: the shots go
    through plans of at most ``max_shots`` consecutive shots each (one checkpointed batched
    gradient per plan, interval chosen from ``budget_fraction`` of the available HBM, one
    shared workspace), and the losses and gradients are summed on the device in shot-batch
    order (the compute_loss_and_grads_GPU reduction, Utils.jl:325-331, on one GPU).  ``c`` is
    the velocity (device tensor, the plans' dtype); ``nstep`` the per-shot record lengths
    (default NSTEP)."""
    loss, grad = None, None
    for idx, plan, K in _plans_in_batches(param, srcs, rcvs, c.dtype, max_shots, nstep, budget_fraction):
        srcv = plan.pack_srcv([srcs[i].srcv for i in idx])
        obs = plan.pack_traces([observed[i] for i in idx])
        l, g, _ = plan.gradient(c, srcv, obs, K)
        loss = l if loss is None else loss + l
        grad = g if grad is None else grad + g
    return loss, grad


def compute_loss_and_grads(model: Callable, src: Sequence[AcousticSource],
                           rcv_sim: Sequence[AcousticReceiver], Rs, vs, aux_loss=None):
    """compute_loss_and_grads_GPU(model, src, rcv_sim, Rs, vs; aux_loss), Utils.jl:300-332, on
    one GPU: ``loss = sum_i sum (SimulatedObservation!(model(src[i]), rcv_sim[i]) - Rs[i])^2``
    (+ ``aux_loss``) and its gradient w.r.t. the tensor(s) ``vs``.  Returns (loss, grads) with
    grads a tensor for a single ``vs`` and a list otherwise."""
    single = isinstance(vs, torch.Tensor)
    vlist = [vs] if single else list(vs)
    loss = None
    for i in range(len(Rs)):
        r = SimulatedObservation_(model(src[i]), rcv_sim[i])
        obs = torch.as_tensor(np.asarray(Rs[i]), dtype=r.dtype, device=r.device)
        term = torch.sum((r - obs) ** 2)
        loss = term if loss is None else loss + term
    if aux_loss is not None:
        loss = loss + aux_loss
    grads = torch.autograd.grad(loss, vlist, allow_unused=True)
    return loss.detach(), (grads[0] if single else list(grads))


def LBFGS_(fg: Callable[[np.ndarray], Tuple[float, np.ndarray]], x0, callback=None,
           maxiter: int = 15000, m: int = 10, gtol: float = 1e-12, ftol: float = 1e-12,
           bounds=None) -> List[float]:
    """LBFGS!(sess, loss, grads, vars; callback, kwargs...), src/Optim.jl:135-193.

    ``fg(x) -> (loss, grad)`` on a flat float64 vector.  Runs limited-memory BFGS (scipy's
    L-BFGS-B, memory ``m``; ``bounds`` optional) and returns the loss after each iteration;
    ``callback(x, iter, loss)`` is called with the same values as the reference's callback.
    The final iterate is ``LBFGS_.x``."""
    from scipy.optimize import minimize
    x0 = np.asarray(x0, np.float64).reshape(-1)
    losses: List[float] = []
    cache = {}

    def fun(x):
        f, g = fg(x)
        cache["x"], cache["f"] = x.copy(), f
        return f, g

    def cb(xk):
        f = cache["f"] if np.array_equal(cache.get("x"), xk) else fg(xk)[0]
        losses.append(float(f))
        if callback is not None:
            callback(xk, len(losses), float(f))

    res = minimize(fun, x0, jac=True, method="L-BFGS-B", callback=cb, bounds=bounds,
                   options=dict(maxiter=int(maxiter), maxcor=int(m), gtol=gtol, ftol=ftol))
    LBFGS_.x, LBFGS_.result = res.x, res
    return losses


def Adam_(fg: Callable[[np.ndarray], Tuple[float, np.ndarray]], x0, lr: float = 1.0,
          maxiter: int = 100, betas=(0.9, 0.999), eps: float = 1e-8, cosine_decay: bool = True,
          callback=None) -> Tuple[np.ndarray, List[float]]:
    """The Adam alternative sketched in FWI_inversion.jl:97-113: learning rate ``lr`` with
    cosine decay over ``maxiter`` steps (tf.train.cosine_decay), one gradient per step.
    Returns (x, losses)."""
    x = torch.tensor(np.asarray(x0, np.float64).reshape(-1), requires_grad=True)
    opt = torch.optim.Adam([x], lr=lr, betas=betas, eps=eps)
    sched = (torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(int(maxiter), 1), eta_min=0.0)
             if cosine_decay else None)
    losses = []
    for it in range(int(maxiter)):
        f, g = fg(x.detach().numpy())
        losses.append(float(f))
        opt.zero_grad()
        x.grad = torch.as_tensor(g, dtype=torch.float64)
        opt.step()
        if sched is not None:
            sched.step()
        if callback is not None:
            callback(x.detach().numpy(), it + 1, float(f))
    return x.detach().numpy(), losses
