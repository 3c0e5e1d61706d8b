"""Elastic solver API: the ADSeismic.jl call surface over the HIP library.

Mirrors src/Core.jl (ElasticPropagatorSolver, SimulatedObservation!(::ElasticPropagator),
compute_PML_Params(::ElasticPropagatorParams)) and adds the batched multi-shot gradient path
w.r.t. (rho, lambda, mu).  Every computation runs in libadseismic_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .acoustic import _dcode, _dev, _stream, _vp
from .params import (ElasticPropagatorParams, ElasticReceiver, ElasticSource, Geometry,
                     MPIElasticPropagatorParams, MPIElasticReceiver, MPIElasticSource)

__all__ = ["compute_PML_Params", "ElasticPlan", "ElasticPropagator", "ElasticPropagatorSolver",
           "SimulatedObservation_elastic", "elastic_traces", "MPIElasticPropagator",
           "MPIElasticPropagatorSolver", "MPISimulatedObservation_"]


def compute_PML_Params(param: ElasticPropagatorParams):
    """compute_PML_Params(param), src/Core.jl:231-407 -> ax, bx, kx, ay, by, ky, each of shape
    (2, n): row 0 on the full grid, row 1 on the half grid (Core.jl:397-402)."""
    pc = param.to_c()
    NX, NY = param.NX, param.NY
    outs = [np.zeros(2 * NX) for _ in range(3)] + [np.zeros(2 * NY) for _ in range(3)]
    _lib.check(_lib.load().ads_elastic_pml_params(C.byref(pc), *[o.ctypes.data for o in outs]))
    return tuple(o.reshape(2, -1) for o in outs)


class ElasticPlan:
    """A batch of elastic shots over one model (rho, lambda, mu): wraps ads_elastic_plan."""

    def __init__(self, param: ElasticPropagatorParams, srcs: Sequence[ElasticSource],
                 rcvs: Sequence[ElasticReceiver], dtype=torch.float32, nstep=None):
        self.lib = _lib.load()
        self.param = param
        self.dtype = dtype
        self.nshots = len(srcs)
        self.geom = Geometry(srcs, rcvs, nstep=nstep, elastic=True)
        self.nstep = (np.full(self.nshots, param.NSTEP, np.int64) if nstep is None
                      else np.asarray(nstep, np.int64))
        self._pc = param.to_c()
        # padded array size: (NX+2) x (NY+2), or (n+4) x (n+4) for one MPIElastic block
        pad = 4 if self._pc.subdomain else 2
        self.gx, self.gy = int(self._pc.NX) + pad, int(self._pc.NY) + pad
        gc = self.geom.to_c()
        h = C.c_void_p()
        _lib.check(self.lib.ads_elastic_plan_create(C.byref(self._pc), C.byref(gc), _dcode(dtype),
                                                    C.byref(h)))
        self.h = h
        se, te = C.c_int64(), C.c_int64()
        so = np.zeros(self.nshots, np.int64)
        to = np.zeros(self.nshots, np.int64)
        _lib.check(self.lib.ads_elastic_plan_sizes(h, C.byref(se), C.byref(te),
                                                   so.ctypes.data_as(_lib.i64p),
                                                   to.ctypes.data_as(_lib.i64p)))
        self.srcv_elems, self.trace_elems = se.value, te.value
        self.srcv_off, self.trace_off = so, to
        self._ws = None
        self._ws_K = None
        self._tok = None  # workspace token of this plan's last forward (see _lib.mark_workspace)
        self._scratch = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.ads_elastic_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def pack_srcv(self, srcvs) -> torch.Tensor:
        parts = []
        for s, v in enumerate(srcvs):
            t = _dev(v, self.dtype).reshape(-1)
            want = int(self.nstep[s] * self.geom.nsrc[s])
            if t.numel() != want:
                raise ValueError(f"shot {s}: srcv has {t.numel()} values, expected {want}")
            parts.append(t)
        return torch.cat(parts).contiguous()

    def unpack_traces(self, flat: torch.Tensor) -> List[torch.Tensor]:
        """Per-shot traces of shape (nrcv_s, nstep_s + 1) (GetReceive.cc:65 layout)."""
        out = []
        for s in range(self.nshots):
            nr, nt = int(self.geom.nrcv[s]), int(self.nstep[s] + 1)
            o = int(self.trace_off[s])
            out.append(flat[o:o + nr * nt].view(nr, nt))
        return out

    def pack_traces(self, traces) -> torch.Tensor:
        return torch.cat([_dev(t, self.dtype).reshape(-1) for t in traces]).contiguous()

    def workspace_size(self, K: int) -> int:
        return int(self.lib.ads_elastic_plan_workspace_size(self.h, int(K)))

    def set_adjoint_kernels(self, streaming: int = 2, rows: int = 8, shots: int = 4) -> None:
        """Adjoint kernels (ads_elastic_plan_set_adjoint_kernels): 2 full-grid streaming, 1
        streaming rectangle with the frame passes, 0 LDS-tiled rectangle (True / False mean 1 /
        0); changes the workspace layout, so call it before the forward."""
        _lib.check(self.lib.ads_elastic_plan_set_adjoint_kernels(self.h, int(streaming),
                                                                 int(rows), int(shots)))
        self._ws_K = None

    def set_forward_kernels(self, fused: int = 1, rows: int = 0) -> None:
        """Forward kernels (ads_elastic_plan_set_forward_kernels): 1 one fused launch per level
        walking `rows` rows per wavefront (0: chosen from the grid and the active shots), 0 the
        stress and velocity phases as two launches."""
        _lib.check(self.lib.ads_elastic_plan_set_forward_kernels(self.h, int(fused), int(rows)))

    def set_snapshots(self, snapshots: int) -> None:
        """Snapshot slots of the reversal schedule: < 0 one per segment start (one-level),
        otherwise the binomial schedule with that many slots."""
        _lib.check(self.lib.ads_elastic_plan_set_snapshots(self.h, int(snapshots)))
        self._ws_K = None

    @property
    def snapshots(self) -> int:
        return int(self.lib.ads_elastic_plan_snapshots(self.h))

    def choose_schedule(self, budget_bytes: Optional[int] = None):
        """(K, S) with the least modelled time whose workspace fits the budget (default 90% of
        the device memory a workspace can use); the plan keeps S."""
        if budget_bytes is None:
            budget_bytes = int(_lib.available_device_bytes() * 0.9)
        K, S = _lib.i64(0), _lib.i64(0)
        _lib.check(self.lib.ads_elastic_plan_choose_schedule(self.h, int(budget_bytes),
                                                             C.byref(K), C.byref(S)))
        self._ws_K = None
        return int(K.value), int(S.value)

    def choose_interval(self, budget_bytes: Optional[int] = None) -> int:
        if budget_bytes is None:
            budget_bytes = int(_lib.available_device_bytes() * 0.6)
        K = int(self.lib.ads_elastic_plan_choose_interval(self.h, int(budget_bytes)))
        if K < 1:
            raise MemoryError("no checkpoint interval fits the memory budget")
        return K

    def workspace(self, K: int) -> torch.Tensor:
        need = self.workspace_size(K)
        if self._ws is None or self._ws.numel() < need:
            self._ws = None
            self._ws = _lib.device_bytes(need)
        self._ws_K = K
        return self._ws

    def _models(self, rho, lam, mu):
        return [_dev(x, self.dtype) for x in (rho, lam, mu)]

    def forward(self, rho, lam, mu, srcv, K: int) -> torch.Tensor:
        r, l, m = self._models(rho, lam, mu)
        tr = torch.empty(self.trace_elems, dtype=self.dtype, device="cuda")
        ws = self.workspace(K)
        _lib.check(self.lib.ads_elastic_plan_forward(self.h, _vp(r), _vp(l), _vp(m), _vp(srcv),
                                                     _vp(tr), _vp(ws), ws.numel(), int(K), _stream()))
        self._tok = _lib.mark_workspace(ws)
        return tr

    def holds_forward(self, token, K: int) -> bool:
        """True while the workspace still holds the state of the forward that returned
        ``token`` (``self._tok`` right after it)."""
        return self._ws_K == K and _lib.workspace_holds(self._ws, token)

    def backward(self, rho, lam, mu, srcv, grad_traces, K: int):
        if self._ws is None or self._ws_K != K:
            raise RuntimeError("backward needs the workspace of a forward with the same K")
        r, l, m = self._models(rho, lam, mu)
        g = [torch.empty_like(r) for _ in range(3)]
        gt = _dev(grad_traces, self.dtype)
        _lib.check(self.lib.ads_elastic_plan_backward(self.h, _vp(r), _vp(l), _vp(m), _vp(srcv),
                                                      _vp(gt), *[_vp(x) for x in g], _vp(self._ws),
                                                      self._ws.numel(), int(K), _stream()))
        _lib.mark_workspace(self._ws)  # the sweep recomputed segments: the forward state is spent
        return tuple(g)

    def gradient(self, rho, lam, mu, srcv, observed, K: int, want_traces: bool = False):
        """(loss (device float64), (d/drho, d/dlambda, d/dmu), traces or None)."""
        r, l, m = self._models(rho, lam, mu)
        obs = _dev(observed, self.dtype)
        g = [torch.empty_like(r) for _ in range(3)]
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        tr = torch.empty(self.trace_elems, dtype=self.dtype, device="cuda") if want_traces else None
        ws = self.workspace(K)
        _lib.check(self.lib.ads_elastic_plan_gradient(self.h, _vp(r), _vp(l), _vp(m), _vp(srcv),
                                                      _vp(obs), _vp(loss), *[_vp(x) for x in g],
                                                      _vp(tr), _vp(ws), ws.numel(), int(K),
                                                      _stream()))
        _lib.mark_workspace(ws)
        return loss, tuple(g), tr

    def forward_history(self, rho, lam, mu, srcv):
        r, l, m = self._models(rho, lam, mu)
        L = self.param.NSTEP + 1
        hist = torch.empty((self.nshots, 5, L, self.gx, self.gy), dtype=self.dtype, device="cuda")
        need = int(self.lib.ads_elastic_plan_history_scratch_size(self.h))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = _lib.device_bytes(need)
        _lib.check(self.lib.ads_elastic_forward_history(self.h, _vp(r), _vp(l), _vp(m), _vp(srcv),
                                                        _vp(hist), _vp(self._scratch), _stream()))
        return hist

    def backward_dense(self, rho, lam, mu, srcv, grad_hist, K: Optional[int] = None):
        r, l, m = self._models(rho, lam, mu)
        if K is None:
            K = self.choose_interval()
        ws = self.workspace(K)
        g = [torch.empty_like(r) for _ in range(3)]
        gh = grad_hist.contiguous()
        _lib.check(self.lib.ads_elastic_plan_backward_dense(self.h, _vp(r), _vp(l), _vp(m),
                                                            _vp(srcv), _vp(gh),
                                                            *[_vp(x) for x in g], _vp(ws),
                                                            ws.numel(), int(K), _stream()))
        _lib.mark_workspace(ws)
        return tuple(g)


class _ElTracesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, lam, mu, plan, srcv, K):
        tr = plan.forward(rho, lam, mu, srcv, K)
        ctx.plan, ctx.K, ctx.tok = plan, K, plan._tok
        ctx.save_for_backward(rho, lam, mu, srcv)
        return tr

    @staticmethod
    def backward(ctx, g):
        rho, lam, mu, srcv = ctx.saved_tensors
        if not ctx.plan.holds_forward(ctx.tok, ctx.K):
            ctx.plan.forward(rho, lam, mu, srcv, ctx.K)  # the workspace was used since
        gr, gl, gm = ctx.plan.backward(rho, lam, mu, srcv, g.contiguous(), ctx.K)
        return gr, gl, gm, None, None, None


def elastic_traces(plan: ElasticPlan, rho, lam, mu, srcv, K: int):
    """Differentiable batched elastic traces (flat plan layout)."""
    return _ElTracesFn.apply(rho, lam, mu, plan, srcv, K)


class _ElHistoryFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, lam, mu, plan, srcv):
        hist = plan.forward_history(rho, lam, mu, srcv)
        ctx.plan = plan
        ctx.save_for_backward(rho, lam, mu, srcv)
        return hist

    @staticmethod
    def backward(ctx, gh):
        rho, lam, mu, srcv = ctx.saved_tensors
        gr, gl, gm = ctx.plan.backward_dense(rho, lam, mu, srcv, gh)
        return gr, gl, gm, None, None


@dataclass
class ElasticPropagator:
    """ElasticPropagator, src/Struct.jl:64-73 (histories of shape (NSTEP+1, NX+2, NY+2))."""
    param: ElasticPropagatorParams
    src: ElasticSource
    vx: Any
    vy: Any
    sigmaxx: Any
    sigmayy: Any
    sigmaxy: Any


def ElasticPropagatorSolver(param: ElasticPropagatorParams, src: ElasticSource, ρ, λ, μ) -> ElasticPropagator:
    """ElasticPropagatorSolver(param, src, rho, lambda, mu), src/Core.jl:31-94.

    Histories are differentiable w.r.t. (rho, lambda, mu) through the hand-written adjoint
    (dense cotangents of the histories, ads_elastic_plan_backward_dense)."""
    dtype = ρ.dtype if isinstance(ρ, torch.Tensor) else torch.float64
    models = [m if isinstance(m, torch.Tensor) and m.is_cuda else _dev(m, dtype) for m in (ρ, λ, μ)]
    models = [m.reshape(param.NX + 2, param.NY + 2) for m in models]
    nsrc = len(np.asarray(src.srci).reshape(-1))
    srcv = _dev(src.srcv, dtype)
    if srcv.dim() == 1:
        srcv = srcv.reshape(-1, 1)
    if tuple(srcv.shape) != (param.NSTEP, nsrc):
        raise ValueError(f"srcv must be (NSTEP, nsrc) = ({param.NSTEP}, {nsrc})")
    plan = ElasticPlan(param, [src], [ElasticReceiver(np.zeros(0, np.int64), np.zeros(0, np.int64),
                                                      np.zeros(0, np.int64))], dtype=dtype)
    hist = _ElHistoryFn.apply(*models, plan, srcv.contiguous())[0]
    return ElasticPropagator(param, src, hist[0], hist[1], hist[2], hist[3], hist[4])


def SimulatedObservation_elastic(ep: ElasticPropagator, rcv: ElasticReceiver):
    """SimulatedObservation!(ep, rcv), src/Core.jl:701-712 (GetReceive): rcvv of shape
    (nrcv, NSTEP+1), rcvv[r, n] = field_type_r[n][(rcvi-1)(NY+2) + rcvj-1]."""
    NY = ep.param.NY
    fields = torch.stack([ep.vx, ep.vy, ep.sigmaxx, ep.sigmayy, ep.sigmaxy]).reshape(5, ep.vx.shape[0], -1)
    idx = (np.asarray(rcv.rcvi, np.int64) - 1) * (NY + 2) + (np.asarray(rcv.rcvj, np.int64) - 1)
    t = np.asarray(rcv.rcvtype, np.int64)
    if len(idx) and (idx.min() < 0 or idx.max() >= fields.shape[2] or t.min() < 0 or t.max() > 4):
        raise IndexError("receiver outside the padded grid or bad type")
    dev = fields.device
    rcv.rcvv = fields[torch.as_tensor(t, device=dev), :, torch.as_tensor(idx, device=dev)]
    return rcv.rcvv


@dataclass
class MPIElasticPropagator:
    """MPIElasticPropagator, MPIElastic.jl:132-142: block histories of shape (NSTEP+1, n, n)."""
    param: MPIElasticPropagatorParams
    src: MPIElasticSource
    vx: Any
    vy: Any
    sigmaxx: Any
    sigmayy: Any
    sigmaxy: Any


def MPIElasticPropagatorSolver(param: MPIElasticPropagatorParams, src: MPIElasticSource, ρ, λ, μ
                               ) -> MPIElasticPropagator:
    """MPIElasticPropagatorSolver(param, src, rho, lambda, mu), MPIElastic.jl:374-466, for one
    block stepped with zero halos: (rho, lambda, mu) are the block's (n+4) x (n+4) arrays; the
    scheme updates every block cell without parameter averaging (MPIElastic.jl:483-620).
    Histories are differentiable through the hand-written adjoint."""
    n = int(param.n)
    dtype = ρ.dtype if isinstance(ρ, torch.Tensor) else torch.float64
    models = [m if isinstance(m, torch.Tensor) and m.is_cuda else _dev(m, dtype) for m in (ρ, λ, μ)]
    models = [m.reshape(n + 4, n + 4) for m in models]
    nsrc = len(src.srci)
    srcv = _dev(src.srcv, dtype).reshape(param.NSTEP, nsrc)
    esrc = ElasticSource(src.srci, src.srcj, src.srctype, srcv)
    none = np.zeros(0, np.int64)
    plan = ElasticPlan(param, [esrc], [ElasticReceiver(none, none, none)], dtype=dtype)
    hist = _ElHistoryFn.apply(*models, plan, srcv.contiguous())[0]
    blk = hist[:, :, 2:n + 2, 2:n + 2]
    return MPIElasticPropagator(param, src, blk[0], blk[1], blk[2], blk[3], blk[4])


def MPISimulatedObservation_(ep: MPIElasticPropagator, rcv: MPIElasticReceiver):
    """MPISimulatedObservation!(ep, rcv), MPIElastic.jl:662-688 (GetReceive over the n x n block
    histories): rcvv of shape (nrcv, NSTEP+1), rcvv[r, n] = field_r[n][(rcvi-1) n + rcvj-1]."""
    n = int(ep.param.n)
    fields = torch.stack([ep.vx, ep.vy, ep.sigmaxx, ep.sigmayy, ep.sigmaxy]).reshape(5, ep.vx.shape[0], -1)
    idx = (np.asarray(rcv.rcvi, np.int64) - 1) * n + (np.asarray(rcv.rcvj, np.int64) - 1)
    t = np.asarray(rcv.rcvtype, np.int64)
    if len(idx) and (idx.min() < 0 or idx.max() >= n * n or t.min() < 0 or t.max() > 4):
        raise IndexError("receiver outside the block or bad type")
    dev = fields.device
    rcv.rcvv = fields[torch.as_tensor(t, device=dev), :, torch.as_tensor(idx, device=dev)]
    return rcv.rcvv
