"""Acoustic solver API: the ADSeismic.jl call surface over the HIP library.

Mirrors src/Core.jl (AcousticPropagatorSolver, SimulatedObservation!, compute_PML_Params!,
acoustic_one_step_customop) and adds the batched multi-shot gradient path.  Fields are torch
tensors on the GPU; every computation runs in libadseismic_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .params import AcousticPropagatorParams, AcousticReceiver, AcousticSource, Geometry

__all__ = [
    "compute_PML_Params_", "AcousticPropagator", "AcousticPropagatorSolver",
    "SimulatedObservation_", "AcousticPlan", "acoustic_one_step", "acoustic_one_step_customop",
    "acoustic_traces", "getid2",
]


def _vp(t) -> Optional[C.c_void_p]:
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dcode(dtype) -> int:
    if dtype == torch.float32:
        return _lib.ADS_F32
    if dtype == torch.float64:
        return _lib.ADS_F64
    raise TypeError(f"unsupported dtype {dtype}: use torch.float32 or torch.float64")


def _dev(x, dtype=None) -> torch.Tensor:
    """A contiguous CUDA tensor (numpy arrays are copied to the GPU, float64 by default)."""
    if not torch.cuda.is_available():
        raise RuntimeError("a GPU is required: the HIP library is the only compute path")
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x), dtype=dtype or torch.float64)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if not t.is_cuda:
        t = t.cuda()
    return t.contiguous()


def getid2(a, b, NX, NY):
    """src/Core.jl:763-771: 1-based flat indices (a-1)(NY+2)+b for a x b (row-major)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    return ((a[:, None] - 1) * (NY + 2) + b[None, :]).reshape(-1)


def compute_PML_Params_(param: AcousticPropagatorParams) -> AcousticPropagatorParams:
    """compute_PML_Params!(param), src/Core.jl:622-655: damping, Sigma_x/Sigma_y grids and the
    stencil index sets.  The profiles come from the library (ads_acoustic_pml_profiles)."""
    lib = _lib.load()
    p = param.to_c()
    sx = np.zeros(param.NX + 2)
    ty = np.zeros(param.NY + 2)
    _lib.check(lib.ads_acoustic_pml_profiles(C.byref(p), sx.ctypes.data, ty.ctypes.data))
    Lx = param.NPOINTS_PML * param.DELTAX
    Ly = param.NPOINTS_PML * param.DELTAY
    param.damping_x = param.vp_ref / Lx * np.log(1 / param.Rcoef)
    param.damping_y = param.vp_ref / Ly * np.log(1 / param.Rcoef)
    param.Σx = np.repeat(sx[:, None], param.NY + 2, axis=1)
    param.Σy = np.repeat(ty[None, :], param.NX + 2, axis=0)
    NX, NY = param.NX, param.NY
    r = lambda lo, hi: np.arange(lo, hi + 1)
    param.IJ = getid2(r(2, NX + 1), r(2, NY + 1), NX, NY)
    param.IJn = getid2(r(2, NX + 1), r(1, NY), NX, NY)
    param.IJp = getid2(r(2, NX + 1), r(3, NY + 2), NX, NY)
    param.IpJ = getid2(r(3, NX + 2), r(2, NY + 1), NX, NY)
    param.IpJp = getid2(r(3, NX + 2), r(3, NY + 2), NX, NY)
    param.IpJn = getid2(r(3, NX + 2), r(1, NY), NX, NY)
    param.InJ = getid2(r(1, NX), r(2, NY + 1), NX, NY)
    param.InJn = getid2(r(1, NX), r(1, NY), NX, NY)
    param.InJp = getid2(r(1, NX), r(3, NY + 2), NX, NY)
    return param


def check_cfl(param: AcousticPropagatorParams, c_max: float) -> None:
    """Raises ADSError(ADS_ERR_CFL) if c_max * dt * sqrt(1/hx^2 + 1/hy^2) > 1."""
    p = param.to_c()
    _lib.check(_lib.load().ads_acoustic_check_cfl(C.byref(p), float(c_max)))


class AcousticPlan:
    """A batch of shots over one model: wraps ads_acoustic_plan (include/adseismic.h).

    Geometry is validated once (1-based padded-grid indices, ADS_ERR_INDEX otherwise).
    ``nstep`` gives per-shot record lengths (uneven records); shots are stepped together.
    """

    def __init__(self, param: AcousticPropagatorParams, srcs: Sequence[AcousticSource],
                 rcvs: Sequence[AcousticReceiver], dtype=torch.float32, nstep=None,
                 square_velocity: bool = True, origin=(0.0, 0.0), global_n=(0, 0)):
        self.lib = _lib.load()
        self.param = param
        self.dtype = dtype
        self.nshots = len(srcs)
        self.geom = Geometry(srcs, rcvs, nstep=nstep)
        self.nstep = (np.full(self.nshots, param.NSTEP, np.int64) if nstep is None
                      else np.asarray(nstep, np.int64))
        self._pc = param.to_c(square_velocity, origin, global_n)
        gc = self.geom.to_c()
        h = C.c_void_p()
        _lib.check(self.lib.ads_acoustic_plan_create(C.byref(self._pc), C.byref(gc), _dcode(dtype),
                                                     C.byref(h)))
        self.h = h
        se, te = C.c_int64(), C.c_int64()
        so = np.zeros(self.nshots, np.int64)
        to = np.zeros(self.nshots, np.int64)
        _lib.check(self.lib.ads_acoustic_plan_sizes(h, C.byref(se), C.byref(te),
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
                self.lib.ads_acoustic_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ---- data packing ----
    def pack_srcv(self, srcvs) -> torch.Tensor:
        """Concatenate per-shot (nstep_s, nsrc_s) wavelets into the plan's flat device layout."""
        parts = []
        for s, v in enumerate(srcvs):
            t = _dev(v, self.dtype).reshape(-1)
            want = int(self.nstep[s] * self.geom.nsrc[s])
            if t.numel() != want:
                raise ValueError(f"shot {s}: srcv has {t.numel()} values, expected "
                                 f"nstep*nsrc = {want}")
            parts.append(t)
        out = torch.cat(parts) if parts else torch.zeros(0, dtype=self.dtype, device="cuda")
        return out.contiguous()

    def unpack_traces(self, flat: torch.Tensor) -> List[torch.Tensor]:
        """Per-shot traces of shape (nstep_s + 1, nrcv_s) (Core.jl:728 layout)."""
        out = []
        for s in range(self.nshots):
            n = int((self.nstep[s] + 1) * self.geom.nrcv[s])
            o = int(self.trace_off[s])
            out.append(flat[o:o + n].view(int(self.nstep[s] + 1), int(self.geom.nrcv[s])))
        return out

    def pack_traces(self, traces) -> torch.Tensor:
        return torch.cat([_dev(t, self.dtype).reshape(-1) for t in traces]).contiguous()

    # ---- workspace ----
    def workspace_size(self, K: int) -> int:
        return int(self.lib.ads_acoustic_plan_workspace_size(self.h, int(K)))

    def choose_interval(self, budget_bytes: Optional[int] = None) -> int:
        if budget_bytes is None:
            budget_bytes = int(_lib.available_device_bytes() * 0.6)
        K = int(self.lib.ads_acoustic_plan_choose_interval(self.h, int(budget_bytes)))
        if K < 1:
            raise MemoryError("no checkpoint interval fits the memory budget")
        return K

    def set_snapshots(self, snapshots: int) -> None:
        """Snapshot slots of the reversal schedule: < 0 one per segment start (one-level),
        otherwise the binomial schedule with that many slots."""
        _lib.check(self.lib.ads_acoustic_plan_set_snapshots(self.h, int(snapshots)))
        self._ws_K = None  # a backward needs a forward made with the same schedule

    @property
    def snapshots(self) -> int:
        return int(self.lib.ads_acoustic_plan_snapshots(self.h))

    def choose_schedule(self, budget_bytes: Optional[int] = None):
        """(K, S) with the least modelled time whose workspace fits the budget (default 90% of
        the device memory a workspace can use); the plan keeps S."""
        if budget_bytes is None:
            budget_bytes = int(_lib.available_device_bytes() * 0.9)
        K, S = _lib.i64(0), _lib.i64(0)
        _lib.check(self.lib.ads_acoustic_plan_choose_schedule(self.h, int(budget_bytes),
                                                              C.byref(K), C.byref(S)))
        self._ws_K = None
        return int(K.value), int(S.value)

    def workspace(self, K: int) -> torch.Tensor:
        need = self.workspace_size(K)
        if self._ws is None or self._ws.numel() < need:
            self._ws = None
            self._ws = _lib.device_bytes(need)
        self._ws_K = K
        return self._ws

    # ---- compute ----
    def forward(self, model: torch.Tensor, srcv: torch.Tensor, K: int) -> torch.Tensor:
        model = _dev(model, self.dtype)
        tr = torch.empty(self.trace_elems, dtype=self.dtype, device="cuda")
        ws = self.workspace(K)
        _lib.check(self.lib.ads_acoustic_plan_forward(self.h, _vp(model), _vp(srcv), _vp(tr),
                                                      _vp(ws), ws.numel(), int(K), _stream()))
        self._tok = _lib.mark_workspace(ws)
        return tr

    def holds_forward(self, token, K: int) -> bool:
        """True while the workspace still holds the state the forward that returned ``token``
        (``self._tok`` right after it) left for its backward."""
        return self._ws_K == K and _lib.workspace_holds(self._ws, token)

    def backward(self, model: torch.Tensor, srcv: torch.Tensor, grad_traces: torch.Tensor,
                 K: int) -> torch.Tensor:
        if self._ws is None or self._ws_K != K:
            raise RuntimeError("backward needs the workspace of a forward with the same K")
        model = _dev(model, self.dtype)
        g = torch.empty_like(model)
        gt = _dev(grad_traces, self.dtype)
        _lib.check(self.lib.ads_acoustic_plan_backward(self.h, _vp(model), _vp(srcv), _vp(gt),
                                                       _vp(g), _vp(self._ws), self._ws.numel(),
                                                       int(K), _stream()))
        # the sweep recomputed segments over the slot ring: the forward's state is spent
        _lib.mark_workspace(self._ws)
        return g

    def gradient(self, model: torch.Tensor, srcv: torch.Tensor, observed: torch.Tensor, K: int,
                 want_traces: bool = False):
        """FWI gradient: (loss tensor (device, float64), d loss / d model, traces or None)."""
        model = _dev(model, self.dtype)
        obs = _dev(observed, self.dtype)
        g = torch.empty_like(model)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        tr = torch.empty(self.trace_elems, dtype=self.dtype, device="cuda") if want_traces else None
        ws = self.workspace(K)
        _lib.check(self.lib.ads_acoustic_plan_gradient(self.h, _vp(model), _vp(srcv), _vp(obs),
                                                       _vp(loss), _vp(g), _vp(tr), _vp(ws),
                                                       ws.numel(), int(K), _stream()))
        _lib.mark_workspace(ws)
        return loss, g, tr

    def autotune(self, model: torch.Tensor, srcv: torch.Tensor, K: int, reps: int = 3):
        """Pick the fastest (rows per wavefront, shots per workgroup) for the forward and adjoint
        kernels on this plan (ads_acoustic_plan_autotune).  Returns the chosen shape."""
        model = _dev(model, self.dtype)
        ws = self.workspace(K)
        out = (C.c_int * 6)()
        _lib.check(self.lib.ads_acoustic_plan_autotune(self.h, _vp(model), _vp(srcv), _vp(ws),
                                                       ws.numel(), int(K), int(reps), out,
                                                       _stream()))
        _lib.mark_workspace(ws)
        self.shape = tuple(out)
        return self.shape

    def set_shape(self, shape):
        if len(shape) != 6:
            raise ValueError("shape is (fwd rows, fwd shots, fwd batch, adj rows, adj shots, adj batch)")
        arr = (C.c_int * 6)(*[int(x) for x in shape])
        _lib.check(self.lib.ads_acoustic_plan_set_shape(self.h, arr))
        self.shape = tuple(shape)

    def set_blocking(self, levels: int = 2):
        """Levels per interior-rectangle pass of the forward sweeps (2 = temporal blocking,
        the default; bit-identical to 1)."""
        _lib.check(self.lib.ads_acoustic_plan_set_blocking(self.h, int(levels)))

    def set_frame_rows(self, rows: int = 0):
        """Rows per item of the two-level PML-frame pass: 4, 8, 16, 32 or 64, or 0 to choose
        per launch from the active shots and the launch's size (bit-identical results)."""
        _lib.check(self.lib.ads_acoustic_plan_set_frame_rows(self.h, int(rows)))

    @property
    def blocking(self) -> int:
        """Levels per pass in effect: 2 only when no source lies in the interior rectangle."""
        return int(self.lib.ads_acoustic_plan_blocking(self.h))

    def set_graph(self, enable: bool = True):
        """Replay checkpoint segments as captured hipGraphs (ads_acoustic_plan_set_graph)."""
        _lib.check(self.lib.ads_acoustic_plan_set_graph(self.h, int(bool(enable))))

    def _scratch_buf(self):
        need = int(self.lib.ads_acoustic_plan_history_scratch_size(self.h))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = _lib.device_bytes(need)
        return self._scratch

    def forward_history(self, model, srcv, with_aux: bool = True):
        model = _dev(model, self.dtype)
        NX, NY, L = self.param.NX, self.param.NY, self.param.NSTEP + 1
        shape = (self.nshots, L, NX + 2, NY + 2)
        u = torch.empty(shape, dtype=self.dtype, device="cuda")
        phi = torch.empty(shape, dtype=self.dtype, device="cuda") if with_aux else None
        psi = torch.empty(shape, dtype=self.dtype, device="cuda") if with_aux else None
        sc = self._scratch_buf()
        _lib.check(self.lib.ads_acoustic_plan_forward_history(self.h, _vp(model), _vp(srcv),
                                                              _vp(u), _vp(phi), _vp(psi),
                                                              _vp(sc), _stream()))
        return u, phi, psi

    def backward_history(self, model, srcv, u_hist, grad_u_hist):
        model = _dev(model, self.dtype)
        g = torch.empty_like(model)
        sc = self._scratch_buf()
        gu = grad_u_hist.contiguous()
        _lib.check(self.lib.ads_acoustic_plan_backward_history(self.h, _vp(model), _vp(srcv),
                                                               _vp(u_hist), _vp(gu), _vp(g),
                                                               _vp(sc), _stream()))
        return g


class _HistoryFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, plan, srcv):
        u, phi, psi = plan.forward_history(c, srcv)
        ctx.plan = plan
        ctx.save_for_backward(c, srcv, u)
        ctx.mark_non_differentiable(phi, psi)
        return u, phi, psi

    @staticmethod
    def backward(ctx, gu, gphi, gpsi):
        c, srcv, u = ctx.saved_tensors
        if gu is None:
            return None, None, None
        g = ctx.plan.backward_history(c, srcv, u, gu)
        return g, None, None


class _TracesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, plan, srcv, K):
        tr = plan.forward(model, srcv, K)
        ctx.plan, ctx.K, ctx.tok = plan, K, plan._tok
        ctx.save_for_backward(model, srcv)
        return tr

    @staticmethod
    def backward(ctx, g):
        model, srcv = ctx.saved_tensors
        plan = ctx.plan
        if not plan.holds_forward(ctx.tok, ctx.K):
            # another call used the workspace since this forward (a second acoustic_traces on
            # the plan, a gradient, an autotune, a plan sharing the buffer, or an earlier
            # backward): rebuild this forward's checkpoints before the sweep
            plan.forward(model, srcv, ctx.K)
        return plan.backward(model, srcv, g.contiguous(), ctx.K), None, None, None


def acoustic_traces(plan: AcousticPlan, model: torch.Tensor, srcv: torch.Tensor, K: int):
    """Differentiable batched traces (flat plan layout); the pullback is the HIP adjoint."""
    return _TracesFn.apply(model, plan, srcv, K)


@dataclass
class AcousticPropagator:
    """AcousticPropagator, src/Struct.jl:140-147 (u, phi, psi of shape (NSTEP+1, NX+2, NY+2))."""
    param: AcousticPropagatorParams
    src: AcousticSource
    u: Any
    φ: Any
    ψ: Any


def AcousticPropagatorSolver(param: AcousticPropagatorParams, src: AcousticSource, c) -> AcousticPropagator:
    """AcousticPropagatorSolver(param, src, c), src/Core.jl:562-620.

    ``c`` is the velocity (NX+2, NY+2); it is squared inside (Core.jl:564).  The dtype follows
    ``c`` (float64 unless a float32 tensor is given).  ``u`` is differentiable w.r.t. ``c``
    through the hand-written adjoint (ads_acoustic_plan_backward_history).
    """
    compute_PML_Params_(param)
    dtype = c.dtype if isinstance(c, torch.Tensor) else torch.float64
    ct = c if isinstance(c, torch.Tensor) and c.is_cuda else _dev(c, dtype)
    if tuple(ct.shape) != (param.NX + 2, param.NY + 2):
        ct = ct.reshape(param.NX + 2, param.NY + 2)
    srcv = _dev(src.srcv, dtype)
    nsrc = len(np.asarray(src.srci).reshape(-1))
    if srcv.dim() == 1:
        srcv = srcv.reshape(-1, 1)
    if tuple(srcv.shape) != (param.NSTEP, nsrc):
        raise ValueError(f"srcv must be (NSTEP, nsrc) = ({param.NSTEP}, {nsrc}), got "
                         f"{tuple(srcv.shape)}")
    plan = AcousticPlan(param, [src], [AcousticReceiver(np.zeros(0, np.int64), np.zeros(0, np.int64))],
                        dtype=dtype)
    u, phi, psi = _HistoryFn.apply(ct, plan, srcv.contiguous())
    return AcousticPropagator(param, src, u[0], phi[0], psi[0])


def SimulatedObservation_(ap: AcousticPropagator, rcv: AcousticReceiver):
    """SimulatedObservation!(ap, rcv), src/Core.jl:726-730: rcv.rcvv = u[:, idx] of shape
    (NSTEP+1, nrcv), idx = (rcvi-1)(NY+2) + rcvj (1-based)."""
    NY = ap.param.NY
    idx = (np.asarray(rcv.rcvi, np.int64) - 1) * (NY + 2) + (np.asarray(rcv.rcvj, np.int64) - 1)
    if len(idx) and (idx.min() < 0 or idx.max() >= (ap.param.NX + 2) * (NY + 2)):
        raise IndexError("receiver outside the padded grid")
    flat = ap.u.reshape(ap.u.shape[0], -1)
    rcv.rcvv = flat[:, torch.as_tensor(idx, device=flat.device)]
    return rcv.rcvv


# ------------------------------------------------------------------------------------------
# the per-step custom op (drop-in for AcousticOneStep / AcousticOneStepGrad)
# ------------------------------------------------------------------------------------------

class _OneStepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx, ny):
        lib = _lib.load()
        ins = [t.contiguous() for t in (w, wold, phi, psi, sigma, tau, c)]
        u, po, so = (torch.empty_like(ins[0]) for _ in range(3))
        f = lib.ads_acoustic_one_step_f64 if ins[0].dtype == torch.float64 else lib.ads_acoustic_one_step_f32
        _lib.check(f(*[_vp(t) for t in ins], float(dt), float(hx), float(hy), int(nx), int(ny),
                     _vp(u), _vp(po), _vp(so), _stream()))
        ctx.save_for_backward(*ins)
        ctx.sc = (dt, hx, hy, nx, ny)
        return u, po, so

    @staticmethod
    def backward(ctx, gu, gp, gs):
        lib = _lib.load()
        ins = ctx.saved_tensors
        z = lambda g: torch.zeros_like(ins[0]) if g is None else g.contiguous()
        gu, gp, gs = z(gu), z(gp), z(gs)
        outs = [torch.empty_like(ins[0]) for _ in range(5)]
        dt, hx, hy, nx, ny = ctx.sc
        f = (lib.ads_acoustic_one_step_grad_f64 if ins[0].dtype == torch.float64
             else lib.ads_acoustic_one_step_grad_f32)
        _lib.check(f(_vp(gu), _vp(gp), _vp(gs), *[_vp(t) for t in ins], float(dt), float(hx),
                     float(hy), int(nx), int(ny), *[_vp(o) for o in outs], _stream()))
        gw, gwold, gphi, gpsi, gc = outs
        return gw, gwold, gphi, gpsi, None, None, gc, None, None, None, None, None


def acoustic_one_step(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx, ny):
    """The AcousticOneStep op (AcousticOneStep.cpp:14-93) with its gradient."""
    return _OneStepFn.apply(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, nx, ny)


def acoustic_one_step_customop(param: AcousticPropagatorParams, w, wold, φ, ψ, σ, τ, c):
    """acoustic_one_step_customop(param, ...), src/Core.jl:475-487."""
    return acoustic_one_step(w, wold, φ, ψ, σ, τ, c, param.DELTAT, param.DELTAX, param.DELTAY,
                             param.NX, param.NY)
