"""One subdomain of the MPI acoustic solver, mirroring src/MPIAcoustic.jl.

The reference splits an NX x NY grid into M x N blocks of n x n cells, one per MPI rank.  Each
rank holds an n x n block of u, phi, psi and c (no padding ring), refills a one-cell halo from
its neighbours before every step (mpi_halo_exchange, MPIAcoustic.jl:221-236) and steps the
block with the acoustic scheme.  On one GPU the block is stepped at its real size with the
halo that a rank without neighbours gets, zero: the block becomes the interior of an
(n+2) x (n+2) grid and runs through the same acoustic plan as AcousticPropagatorSolver, with

- the block's PML profile: padded cell 0 sits at x = (II-1) n DELTAX, y = (JJ-1) n DELTAY and
  the far PML edges follow the global NX, NY (compute_PML_Params!, MPIAcoustic.jl:154-199);
- c used as given (the reference does not square it here, MPIAcoustic.jl:336), so gradients are
  d loss / d c;
- sources and receivers given 1-based on the global unpadded grid and kept when they fall in
  the block (MPIAcousticSource / MPIAcousticReceiver, MPIAcoustic.jl:56-116), stored 1-based on
  the n x n block.

PropagatorKernel 0 (the default) advances phi/psi from the new u as MPIAcoustic.jl:212-246
does; 1 and 2 from w (one_step_customop / one_step_ref, :250-330).  The reference's custom-op
gradient for this path is an uninitialised stub (MpiAcousticOneStep.cpp:191), so gradients here
are those of the kernel-0 graph, the hand-written adjoint of the plan.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib
from .acoustic import AcousticPlan, _HistoryFn, _dev, _stream, _vp
from .params import AcousticPropagatorParams, AcousticReceiver, AcousticSource

__all__ = ["MPIAcousticPropagatorParams", "compute_PML_Params_mpi_acoustic", "MPIAcousticSource",
           "MPIAcousticReceiver", "MPIAcousticPropagator", "MPIAcousticPropagatorSolver",
           "MPISimulatedObservation_acoustic", "mpi_acoustic_one_step"]


@dataclass
class MPIAcousticPropagatorParams:
    """MPIAcousticPropagatorParams, MPIAcoustic.jl:3-49 (field names and defaults kept)."""
    NX: int = 0
    NY: int = 0
    M: int = 0
    N: int = 0
    n: int = 300
    NSTEP: int = 2000 * 2
    DELTAX: float = 10.0
    DELTAY: float = 10.0
    DELTAT: float = 2.0e-3 / 2
    USE_PML_XMIN: bool = True
    USE_PML_XMAX: bool = True
    USE_PML_YMIN: bool = True
    USE_PML_YMAX: bool = True
    NPOINTS_PML: int = 12
    NPOWER: int = 2
    damping_x: Optional[float] = None
    damping_y: Optional[float] = None
    Rcoef: float = 0.001
    vp_ref: float = 1000.0
    Σx: Any = None
    Σy: Any = None
    IT_DISPLAY: int = 0
    II: int = -1
    JJ: int = -1
    PropagatorKernel: int = 0

    def block_params(self) -> AcousticPropagatorParams:
        """The block as an (n+2) x (n+2) padded grid of the acoustic plan."""
        return AcousticPropagatorParams(
            NX=int(self.n), NY=int(self.n), NSTEP=int(self.NSTEP), DELTAX=float(self.DELTAX),
            DELTAY=float(self.DELTAY), DELTAT=float(self.DELTAT),
            USE_PML_XMIN=self.USE_PML_XMIN, USE_PML_XMAX=self.USE_PML_XMAX,
            USE_PML_YMIN=self.USE_PML_YMIN, USE_PML_YMAX=self.USE_PML_YMAX,
            NPOINTS_PML=int(self.NPOINTS_PML), Rcoef=float(self.Rcoef), vp_ref=float(self.vp_ref),
            PropagatorKernel=int(self.PropagatorKernel))

    def origin(self):
        """(x, y) of the block's padded cell 0 (MPIAcoustic.jl:177-178)."""
        self._check_block()
        n = int(self.n)
        return ((self.II - 1) * n * self.DELTAX, (self.JJ - 1) * n * self.DELTAY)

    def _check_block(self):
        if self.II < 1 or self.JJ < 1:
            raise ValueError("block (II, JJ) not set: call compute_PML_Params_mpi_acoustic")

    def to_c(self):
        self._check_block()
        return self.block_params().to_c(square_velocity=False, origin=self.origin(),
                                         global_n=(int(self.NX), int(self.NY)))


def compute_PML_Params_mpi_acoustic(param: MPIAcousticPropagatorParams, rank: int = 0,
                                    size: Optional[int] = None) -> MPIAcousticPropagatorParams:
    """compute_PML_Params!(::MPIAcousticPropagatorParams), MPIAcoustic.jl:154-199: the block grid
    M = NX/n, N = NY/n, this rank's block (II, JJ) = (rank div N + 1, rank mod N + 1), the damping
    constants and the block's (n+2) x (n+2) profiles Sigma_x, Sigma_y (from the library's
    ads_acoustic_pml_profiles at the block origin).  ``size`` (the MPI size) is checked against
    M N when given."""
    n = int(param.n)
    if param.NX <= 0 or param.NY <= 0 or param.NX % n or param.NY % n:
        raise ValueError("NX and NY must be positive multiples of n")
    param.M, param.N = param.NX // n, param.NY // n
    if size is not None and size != param.M * param.N:
        raise ValueError(f"MPI size mismatch: size = {size}, M = {param.M}, N = {param.N}")
    if not 0 <= rank < param.M * param.N:
        raise ValueError(f"rank {rank} outside the {param.M} x {param.N} block grid")
    param.II, param.JJ = rank // param.N + 1, rank % param.N + 1
    Lx = param.NPOINTS_PML * param.DELTAX
    Ly = param.NPOINTS_PML * param.DELTAY
    param.damping_x = param.vp_ref / Lx * math.log(1 / param.Rcoef)
    param.damping_y = param.vp_ref / Ly * math.log(1 / param.Rcoef)
    pc = param.to_c()
    sx = np.zeros(n + 2)
    ty = np.zeros(n + 2)
    _lib.check(_lib.load().ads_acoustic_pml_profiles(C.byref(pc), sx.ctypes.data, ty.ctypes.data))
    param.Σx = np.repeat(sx[:, None], n + 2, axis=1)
    param.Σy = np.repeat(ty[None, :], n + 2, axis=0)
    return param


def _in_block(param, i, j):
    n, II, JJ = int(param.n), int(param.II), int(param.JJ)
    return ((II - 1) * n + 1 <= i) & (i <= II * n) & ((JJ - 1) * n + 1 <= j) & (j <= JJ * n)


class MPIAcousticSource:
    """MPIAcousticSource(param, srci, srcj, srcv), MPIAcoustic.jl:56-82: global 1-based
    (srci, srcj) on the unpadded NX x NY grid; the sources inside block (II, JJ) are kept as
    1-based indices on the n x n block, with their columns of srcv (NSTEP x nsrc).  A block
    without sources has srci = srcj = srcv = None (the reference's `missing`)."""

    def __init__(self, param: MPIAcousticPropagatorParams, srci, srcj, srcv):
        param._check_block()
        n = int(param.n)
        srci, srcj = (np.asarray(x, np.int64).reshape(-1) for x in (srci, srcj))
        srcv = srcv if isinstance(srcv, torch.Tensor) else np.asarray(srcv, np.float64)
        if srcv.ndim == 1:
            srcv = srcv.reshape(-1, 1)
        keep = _in_block(param, srci, srcj)
        if not keep.any():
            self.srci = self.srcj = self.srcv = None
            return
        self.srci = srci[keep] - (param.II - 1) * n
        self.srcj = srcj[keep] - (param.JJ - 1) * n
        idx = np.nonzero(keep)[0]
        self.srcv = srcv[:, torch.as_tensor(idx) if isinstance(srcv, torch.Tensor) else idx]


class MPIAcousticReceiver:
    """MPIAcousticReceiver(param, rcvi, rcvj), MPIAcoustic.jl:84-104: the receivers inside
    block (II, JJ) as 1-based indices on the n x n block (None when there are none)."""

    def __init__(self, param: MPIAcousticPropagatorParams, rcvi, rcvj):
        param._check_block()
        n = int(param.n)
        rcvi, rcvj = (np.asarray(x, np.int64).reshape(-1) for x in (rcvi, rcvj))
        keep = _in_block(param, rcvi, rcvj)
        self.rcvv = None
        if not keep.any():
            self.rcvi = self.rcvj = None
            return
        self.rcvi = rcvi[keep] - (param.II - 1) * n
        self.rcvj = rcvj[keep] - (param.JJ - 1) * n


@dataclass
class MPIAcousticPropagator:
    """MPIAcousticPropagator, MPIAcoustic.jl:113-121: u, phi, psi of shape (NSTEP+1, n, n)."""
    param: MPIAcousticPropagatorParams
    src: MPIAcousticSource
    u: Any
    φ: Any
    ψ: Any


def MPIAcousticPropagatorSolver(param: MPIAcousticPropagatorParams, src: MPIAcousticSource,
                                c) -> MPIAcousticPropagator:
    """MPIAcousticPropagatorSolver(param, src, c), MPIAcoustic.jl:334-404, on this rank's block.

    ``c`` is the n x n block of the (unsquared) model the scheme multiplies the Laplacian by.
    Level 0 and 1 are zero; level k >= 2 is one step from levels k-1, k-2, then
    u[srci, srcj] += DELTAT^2 srcv[k-1, :] (row k-1, 0-based).  ``u`` is differentiable w.r.t.
    ``c`` through the hand-written adjoint of the kernel-0 (or kernel-1) scheme."""
    param._check_block()
    n = int(param.n)
    dtype = c.dtype if isinstance(c, torch.Tensor) else torch.float64
    ct = c if isinstance(c, torch.Tensor) and c.is_cuda else _dev(c, dtype)
    if tuple(ct.shape) != (n, n):
        raise ValueError(f"c must be the n x n block ({n}, {n}), got {tuple(ct.shape)}")
    if src.srci is None:
        srci = srcj = np.zeros(0, np.int64)
        # nothing reads it; one element keeps the device pointer non-null
        srcv = torch.zeros(1, dtype=dtype, device="cuda")
    else:
        srci, srcj = src.srci + 1, src.srcj + 1          # block index -> padded block grid
        srcv = _dev(src.srcv, dtype)
        if tuple(srcv.shape) != (param.NSTEP, len(srci)):
            raise ValueError(f"srcv must be (NSTEP, nsrc) = ({param.NSTEP}, {len(srci)}), got "
                             f"{tuple(srcv.shape)}")
    bp = param.block_params()
    none = np.zeros(0, np.int64)
    plan = AcousticPlan(bp, [AcousticSource(srci, srcj, None)], [AcousticReceiver(none, none)],
                        dtype=dtype, square_velocity=False, origin=param.origin(),
                        global_n=(int(param.NX), int(param.NY)))
    cpad = F.pad(ct, (1, 1, 1, 1))                       # the ring's c is never read
    u, phi, psi = _HistoryFn.apply(cpad, plan, srcv.contiguous())
    inner = (slice(None), slice(1, n + 1), slice(1, n + 1))
    return MPIAcousticPropagator(param, src, u[0][inner], phi[0][inner], psi[0][inner])


def MPISimulatedObservation_acoustic(ap: MPIAcousticPropagator, rcv: MPIAcousticReceiver):
    """MPISimulatedObservation!(ap, rcv), MPIAcoustic.jl:422-431: rcv.rcvv = u[:, idx] of shape
    (NSTEP+1, nrcv) with idx = (rcvi-1) n + rcvj (1-based on the block); an empty
    (NSTEP+1, 0) record for a block without receivers."""
    n = int(ap.param.n)
    flat = ap.u.reshape(ap.u.shape[0], -1)
    if rcv.rcvi is None:
        rcv.rcvv = flat[:, :0]
        return rcv.rcvv
    idx = (np.asarray(rcv.rcvi, np.int64) - 1) * n + (np.asarray(rcv.rcvj, np.int64) - 1)
    if idx.min() < 0 or idx.max() >= n * n:
        raise IndexError("receiver outside the block")
    rcv.rcvv = flat[:, torch.as_tensor(idx, device=flat.device)]
    return rcv.rcvv


def mpi_acoustic_one_step(w, wold, phi, psi, sigma, tau, c, dt, hx, hy, n):
    """The MpiAcousticOneStep op (MpiAcousticOneStep.cpp forward, MpiAcousticOneStep.h:1-42):
    w, wold, phi, psi, sigma, tau are (n+2)^2 halo'd blocks, c is n x n; returns u, phi, psi of
    n x n (phi/psi from w, the custom-op variant).  Forward only: the reference registers no
    working gradient for it."""
    lib = _lib.load()
    ins = [t.contiguous() for t in (w, wold, phi, psi, sigma, tau, c)]
    if ins[0].numel() != (n + 2) ** 2 or ins[6].numel() != n * n:
        raise ValueError("w, wold, phi, psi, sigma, tau must hold (n+2)^2 values and c n^2")
    outs = [torch.empty(n * n, dtype=ins[0].dtype, device=ins[0].device) for _ in range(3)]
    f = (lib.ads_mpi_acoustic_one_step_f64 if ins[0].dtype == torch.float64
         else lib.ads_mpi_acoustic_one_step_f32)
    _lib.check(f(*[_vp(t) for t in ins], float(dt), float(hx), float(hy), int(n),
                 *[_vp(o) for o in outs], _stream()))
    return tuple(o.view(n, n) for o in outs)
