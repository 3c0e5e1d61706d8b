"""Parameter and geometry types, mirroring src/Struct.jl (field names and defaults kept)."""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _lib


@dataclass
class AcousticPropagatorParams:
    """AcousticPropagatorParams, src/Struct.jl:82-121.

    ``PropagatorKernel``: 0 advances phi/psi from the new u (Core.jl:528-549, the default);
    1 and 2 from w (the AcousticOneStep op, Core.jl:475-525).  The auxiliary fields
    (damping_x, Sigma_x, IJ, ...) are filled by :func:`compute_PML_Params_`.
    """
    NX: int = 101
    NY: int = 641
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
    IJ: Any = None
    IJn: Any = None
    IJp: Any = None
    IpJ: Any = None
    IpJp: Any = None
    IpJn: Any = None
    InJ: Any = None
    InJn: Any = None
    InJp: Any = None
    IT_DISPLAY: int = 0
    PropagatorKernel: int = 0

    def to_c(self, square_velocity: bool = True, origin=(0.0, 0.0), global_n=(0, 0)):
        return _lib.AcousticParamsC(
            int(self.NX), int(self.NY), int(self.NSTEP), float(self.DELTAX), float(self.DELTAY),
            float(self.DELTAT), int(bool(self.USE_PML_XMIN)), int(bool(self.USE_PML_XMAX)),
            int(bool(self.USE_PML_YMIN)), int(bool(self.USE_PML_YMAX)), int(self.NPOINTS_PML),
            float(self.Rcoef), float(self.vp_ref), int(self.PropagatorKernel),
            int(bool(square_velocity)), float(origin[0]), float(origin[1]), int(global_n[0]),
            int(global_n[1]))


@dataclass
class ElasticPropagatorParams:
    """ElasticPropagatorParams, src/Struct.jl:4-32 (ALPHA_MAX_PML defaults to 2*pi*f0/2)."""
    NX: int = 101
    NY: int = 641
    NSTEP: int = 2000 * 2
    DELTAX: float = 10.0
    DELTAY: float = 10.0
    DELTAT: float = 2.0e-3 / 2
    f0: float = 5.0
    vp_ref: float = 2000.0
    USE_PML_XMIN: bool = True
    USE_PML_XMAX: bool = True
    USE_PML_YMIN: bool = True
    USE_PML_YMAX: bool = True
    NPOINTS_PML: int = 12
    NPOWER: float = 2.0
    K_MAX_PML: float = 1.0
    ALPHA_MAX_PML: Optional[float] = None
    Rcoef: float = 0.001
    IT_DISPLAY: int = 0

    def __post_init__(self):
        if self.ALPHA_MAX_PML is None:
            self.ALPHA_MAX_PML = 2.0 * math.pi * (self.f0 / 2.0)

    def to_c(self, subdomain: bool = False, origin=(0.0, 0.0), global_n=(0, 0)):
        return _lib.ElasticParamsC(
            int(self.NX), int(self.NY), int(self.NSTEP), float(self.DELTAX), float(self.DELTAY),
            float(self.DELTAT), float(self.f0), float(self.vp_ref), int(bool(self.USE_PML_XMIN)),
            int(bool(self.USE_PML_XMAX)), int(bool(self.USE_PML_YMIN)),
            int(bool(self.USE_PML_YMAX)), int(self.NPOINTS_PML), float(self.NPOWER),
            float(self.K_MAX_PML), float(self.ALPHA_MAX_PML), float(self.Rcoef),
            int(bool(subdomain)), float(origin[0]), float(origin[1]), int(global_n[0]),
            int(global_n[1]))


@dataclass
class MPIElasticPropagatorParams:
    """MPIElasticPropagatorParams, src/MPIElastic.jl:3-57: a global NX x NY grid split into
    M x N blocks of n x n cells; (II, JJ) (1-based) is the block this process steps.  On one
    GPU a block is stepped with zero halos (the one-subdomain form)."""
    NX: int = 0
    NY: int = 0
    M: int = 0
    N: int = 0
    n: int = 300
    NSTEP: int = 2000 * 2
    DELTAX: float = 10.0
    DELTAY: float = 10.0
    DELTAT: float = 2.0e-3 / 2
    f0: float = 5.0
    vp_ref: float = 2000.0
    USE_PML_XMIN: bool = True
    USE_PML_XMAX: bool = True
    USE_PML_YMIN: bool = True
    USE_PML_YMAX: bool = True
    NPOINTS_PML: int = 12
    NPOWER: float = 2.0
    K_MAX_PML: float = 1.0
    ALPHA_MAX_PML: Optional[float] = None
    Rcoef: float = 0.001
    IT_DISPLAY: int = 0
    II: int = -1
    JJ: int = -1

    def __post_init__(self):
        if self.ALPHA_MAX_PML is None:
            self.ALPHA_MAX_PML = 2.0 * math.pi * (self.f0 / 2.0)

    def to_c(self):
        """The block as an ads_elastic_params with subdomain = 1.  The profile origins follow
        compute_PML_Params!, MPIElastic.jl:231/306: x = (II-1) n DELTAX, and the y origin is
        (JJ-1) n DELTAX (the reference uses DELTAX there too)."""
        if self.II < 1 or self.JJ < 1:
            raise ValueError("call compute_PML_Params_mpi(param) first (II, JJ unset)")
        n = int(self.n)
        return _lib.ElasticParamsC(
            n, n, int(self.NSTEP), float(self.DELTAX), float(self.DELTAY), float(self.DELTAT),
            float(self.f0), float(self.vp_ref), int(bool(self.USE_PML_XMIN)),
            int(bool(self.USE_PML_XMAX)), int(bool(self.USE_PML_YMIN)),
            int(bool(self.USE_PML_YMAX)), int(self.NPOINTS_PML), float(self.NPOWER),
            float(self.K_MAX_PML), float(self.ALPHA_MAX_PML), float(self.Rcoef), 1,
            float((self.II - 1) * n * self.DELTAX), float((self.JJ - 1) * n * self.DELTAX),
            int(self.NX), int(self.NY))


def compute_PML_Params_mpi(param: MPIElasticPropagatorParams, rank: int = 0,
                           size: Optional[int] = None) -> MPIElasticPropagatorParams:
    """compute_PML_Params!(::MPIElasticPropagatorParams), MPIElastic.jl:154-176: the block grid
    and this rank's block (II, JJ) = (rank div N + 1, rank mod N + 1).  The CPML coefficients
    themselves are computed by the library (ads_elastic_pml_params with the block's origin)."""
    n = int(param.n)
    if param.NX <= 0 or param.NY <= 0 or param.NX % n or param.NY % n:
        raise ValueError("NX and NY must be positive multiples of n")
    param.M, param.N = param.NX // n, param.NY // n
    if size is not None and size != param.M * param.N:
        raise ValueError(f"MPI size mismatch: size = {size}, M = {param.M}, N = {param.N}")
    param.II, param.JJ = rank // param.N + 1, rank % param.N + 1
    return param


@dataclass
class AcousticSource:
    """AcousticSource, src/Struct.jl:123-127: 1-based (srci, srcj) on the padded grid and
    srcv of shape (NSTEP, nsrc)."""
    srci: Any
    srcj: Any
    srcv: Any


@dataclass
class AcousticReceiver:
    """AcousticReceiver, src/Struct.jl:129-133; rcvv is filled by SimulatedObservation_."""
    rcvi: Any
    rcvj: Any
    rcvv: Any = None


@dataclass
class ElasticSource:
    """ElasticSource, src/Struct.jl:50-55; srctype 0:vx 1:vy 2:sxx 3:syy 4:sxy (:44-48)."""
    srci: Any
    srcj: Any
    srctype: Any
    srcv: Any


@dataclass
class ElasticReceiver:
    """ElasticReceiver, src/Struct.jl:57-62."""
    rcvi: Any
    rcvj: Any
    rcvtype: Any
    rcvv: Any = None


class MPIElasticSource:
    """MPIElasticSource(param, srci, srcj, srctype, srcv), MPIElastic.jl:66-97: keeps the
    sources whose global 1-based (srci, srcj) fall in block (II, JJ) and stores them as 1-based
    indices into the block's (n+4) x (n+4) arrays (local index + 2)."""

    def __init__(self, param: MPIElasticPropagatorParams, srci, srcj, srctype, srcv):
        n, II, JJ = int(param.n), int(param.II), int(param.JJ)
        srci, srcj, srctype = (np.asarray(x, np.int64).reshape(-1) for x in (srci, srcj, srctype))
        srcv = np.asarray(srcv, np.float64)
        if srcv.ndim == 1:
            srcv = srcv.reshape(-1, 1)
        keep = ((II - 1) * n + 1 <= srci) & (srci <= II * n) & ((JJ - 1) * n + 1 <= srcj) & (srcj <= JJ * n)
        self.srci = srci[keep] - (II - 1) * n + 2
        self.srcj = srcj[keep] - (JJ - 1) * n + 2
        self.srctype = srctype[keep]
        self.srcv = srcv[:, keep]


class MPIElasticReceiver:
    """MPIElasticReceiver(param, rcvi, rcvj, rcvtype), MPIElastic.jl:99-127: the receivers of
    block (II, JJ) as 1-based indices into the n x n block."""

    def __init__(self, param: MPIElasticPropagatorParams, rcvi, rcvj, rcvtype):
        n, II, JJ = int(param.n), int(param.II), int(param.JJ)
        rcvi, rcvj, rcvtype = (np.asarray(x, np.int64).reshape(-1) for x in (rcvi, rcvj, rcvtype))
        keep = ((II - 1) * n + 1 <= rcvi) & (rcvi <= II * n) & ((JJ - 1) * n + 1 <= rcvj) & (rcvj <= JJ * n)
        self.rcvi = rcvi[keep] - (II - 1) * n
        self.rcvj = rcvj[keep] - (JJ - 1) * n
        self.rcvtype = rcvtype[keep]
        self.rcvv = None


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).reshape(-1))


class Geometry:
    """Host-side shot geometry (CSR over shots) kept alive for the C struct's pointers."""

    def __init__(self, srcs, rcvs, nstep=None, elastic=False):
        if len(srcs) != len(rcvs):
            raise ValueError(f"{len(srcs)} sources but {len(rcvs)} receivers")
        self.nshots = len(srcs)
        si, sj, st, ri, rj, rt = [], [], [], [], [], []
        sp, rp = [0], [0]
        for s, r in zip(srcs, rcvs):
            a, b = _i64(s.srci), _i64(s.srcj)
            if len(a) != len(b):
                raise ValueError("srci and srcj differ in length")
            si.append(a)
            sj.append(b)
            st.append(_i64(s.srctype) if elastic else np.zeros(len(a), np.int64))
            sp.append(sp[-1] + len(a))
            a, b = _i64(r.rcvi), _i64(r.rcvj)
            if len(a) != len(b):
                raise ValueError("rcvi and rcvj differ in length")
            ri.append(a)
            rj.append(b)
            rt.append(_i64(r.rcvtype) if elastic else np.zeros(len(a), np.int64))
            rp.append(rp[-1] + len(a))
        cat = lambda xs: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0, np.int64))
        self.srci, self.srcj, self.srctype = cat(si), cat(sj), cat(st)
        self.rcvi, self.rcvj, self.rcvtype = cat(ri), cat(rj), cat(rt)
        self.src_ptr, self.rcv_ptr = _i64(sp), _i64(rp)
        self.nsrc = np.diff(self.src_ptr)
        self.nrcv = np.diff(self.rcv_ptr)
        self.nstep = None if nstep is None else _i64(nstep)
        for arr in (self.srctype, self.rcvtype):
            if len(arr) and (arr.min() < 0 or arr.max() > 4):
                raise ValueError("source/receiver types must be in 0..4 (vx, vy, sxx, syy, sxy)")

    def to_c(self):
        P = lambda a: a.ctypes.data_as(_lib.i64p) if a is not None else None
        return _lib.GeometryC(self.nshots, P(self.nstep), P(self.src_ptr), P(self.srci),
                              P(self.srcj), P(self.srctype), P(self.rcv_ptr), P(self.rcvi),
                              P(self.rcvj), P(self.rcvtype))
