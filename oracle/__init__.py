"""CPU FP64 oracle for the wave-propagation hot path (test infrastructure only).

Thin numpy/ctypes front end over ``oracle/liboracle.so`` (built from ``ads_oracle.c`` by
``oracle/Makefile``).  Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, as the checker and the CPU baseline; the product path
never does.  Functions cite the reference file:line they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libadseismic_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
i64 = C.c_int64
f64 = C.c_double


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, when the reference tree is present, the reference headers)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def ref_lib():
    """The reference's own header-only kernels (oracle/_ref), or None if not built here."""
    if not os.path.exists(_REF):
        return None
    return C.CDLL(_REF)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def set_num_threads(n: int) -> None:
    lib().or_set_num_threads(C.c_int(n))


def num_threads() -> int:
    return int(lib().or_num_threads())


# ------------------------------------------------------------------------------------------
# acoustic
# ------------------------------------------------------------------------------------------

def acoustic_pml_axis(length, h, origin, NXg, npml, Rcoef, vp_ref, use_min=True, use_max=True):
    """pml_helper on one axis (src/Core.jl:738-761), coord(k) = k*h + origin."""
    out = np.zeros(length)
    lib().or_acoustic_pml_axis(i64(length), f64(h), f64(origin), i64(NXg), i64(npml), f64(Rcoef),
                               f64(vp_ref), C.c_int(int(use_min)), C.c_int(int(use_max)), _d(out))
    return out


def acoustic_pml(NX, NY, hx, hy, npml, Rcoef, vp_ref, use=(True, True, True, True)):
    """Full-grid Sigma_x, Sigma_y of compute_PML_Params! (src/Core.jl:622-655)."""
    sx = acoustic_pml_axis(NX + 2, hx, 0.0, NX, npml, Rcoef, vp_ref, use[0], use[1])
    ty = acoustic_pml_axis(NY + 2, hy, 0.0, NY, npml, Rcoef, vp_ref, use[2], use[3])
    Sx = np.repeat(sx[:, None], NY + 2, axis=1)
    Sy = np.repeat(ty[None, :], NX + 2, axis=0)
    return np.ascontiguousarray(Sx), np.ascontiguousarray(Sy)


def acoustic_step(kernel, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY):
    """One step (a3 kernel 0: Core.jl:528-549; a4 kernel 1: AcousticOneStepCpu.h:1-48)."""
    N = (NX + 2) * (NY + 2)
    u, po, so = np.empty(N), np.empty(N), np.empty(N)
    args = [_f64(x).ravel() for x in (w, wold, phi, psi, sigma, tau, c)]
    lib().or_acoustic_step(C.c_int(kernel), *[_d(a) for a in args], f64(dt), f64(hx), f64(hy),
                           i64(NX), i64(NY), _d(u), _d(po), _d(so))
    return u, po, so


def acoustic_step_vjp(kernel, gu, gphi, gpsi, w, wold, phi, psi, sigma, tau, c, dt, hx, hy, NX, NY):
    """VJP of one step (a5, AcousticOneStepCpu.h:51-125 for kernel 1)."""
    N = (NX + 2) * (NY + 2)
    outs = [np.empty(N) for _ in range(5)]
    scratch = np.empty(4 * N)
    args = [_f64(x).ravel() for x in (gu, gphi, gpsi, w, wold, phi, psi, sigma, tau, c)]
    lib().or_acoustic_step_vjp(C.c_int(kernel), *[_d(a) for a in args], f64(dt), f64(hx), f64(hy),
                               i64(NX), i64(NY), *[_d(o) for o in outs], _d(scratch))
    return tuple(outs)  # grad_w, grad_wold, grad_phi, grad_psi, grad_c


def flat_index(i1, j1, NY):
    """1-based (i, j) on the padded grid -> 0-based flat index (Core.jl:600, 727)."""
    return (np.asarray(i1, dtype=np.int64) - 1) * (NY + 2) + (np.asarray(j1, dtype=np.int64) - 1)


def acoustic_forward(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, C2, srci, srcj, srcv):
    """AcousticPropagatorSolver (Core.jl:562-620) with C2 = c^2.  Returns u, phi, psi histories."""
    N = (NX + 2) * (NY + 2)
    sidx = flat_index(srci, srcj, NY)
    srcv = _f64(srcv).reshape(NSTEP, len(sidx))
    u = np.empty((NSTEP + 1) * N)
    p = np.empty_like(u)
    s = np.empty_like(u)
    lib().or_acoustic_forward(C.c_int(kernel), i64(NX), i64(NY), i64(NSTEP), f64(dt), f64(hx),
                              f64(hy), _d(_f64(sigma).ravel()), _d(_f64(tau).ravel()),
                              _d(_f64(C2).ravel()), i64(len(sidx)), _i(sidx), _d(srcv), _d(u),
                              _d(p), _d(s))
    sh = (NSTEP + 1, NX + 2, NY + 2)
    return u.reshape(sh), p.reshape(sh), s.reshape(sh)


def acoustic_receive(u_hist, rcvi, rcvj, NY):
    """SimulatedObservation! (Core.jl:726-730): (NSTEP+1, nrcv)."""
    idx = flat_index(rcvi, rcvj, NY)
    return u_hist.reshape(u_hist.shape[0], -1)[:, idx]


def acoustic_adjoint(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, C2, u_hist, phi_hist, psi_hist,
                     gu_hist):
    """Reverse sweep of the time loop: d loss / d C2 given the u-history cotangent."""
    N = (NX + 2) * (NY + 2)
    gC = np.empty(N)
    lib().or_acoustic_adjoint(C.c_int(kernel), i64(NX), i64(NY), i64(NSTEP), f64(dt), f64(hx),
                              f64(hy), _d(_f64(sigma).ravel()), _d(_f64(tau).ravel()),
                              _d(_f64(C2).ravel()), _d(_f64(u_hist).ravel()),
                              _d(_f64(phi_hist).ravel()), _d(_f64(psi_hist).ravel()),
                              _d(_f64(gu_hist).ravel()), _d(gC))
    return gC.reshape(NX + 2, NY + 2)


def acoustic_step_vjp_gather(gu, gphi, gpsi, w, sigma, tau, c, dt, hx, hy, NX, NY):
    """Kernel-1 step VJP in gather form (parallel over rows): grad_w, grad_wold, grad_phi,
    grad_psi, grad_c."""
    N = (NX + 2) * (NY + 2)
    outs = [np.empty(N) for _ in range(5)]
    args = [_f64(x).ravel() for x in (gu, gphi, gpsi, w, sigma, tau, c)]
    lib().or_acoustic_step_vjp_gather(*[_d(a) for a in args], f64(dt), f64(hx), f64(hy), i64(NX),
                                      i64(NY), *[_d(o) for o in outs])
    return tuple(outs)


def acoustic_adjoint_par(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, C2, u_hist, gu_hist):
    """acoustic_adjoint for kernel 1 with the gather-form VJP, multi-threaded."""
    N = (NX + 2) * (NY + 2)
    gC = np.empty(N)
    rc = lib().or_acoustic_adjoint_par(C.c_int(kernel), i64(NX), i64(NY), i64(NSTEP), f64(dt),
                                       f64(hx), f64(hy), _d(_f64(sigma).ravel()),
                                       _d(_f64(tau).ravel()), _d(_f64(C2).ravel()),
                                       _d(_f64(u_hist).ravel()), _d(_f64(gu_hist).ravel()), _d(gC))
    if rc != 0:
        raise ValueError(f"or_acoustic_adjoint_par failed ({rc}): kernel 1 only")
    return gC.reshape(NX + 2, NY + 2)


def acoustic_misfit_gradient(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, c, srci, srcj, srcv,
                             rcvi, rcvj, observed):
    """loss = sum((rcvv - observed)^2) (Utils.jl:308) and d loss / d c (c = velocity)."""
    C2 = _f64(c) ** 2
    u, p, s = acoustic_forward(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, C2, srci, srcj, srcv)
    rc = acoustic_receive(u, rcvi, rcvj, NY)
    res = rc - observed
    loss = float(np.sum(res * res))
    gu = np.zeros(u.shape).reshape(NSTEP + 1, -1)
    idx = flat_index(rcvi, rcvj, NY)
    np.add.at(gu, (slice(None), idx), 2.0 * res)
    gC = acoustic_adjoint(kernel, NX, NY, NSTEP, dt, hx, hy, sigma, tau, C2, u, p, s, gu)
    return loss, 2.0 * _f64(c) * gC, rc


# ------------------------------------------------------------------------------------------
# elastic
# ------------------------------------------------------------------------------------------

def elastic_pml_axis(length, h, origin, NXg, npml, npower, kmax, alpha_max, Rcoef, vp_ref, dt,
                     use_min=True, use_max=True):
    """CPML coefficients on one axis (src/Core.jl:231-407): returns a, b, K as (2, length)."""
    outs = [np.empty(length) for _ in range(6)]
    lib().or_elastic_pml_axis(i64(length), f64(h), f64(origin), i64(NXg), i64(npml), f64(npower),
                              f64(kmax), f64(alpha_max), f64(Rcoef), f64(vp_ref), f64(dt),
                              C.c_int(int(use_min)), C.c_int(int(use_max)), *[_d(o) for o in outs])
    a, ah, b, bh, k, kh = outs
    return np.stack([a, ah]), np.stack([b, bh]), np.stack([k, kh])


def elastic_pml(NX, NY, hx, hy, dt, npml, npower, kmax, alpha_max, Rcoef, vp_ref,
                use=(True, True, True, True)):
    """compute_PML_Params(::ElasticPropagatorParams) -> ax, bx, kx, ay, by, ky (each (2, n))."""
    ax, bx, kx = elastic_pml_axis(NX, hx, 0.0, NX, npml, npower, kmax, alpha_max, Rcoef, vp_ref, dt,
                                  use[0], use[1])
    ay, by, ky = elastic_pml_axis(NY, hy, 0.0, NY, npml, npower, kmax, alpha_max, Rcoef, vp_ref, dt,
                                  use[2], use[3])
    return ax, bx, kx, ay, by, ky


def elastic_step(NX, NY, hx, hy, dt, rho, lam, mu, ax, bx, ay, by, f, m, srci, srcj, srctype, srcv,
                 sub=False):
    """One elastic step in place (Core.jl:215-228 + AddSource.cpp:57-85).  sub: the
    one-subdomain form of MPIElastic.jl (NX, NY = block + 2; profiles of the block length)."""
    f = _f64(f)
    m = _f64(m)
    srci, srcj, srctype = _i64(srci), _i64(srcj), _i64(srctype)
    lib().or_elastic_step(i64(NX), i64(NY), f64(hx), f64(hy), f64(dt), _d(_f64(rho).ravel()),
                          _d(_f64(lam).ravel()), _d(_f64(mu).ravel()), _d(_f64(ax)), _d(_f64(bx)),
                          _d(_f64(ay)), _d(_f64(by)), _d(f), _d(m), i64(len(srci)), _i(srci),
                          _i(srcj), _i(srctype), _d(_f64(srcv)), C.c_int(int(sub)))
    return f, m


def elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci, srcj, srctype, srcv,
                    keep_mem=True, sub=False):
    """ElasticPropagatorSolver (Core.jl:31-94): histories (5, NSTEP+1, NX+2, NY+2) in the order
    vx, vy, sxx, syy, sxy, and memory variables (8, NSTEP+1, ...)."""
    ax, bx, _, ay, by, _ = pml
    N = (NX + 2) * (NY + 2)
    srci, srcj, srctype = _i64(srci), _i64(srcj), _i64(srctype)
    srcv = _f64(srcv).reshape(NSTEP, len(srci))
    hist = np.empty(5 * (NSTEP + 1) * N)
    mem = np.empty(8 * (NSTEP + 1) * N) if keep_mem else None
    lib().or_elastic_forward(i64(NX), i64(NY), i64(NSTEP), f64(hx), f64(hy), f64(dt),
                             _d(_f64(rho).ravel()), _d(_f64(lam).ravel()), _d(_f64(mu).ravel()),
                             _d(_f64(ax)), _d(_f64(bx)), _d(_f64(ay)), _d(_f64(by)),
                             i64(len(srci)), _i(srci), _i(srcj), _i(srctype), _d(srcv), _d(hist),
                             _d(mem) if mem is not None else None, C.c_int(int(sub)))
    sh = (NSTEP + 1, NX + 2, NY + 2)
    return hist.reshape((5,) + sh), (mem.reshape((8,) + sh) if mem is not None else None)


def elastic_receive(hist, rcvi, rcvj, rcvtype, NX, NY):
    """GetReceive (GetReceive.cpp:10-46): (nrcv, NSTEP+1)."""
    NSTEP = hist.shape[1] - 1
    rcvi, rcvj, rcvtype = _i64(rcvi), _i64(rcvj), _i64(rcvtype)
    out = np.zeros(len(rcvi) * (NSTEP + 1))
    lib().or_elastic_receive(i64(NX), i64(NY), i64(NSTEP), _d(_f64(hist).ravel()), i64(len(rcvi)),
                             _i(rcvi), _i(rcvj), _i(rcvtype), _d(out))
    return out.reshape(len(rcvi), NSTEP + 1)


def elastic_adjoint(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, hist, mem_hist, g_hist,
                    srci=None, srcj=None, srctype=None, sub=False):
    """Reverse sweep of the elastic loop: grads wrt rho, lambda, mu (and srcv if sources given)."""
    ax, bx, _, ay, by, _ = pml
    N = (NX + 2) * (NY + 2)
    gr, gl, gm = np.empty(N), np.empty(N), np.empty(N)
    nsrc = 0 if srci is None else len(srci)
    gsrc = np.zeros(NSTEP * nsrc) if nsrc else None
    si = _i64(srci if nsrc else [0])
    sj = _i64(srcj if nsrc else [0])
    st = _i64(srctype if nsrc else [0])
    lib().or_elastic_adjoint(i64(NX), i64(NY), i64(NSTEP), f64(hx), f64(hy), f64(dt),
                             _d(_f64(rho).ravel()), _d(_f64(lam).ravel()), _d(_f64(mu).ravel()),
                             _d(_f64(ax)), _d(_f64(bx)), _d(_f64(ay)), _d(_f64(by)),
                             _d(_f64(hist).ravel()), _d(_f64(mem_hist).ravel()),
                             _d(_f64(g_hist).ravel()), _d(gr), _d(gl), _d(gm), i64(nsrc), _i(si),
                             _i(sj), _i(st), _d(gsrc) if gsrc is not None else None,
                             C.c_int(int(sub)))
    sh = (NX + 2, NY + 2)
    out = (gr.reshape(sh), gl.reshape(sh), gm.reshape(sh))
    if gsrc is not None:
        out = out + (gsrc.reshape(NSTEP, nsrc),)
    return out


def elastic_misfit_gradient(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci, srcj, srctype, srcv,
                            rcvi, rcvj, rcvtype, observed, sub=False):
    """loss = sum((rcvv - observed)^2) and grads wrt rho, lambda, mu.  Receiver indices are
    1-based into the padded grid (for sub, the block index + 2)."""
    hist, mem = elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci, srcj, srctype,
                                srcv, sub=sub)
    rc = elastic_receive(hist, rcvi, rcvj, rcvtype, NX, NY)
    res = rc - observed
    loss = float(np.sum(res * res))
    g = np.zeros_like(hist).reshape(5, NSTEP + 1, -1)
    idx = flat_index(rcvi, rcvj, NY)
    for r in range(len(idx)):
        g[int(rcvtype[r]), :, idx[r]] += 2.0 * res[r]
    gr, gl, gm = elastic_adjoint(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, hist, mem, g, sub=sub)
    return loss, (gr, gl, gm), rc


# ------------------------------------------------------------------------------------------
# wavelets
# ------------------------------------------------------------------------------------------

def ricker(NT, a, shift, amp=1.0):
    """Ricker(param, a, shift, amp), src/Utils.jl:188-204."""
    out = np.empty(NT)
    lib().or_ricker(i64(NT), f64(a), f64(shift), f64(amp), _d(out))
    return out


def ricker_freq(nt, f0, dt, amp=1e4):
    """Ricker of the shipped fixtures (examples/nn_fwi/Marmousi_model.ipynb cell 4, cell 11)."""
    out = np.empty(nt)
    lib().or_ricker_freq(i64(nt), f64(f0), f64(dt), f64(amp), _d(out))
    return out


lib_restype = {"or_ricker_freq": i64, "or_num_threads": C.c_int}


def _setup_restypes():
    L = lib()
    for name, rt in lib_restype.items():
        getattr(L, name).restype = rt


try:
    _setup_restypes()
except OSError:
    pass
