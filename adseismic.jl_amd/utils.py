"""Wavelets and material helpers, mirroring src/Utils.jl:188-258 (host-side, numpy)."""
from __future__ import annotations

import math

import numpy as np


def Ricker(epp, a, shift, amp=1.0):
    """Ricker(epp, a, shift, amp), src/Utils.jl:188-204: sample-domain wavelet, k = 1..NSTEP."""
    NT = int(epp.NSTEP)
    A = 2 / (math.sqrt(3 * a) * (math.pi ** 0.25))
    wsq = a ** 2
    vec = np.arange(1, NT + 1, dtype=np.float64) - shift
    xsq = vec ** 2
    return amp * A * (1 - xsq / wsq) * np.exp(-xsq / (2 * wsq))


def Gauss(epp, a, shift=None, amp=1.0):
    """Gauss(epp, a, shift, amp), src/Utils.jl:206-220."""
    if shift is None:
        shift = 1.2 / a
    t = np.arange(0, int(epp.NSTEP), dtype=np.float64) * epp.DELTAT
    A = math.pi ** 2 * a ** 2
    return amp * 2.0 * A * np.exp(-A * (t - shift) ** 2)


def ricker_freq(f0, dt, nt, amp=1e4):
    """Frequency-domain Ricker of the reference's fixtures (examples/nn_fwi/
    Marmousi_model.ipynb cells 4 and 11): (1-2 pi^2 f^2 t^2) exp(-pi^2 f^2 t^2) * amp on
    t = arange(-T, T, dt), T = sqrt(20)/(pi f), zero-padded/cut to nt."""
    T = math.sqrt(20) / (math.pi * f0)
    t = np.arange(-T, T, dt)
    src = (1 - 2 * math.pi ** 2 * f0 ** 2 * t ** 2) * np.exp(-math.pi ** 2 * f0 ** 2 * t ** 2) * amp
    vec = np.zeros(nt)
    vec[:len(src)] = src[:nt]
    return vec


def compute_lame_parameters(*args):
    """compute_lame_parameters, src/Utils.jl:223-247.

    (NX, NY, vp, vs, rho) -> constant (NX+2, NY+2) lambda, mu, rho;
    (param, vp, vs, rho) with an MPIElasticPropagatorParams -> the same on its (n+4)^2 block
                          arrays (Utils.jl:243-245);
    (vp, vs, rho) arrays  -> lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2, rho."""
    if len(args) == 4:
        param, vp, vs, rho = args
        return compute_lame_parameters(param.n + 2, param.n + 2, vp, vs, rho)
    if len(args) == 5:
        NX, NY, vp, vs, rho = args
        shape = (NX + 2, NY + 2)
        return (np.full(shape, rho * (vp * vp - 2.0 * vs * vs)), np.full(shape, rho * vs * vs),
                np.full(shape, float(rho)))
    vp, vs, rho = args
    mu = rho * vs * vs
    lam = rho * (vp * vp - 2.0 * vs * vs)
    return lam, mu, rho


def compute_properties(vp, vs, rho):
    """compute_properties, src/Utils.jl:249-258: mu = rho vs^2, lambda = rho vp^2 - 2 mu."""
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    return lam, mu, rho
