"""MAT-file loaders for models, sources and receivers, mirroring src/IO.jl:4-161.

Files are read with ``scipy.io.loadmat`` (a data reader; nothing in the file is executed).
Field conventions follow the reference: ``nx``/``ny`` are padded-grid sizes (NX = nx - 2),
``nt`` is NSTEP, ``dx``/``dy``/``dt`` the spacings; ``source`` / ``receiver`` are struct arrays
with 1-based ``ix``, ``iy`` (padded grid), ``vec`` (source wavelets) and ``type`` (elastic).
"""
from __future__ import annotations

from typing import Any, Dict, List

import numpy as np

from .params import (AcousticPropagatorParams, AcousticReceiver, AcousticSource,
                     ElasticPropagatorParams, ElasticReceiver, ElasticSource,
                     MPIElasticPropagatorParams, MPIElasticReceiver, MPIElasticSource)


def _value(v: Any) -> Any:
    """A MATLAB value as Julia's MAT.jl hands it to IO.jl: 1x1 numeric arrays become scalars,
    struct arrays become lists of dicts, other arrays stay 2-D."""
    if isinstance(v, np.ndarray):
        if v.dtype == object:
            return [_value(x) for x in v.reshape(-1, order="F")]
        if v.dtype.names is not None:  # struct array (struct_as_record=True)
            out = []
            for rec in v.reshape(-1, order="F"):
                out.append({n: _value(rec[n]) for n in v.dtype.names})
            return out
        if v.size == 1 and v.dtype.kind in "biuf":
            return v.reshape(()).item()
        return v
    return v


def matread(filename: str) -> Dict[str, Any]:
    """matread(filename) of MAT.jl, as used by IO.jl: a dict of the file's variables."""
    import scipy.io as sio
    d = sio.loadmat(filename, squeeze_me=False, struct_as_record=True)
    out = {}
    for k, v in d.items():
        if k.startswith("__"):
            continue
        out[k] = _value(v)
    return out


def _field(x):
    """A struct field: object cells hold the stored array."""
    while isinstance(x, list) and len(x) == 1:
        x = x[0]
    return x


def safe_vec(a) -> np.ndarray:
    """safe_vec, src/IO.jl:163-169: a number becomes a one-element vector, an array its
    column-major vec."""
    a = _field(a)
    if np.isscalar(a) or (isinstance(a, np.ndarray) and a.ndim == 0):
        return np.array([a])
    return np.asarray(a).reshape(-1, order="F")


def _int(x) -> int:
    return int(np.asarray(_field(x)).reshape(-1)[0])


def _float(x) -> float:
    return float(np.asarray(_field(x)).reshape(-1)[0])


def _structs(v) -> List[Dict[str, Any]]:
    """The shots of a ``source`` / ``receiver`` variable: a struct array or a cell array of
    structs (the layout of the shipped example files) both give a flat list of dicts."""
    if isinstance(v, dict):
        return [v]
    if isinstance(v, np.ndarray):
        return [] if v.size == 0 else [v]
    out: List[Dict[str, Any]] = []
    for x in v:
        out.extend(_structs(x))
    return out


def _grid_kwargs(d, option):
    kw = dict(NX=_int(d["nx"]) - 2, NY=_int(d["ny"]) - 2, NSTEP=_int(d["nt"]),
              DELTAX=_float(d["dx"]), DELTAY=_float(d["dy"]), DELTAT=_float(d["dt"]))
    return kw


def load_params(filename: str, option: str = "Acoustic", **kwargs):
    """load_params(filename, option; kwargs...), src/IO.jl:4-27."""
    d = matread(filename)
    if option == "Acoustic":
        return AcousticPropagatorParams(**{**_grid_kwargs(d, option), **kwargs})
    if option == "Elastic":
        return ElasticPropagatorParams(**{**_grid_kwargs(d, option), **kwargs})
    if option == "MPIElastic":
        n = _int(d["node_per_cpu"])
        nx, ny = _int(d["nx"]), _int(d["ny"])
        return MPIElasticPropagatorParams(**{**dict(NX=nx, NY=ny, n=n, NSTEP=_int(d["nt"]),
                                                    DELTAT=_float(d["dt"]), DELTAX=_float(d["dx"]),
                                                    DELTAY=_float(d["dy"]), M=nx // n, N=ny // n),
                                             **kwargs})
    raise ValueError("option should be either Acoustic or Elastic!")


def _vec_column(vec) -> np.ndarray:
    v = np.asarray(_field(vec), dtype=np.float64)
    if v.ndim == 1 or v.shape[0] == 1:
        v = v.reshape(-1, 1)
    return v


def load_acoustic_source(filename: str) -> List[AcousticSource]:
    """load_acoustic_source, src/IO.jl:144-157: wavelets as (nt, nsrc) columns."""
    d = matread(filename)
    nt = _int(d["nt"])
    out = []
    for s in _structs(d["source"]):
        vec = _vec_column(s["vec"])
        if vec.shape[0] != nt:
            raise ValueError(f"source wavelet has {vec.shape[0]} samples, nt = {nt}")
        out.append(AcousticSource(safe_vec(s["ix"]).astype(np.int64),
                                  safe_vec(s["iy"]).astype(np.int64), vec))
    return out


def load_acoustic_receiver(filename: str) -> List[AcousticReceiver]:
    """load_acoustic_receiver, src/IO.jl:131-142."""
    d = matread(filename)
    return [AcousticReceiver(safe_vec(r["ix"]).astype(np.int64), safe_vec(r["iy"]).astype(np.int64))
            for r in _structs(d.get("receiver", []))]


def load_elastic_source(filename: str, use_mpi: bool = False, param=None):
    """load_elastic_source, src/IO.jl:69-94: wavelets as (nt, nsrc), transposed when stored
    the other way round."""
    d = matread(filename)
    nt = _int(d["nt"])
    out = []
    for s in _structs(d["source"]):
        vec = np.asarray(_field(s["vec"]), dtype=np.float64)
        if vec.ndim == 1:
            vec = vec.reshape(-1, 1)
        if vec.shape[0] != nt and vec.shape[1] == nt:
            vec = vec.T.copy()
        if vec.shape[0] != nt:
            raise ValueError(f"source wavelet has {vec.shape[0]} samples, nt = {nt}")
        ix, iy, ty = (safe_vec(s[k]).astype(np.int64) for k in ("ix", "iy", "type"))
        out.append(MPIElasticSource(param, ix, iy, ty, vec) if use_mpi else ElasticSource(ix, iy, ty, vec))
    return out


def load_elastic_receiver(filename: str, use_mpi: bool = False, param=None):
    """load_elastic_receiver, src/IO.jl:49-67."""
    d = matread(filename)
    out = []
    for r in _structs(d.get("receiver", [])):
        ix, iy, ty = (safe_vec(r[k]).astype(np.int64) for k in ("ix", "iy", "type"))
        out.append(MPIElasticReceiver(param, ix, iy, ty) if use_mpi else ElasticReceiver(ix, iy, ty))
    return out


def _device_model(a, dtype, trainable: bool):
    import torch
    t = torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=dtype, device="cuda")
    return t.requires_grad_(True) if trainable else t


def load_acoustic_model(filename: str, inv_vp: bool = False, mask=None, vp=None, dtype=None,
                        **kwargs):
    """load_acoustic_model(filename; inv_vp, mask, vp, kwargs...), src/IO.jl:105-129.

    Returns ``ap_fun(src) -> AcousticPropagatorSolver(param, src, vp)``.  As with the
    reference's TF variable, the model lives with the function: ``ap_fun.param``,
    ``ap_fun.variables`` (the trainable leaf tensors, empty unless ``inv_vp``) and
    ``ap_fun.velocity()``, the velocity the solver uses (``mask * vp_ + (1 - mask) * vp0``
    with a mask), rebuilt on every call so each evaluation has its own autograd graph."""
    import torch
    from .acoustic import AcousticPropagatorSolver
    dtype = dtype or torch.float64
    d = matread(filename)
    param = AcousticPropagatorParams(**{**_grid_kwargs(d, "Acoustic"), **kwargs})
    variables = []
    if vp is not None:
        fixed = vp if isinstance(vp, torch.Tensor) else _device_model(vp, dtype, False)
        velocity = lambda: fixed  # noqa: E731
    else:
        vp0 = _device_model(d["vp"], dtype, False)
        if not inv_vp:
            velocity = lambda: vp0  # noqa: E731
        else:
            var = vp0.clone().requires_grad_(True)
            variables.append(var)
            if mask is None:
                velocity = lambda: var  # noqa: E731
            else:
                m = torch.as_tensor(np.asarray(mask, np.float64), dtype=dtype, device="cuda")
                velocity = lambda: m * var + (1.0 - m) * vp0  # noqa: E731

    def ap_fun(src):
        return AcousticPropagatorSolver(param, src, velocity())

    ap_fun.param, ap_fun.variables, ap_fun.velocity = param, variables, velocity
    return ap_fun


def load_elastic_model(filename: str, inv_vp: bool = False, inv_vs: bool = False,
                       inv_rho: bool = False, dtype=None, **kwargs):
    """load_elastic_model(filename; inv_vp, inv_vs, inv_rho, kwargs...), src/IO.jl:29-47:
    parameters with vp_ref = mean(vp) and f0 = f0/2 from the file, (lambda, mu, rho) from
    compute_properties(vp, vs, rho).  Returns ``ep_fun(src)`` with ``.param``,
    ``.variables`` and ``.models()`` = (rho, lambda, mu), rebuilt on every call."""
    import torch
    from .elastic import ElasticPropagatorSolver
    from .utils import compute_properties
    dtype = dtype or torch.float64
    d = matread(filename)
    base = _grid_kwargs(d, "Elastic")
    base.update(vp_ref=float(np.mean(np.asarray(d["vp"], np.float64))), f0=_float(d["f0"]) / 2)
    param = ElasticPropagatorParams(**{**base, **kwargs})
    vp = _device_model(d["vp"], dtype, inv_vp)
    vs = _device_model(d["vs"], dtype, inv_vs)
    rho = _device_model(d["rho"], dtype, inv_rho)

    def models():
        lam, mu, rho_ = compute_properties(vp, vs, rho)
        return rho_, lam, mu

    def ep_fun(src):
        return ElasticPropagatorSolver(param, src, *models())

    ep_fun.param, ep_fun.models = param, models
    ep_fun.variables = [t for t, f in ((vp, inv_vp), (vs, inv_vs), (rho, inv_rho)) if f]
    return ep_fun
