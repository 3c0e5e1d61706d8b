"""ctypes binding of libadseismic_hip.so (the C-ABI declared in include/adseismic.h).

The library is the only compute path: if it is missing or cannot be loaded this module raises
immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libadseismic_hip.so")

i64 = C.c_int64
i32 = C.c_int32
f64 = C.c_double
vp = C.c_void_p
i64p = C.POINTER(C.c_int64)

ADS_OK, ADS_ERR_ARG, ADS_ERR_SHAPE, ADS_ERR_INDEX, ADS_ERR_CFL, ADS_ERR_HIP, ADS_ERR_WORKSPACE, \
    ADS_ERR_NONFINITE = range(8)
ADS_F32, ADS_F64 = 0, 1


class ADSError(RuntimeError):
    """Error returned by the HIP library; ``code`` is the ADS_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[ads error {code}] {msg}")
        self.code = code


class AcousticParamsC(C.Structure):
    _fields_ = [("NX", i64), ("NY", i64), ("NSTEP", i64),
                ("DELTAX", f64), ("DELTAY", f64), ("DELTAT", f64),
                ("USE_PML_XMIN", i32), ("USE_PML_XMAX", i32), ("USE_PML_YMIN", i32),
                ("USE_PML_YMAX", i32), ("NPOINTS_PML", i64), ("Rcoef", f64), ("vp_ref", f64),
                ("PropagatorKernel", i32), ("square_velocity", i32),
                ("origin_x", f64), ("origin_y", f64), ("NX_global", i64), ("NY_global", i64)]


class ElasticParamsC(C.Structure):
    _fields_ = [("NX", i64), ("NY", i64), ("NSTEP", i64),
                ("DELTAX", f64), ("DELTAY", f64), ("DELTAT", f64), ("f0", f64), ("vp_ref", f64),
                ("USE_PML_XMIN", i32), ("USE_PML_XMAX", i32), ("USE_PML_YMIN", i32),
                ("USE_PML_YMAX", i32), ("NPOINTS_PML", i64), ("NPOWER", f64),
                ("K_MAX_PML", f64), ("ALPHA_MAX_PML", f64), ("Rcoef", f64), ("subdomain", i32),
                ("origin_x", f64), ("origin_y", f64), ("NX_global", i64), ("NY_global", i64)]


class GeometryC(C.Structure):
    _fields_ = [("nshots", i64), ("nstep", i64p), ("src_ptr", i64p), ("srci", i64p),
                ("srcj", i64p), ("srctype", i64p), ("rcv_ptr", i64p), ("rcvi", i64p),
                ("rcvj", i64p), ("rcvtype", i64p)]


# name -> (restype, argtypes)
_SIGS = {
    "ads_last_error": (C.c_char_p, []),
    "ads_version": (C.c_int, []),
    "ads_acoustic_one_step_f64": (C.c_int, [vp] * 7 + [f64] * 3 + [i64] * 2 + [vp] * 3 + [vp]),
    "ads_acoustic_one_step_f32": (C.c_int, [vp] * 7 + [f64] * 3 + [i64] * 2 + [vp] * 3 + [vp]),
    "ads_acoustic_one_step_grad_f64": (C.c_int, [vp] * 10 + [f64] * 3 + [i64] * 2 + [vp] * 5 + [vp]),
    "ads_acoustic_one_step_grad_f32": (C.c_int, [vp] * 10 + [f64] * 3 + [i64] * 2 + [vp] * 5 + [vp]),
    "ads_mpi_acoustic_one_step_f64": (C.c_int, [vp] * 7 + [f64] * 3 + [i64] + [vp] * 3 + [vp]),
    "ads_mpi_acoustic_one_step_f32": (C.c_int, [vp] * 7 + [f64] * 3 + [i64] + [vp] * 3 + [vp]),
    "ads_gather_f64": (C.c_int, [vp, i64, vp, i64, vp, vp]),
    "ads_gather_f32": (C.c_int, [vp, i64, vp, i64, vp, vp]),
    "ads_scatter_add_f64": (C.c_int, [vp, i64, vp, vp, i64, vp, vp]),
    "ads_scatter_add_f32": (C.c_int, [vp, i64, vp, vp, i64, vp, vp]),
    "ads_scatter_nd_f64": (C.c_int, [vp, vp, i64, i64, vp, vp]),
    "ads_scatter_nd_f32": (C.c_int, [vp, vp, i64, i64, vp, vp]),
    "ads_add_source_f64": (C.c_int, [vp] * 5 + [i64p] * 3 + [i64] * 3 + [vp] * 6 + [vp]),
    "ads_add_source_noalloc_f64": (C.c_int, [vp] * 5 + [i64p] * 3 + [i64] * 3 + [vp] * 6 + [vp]),
    "ads_add_source_grad_f64": (C.c_int, [vp] * 5 + [i64p] * 3 + [i64] * 3 + [vp, vp]),
    "ads_get_receive_f64": (C.c_int, [vp] * 5 + [i64] + [i64p] * 3 + [i64] * 3 + [vp, vp]),
    "ads_acoustic_pml_profiles": (C.c_int, [C.POINTER(AcousticParamsC), vp, vp]),
    "ads_acoustic_check_cfl": (C.c_int, [C.POINTER(AcousticParamsC), f64]),
    "ads_elastic_pml_params": (C.c_int, [C.POINTER(ElasticParamsC)] + [vp] * 6),
    "ads_acoustic_plan_create": (C.c_int, [C.POINTER(AcousticParamsC), C.POINTER(GeometryC), C.c_int,
                                           C.POINTER(vp)]),
    "ads_acoustic_plan_destroy": (C.c_int, [vp]),
    "ads_acoustic_plan_sizes": (C.c_int, [vp, i64p, i64p, i64p, i64p]),
    "ads_acoustic_plan_workspace_size": (C.c_size_t, [vp, i64]),
    "ads_acoustic_plan_choose_interval": (i64, [vp, C.c_size_t]),
    "ads_acoustic_plan_set_snapshots": (C.c_int, [vp, i64]),
    "ads_acoustic_plan_snapshots": (i64, [vp]),
    "ads_acoustic_plan_choose_schedule": (C.c_int, [vp, C.c_size_t, i64p, i64p]),
    "ads_checkpoint_schedule": (C.c_int, [i64, i64, i64p, i64, i64p]),
    "ads_acoustic_plan_forward": (C.c_int, [vp, vp, vp, vp, vp, C.c_size_t, i64, vp]),
    "ads_acoustic_plan_backward": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_size_t, i64, vp]),
    "ads_acoustic_plan_gradient": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, C.c_size_t, i64, vp]),
    "ads_acoustic_plan_autotune": (C.c_int, [vp, vp, vp, vp, C.c_size_t, i64, C.c_int,
                                             C.POINTER(C.c_int), vp]),
    "ads_acoustic_plan_set_shape": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ads_acoustic_plan_set_graph": (C.c_int, [vp, C.c_int]),
    "ads_acoustic_plan_set_blocking": (C.c_int, [vp, C.c_int]),
    "ads_acoustic_plan_set_frame_rows": (C.c_int, [vp, C.c_int]),
    "ads_acoustic_plan_blocking": (C.c_int, [vp]),
    "ads_acoustic_plan_history_scratch_size": (C.c_size_t, [vp]),
    "ads_acoustic_plan_forward_history": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "ads_acoustic_plan_backward_history": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "ads_elastic_plan_create": (C.c_int, [C.POINTER(ElasticParamsC), C.POINTER(GeometryC), C.c_int,
                                          C.POINTER(vp)]),
    "ads_elastic_plan_destroy": (C.c_int, [vp]),
    "ads_elastic_plan_sizes": (C.c_int, [vp, i64p, i64p, i64p, i64p]),
    "ads_elastic_plan_workspace_size": (C.c_size_t, [vp, i64]),
    "ads_elastic_plan_choose_interval": (i64, [vp, C.c_size_t]),
    "ads_elastic_plan_set_adjoint_kernels": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "ads_elastic_plan_set_forward_kernels": (C.c_int, [vp, C.c_int, C.c_int]),
    "ads_elastic_plan_set_snapshots": (C.c_int, [vp, i64]),
    "ads_elastic_plan_snapshots": (i64, [vp]),
    "ads_elastic_plan_choose_schedule": (C.c_int, [vp, C.c_size_t, i64p, i64p]),
    "ads_elastic_plan_forward": (C.c_int, [vp] * 6 + [vp, C.c_size_t, i64, vp]),
    "ads_elastic_plan_backward": (C.c_int, [vp] * 9 + [vp, C.c_size_t, i64, vp]),
    "ads_elastic_plan_gradient": (C.c_int, [vp] * 12 + [C.c_size_t, i64, vp]),
    "ads_elastic_plan_backward_dense": (C.c_int, [vp] * 9 + [vp, C.c_size_t, i64, vp]),
    "ads_elastic_plan_history_scratch_size": (C.c_size_t, [vp]),
    "ads_elastic_forward_history": (C.c_int, [vp] * 8),
    "ads_profile_begin": (C.c_int, []),
    "ads_profile_begin_sampled": (C.c_int, [C.c_int]),
    "ads_profile_end": (C.c_int, []),
    "ads_profile_query": (C.c_int, [C.c_int, C.POINTER(f64), i64p, C.POINTER(f64)]),
}

_lib = None


def load():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
            "`python adseismic.jl_amd/build.py` (the HIP library is the only compute path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols():
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != ADS_OK:
        msg = load().ads_last_error()
        raise ADSError(rc, msg.decode() if msg else "")


# Which call last filled a plan workspace.  A forward leaves checkpoints, the last segment's
# levels and the model copy in the workspace for the backward that follows; any other call
# on the same buffer (another forward, a gradient, an autotune, the backward itself, which
# recomputes segments over the slot ring, or another plan sharing the buffer) overwrites them.
_ws_state = {}
_ws_count = itertools.count(1)


def mark_workspace(ws) -> int:
    """Records that the caller has just (re)filled ``ws``; returns its token."""
    tok = next(_ws_count)
    _ws_state[ws.data_ptr()] = tok
    return tok


def workspace_holds(ws, token) -> bool:
    """True while ``ws`` still holds the state of the call that returned ``token``."""
    return ws is not None and token is not None and _ws_state.get(ws.data_ptr()) == token


def available_device_bytes() -> int:
    """Device memory a new workspace can use: the free HBM plus what torch's caching allocator
    holds but does not use (released workspaces of earlier plans stay cached)."""
    import torch
    free, _ = torch.cuda.mem_get_info()
    return int(free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated())


def device_bytes(n: int):
    """A uint8 device buffer of n bytes for a workspace.  A large request that torch's cache
    cannot place (cached blocks of earlier, smaller workspaces) is retried once after the
    cache is returned to the device."""
    import torch
    try:
        return torch.empty(int(n), dtype=torch.uint8, device="cuda")
    except torch.OutOfMemoryError:
        torch.cuda.empty_cache()
        return torch.empty(int(n), dtype=torch.uint8, device="cuda")
