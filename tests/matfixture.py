"""Writes MAT files with the layout of the reference's example models (examples/nn_fwi/models/
*.mat: nx, ny, nt, dx, dy, dt, f0, vp [, vs, rho, mask], 1 x nshots struct arrays ``source``
(ix, iy, vec [, type]) and ``receiver`` (ix, iy, type)) from the committed fixtures.

The fixtures hold the Marmousi2 models, the first source's wavelet and the first shot's
receiver spread; the replica gives every shot that wavelet and that spread, as the shipped
files do (all 8 shots of marmousi2-model-*.mat and BP-model-*.mat share both)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def marmousi_fixture():
    w = np.load(os.path.join(GOLDEN, "wavelet_fixtures.npz"), allow_pickle=False)
    m = np.load(os.path.join(GOLDEN, "marmousi2_models.npz"), allow_pickle=False)
    nt, dt, f0, nx, ny = w["marmousi2/nt_dt_f0_nx_ny"]
    return dict(nt=int(nt), dt=float(dt), f0=float(f0), nx=int(nx), ny=int(ny),
                dx=float(m["dx_dy"][0]), dy=float(m["dx_dy"][1]),
                vec=w["marmousi2/vec"], src_ix=w["marmousi2/src_ix"], src_iy=w["marmousi2/src_iy"],
                rcv_ix=w["marmousi2/rcv_ix"], rcv_iy=w["marmousi2/rcv_iy"],
                rcv_type=w["marmousi2/rcv_type"], vp_true=m["vp_true"].astype(np.float64),
                vp_smooth=m["vp_smooth"].astype(np.float64), mask=m["mask"].astype(np.float64))


def _structs(rows, fields, cells):
    """1 x n struct array, or (cells=True, the layout of the shipped files) a 1 x n cell array
    of 1 x 1 structs."""
    dt = [(f, object) for f in fields]
    if not cells:
        arr = np.empty((1, len(rows)), dtype=dt)
        for k, row in enumerate(rows):
            for f in fields:
                arr[0, k][f] = row[f]
        return arr
    arr = np.empty((1, len(rows)), dtype=object)
    for k, row in enumerate(rows):
        one = np.empty((1, 1), dtype=dt)
        for f in fields:
            one[0, 0][f] = row[f]
        arr[0, k] = one
    return arr


def write_mat(path, fx, vp, elastic=False, mask=None, vec_as_row=False, nshots=None, cells=True):
    import scipy.io as sio
    ns = len(fx["src_ix"]) if nshots is None else nshots
    vec = fx["vec"].reshape(1, -1) if vec_as_row else fx["vec"].reshape(-1, 1)
    srcs, rcvs = [], []
    for s in range(ns):
        src = {"ix": np.array([[fx["src_ix"][s]]], np.float64),
               "iy": np.array([[fx["src_iy"][s]]], np.float64), "vec": vec}
        rcv = {"ix": fx["rcv_ix"].reshape(1, -1).astype(np.float64),
               "iy": fx["rcv_iy"].reshape(1, -1).astype(np.float64),
               "type": fx["rcv_type"].reshape(1, -1).astype(np.float64)}
        if elastic:
            src["type"] = np.array([[2.0]])
        srcs.append(src)
        rcvs.append(rcv)
    d = {"nx": np.array([[fx["nx"]]], np.int64), "ny": np.array([[fx["ny"]]], np.int64),
         "nt": np.array([[fx["nt"]]], np.int64), "dx": np.array([[fx["dx"]]]),
         "dy": np.array([[fx["dy"]]]), "dt": np.array([[fx["dt"]]]),
         "f0": np.array([[fx["f0"]]]), "vp": vp,
         "source": _structs(srcs, ["ix", "iy", "vec"] + (["type"] if elastic else []), cells),
         "receiver": _structs(rcvs, ["ix", "iy", "type"], cells)}
    if elastic:
        d["vs"] = vp / 1.73
        d["rho"] = 310.0 * vp ** 0.25
    if mask is not None:
        d["mask"] = mask
    sio.savemat(path, d)
    return path
