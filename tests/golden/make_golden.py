"""Writes the golden vectors under tests/golden/ (run in a container that has the reference tree).

1. acoustic_one_step.npz / mpi_acoustic_one_step.npz / index_ops.npz: seeded inputs and the
   outputs of the reference's own header-only CPU kernels (deps/CustomOps/*/ *.h), compiled
   where they lie by oracle/Makefile (`make -C oracle ref`) -- AcousticOneStepCpuForward/
   Backward, MpiAcousticOneStepCpuForward, GatherOps/ScatterAddOps/ScatterNdOps forward.
2. wavelet_fixtures.npz: the source wavelets, sampling and geometry the reference ships in
   examples/nn_fwi/models/{marmousi2,BP}-model-true.mat (read with scipy.io.loadmat; no
   pickle), used to pin the Ricker restatement and the index conventions.
3. marmousi2_models.npz: the velocity models (true, smooth), the update mask and dx, dy of
   examples/nn_fwi/models/marmousi2-model-{true,smooth}.mat (models as float32), for the MAT loader and
   FWI-driver tests (which rebuild a .mat of the same layout from them).

``python make_golden.py models`` writes only item 3.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_longlong)


def P(a):
    return a.ctypes.data_as(dp)


def models():
    import scipy.io as sio
    t = sio.loadmat(os.path.join(REF, "examples/nn_fwi/models/marmousi2-model-true.mat"))
    m = sio.loadmat(os.path.join(REF, "examples/nn_fwi/models/marmousi2-model-smooth.mat"))
    np.savez_compressed(os.path.join(HERE, "marmousi2_models.npz"),
                        vp_true=np.asarray(t["vp"], np.float32), vp_smooth=np.asarray(m["vp"], np.float32),
                        mask=np.asarray(m["mask"], np.float32), dx_dy=np.array([float(t["dx"][0, 0]),
                                                                                float(t["dy"][0, 0])]))


def main():
    models()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    R = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libadseismic_ref.so"))
    rng = np.random.default_rng(233)
    dt, hx, hy = 0.1, 0.1, 0.1   # the reference gradtest's values (AcousticOneStep/gradtest.jl:20-22)
    cases = {}
    for name, (NX, NY) in {"10x10": (10, 10), "7x12": (7, 12)}.items():
        N = (NX + 2) * (NY + 2)
        ins = {k: rng.random(N) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")}
        u, po, so = np.empty(N), np.empty(N), np.empty(N)
        R.ref_acoustic_one_step_forward(*[P(ins[k]) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")],
                                        C.c_double(dt), C.c_double(hx), C.c_double(hy),
                                        C.c_longlong(NX), C.c_longlong(NY), P(u), P(po), P(so))
        gu, gp, gs = rng.random(N), rng.random(N), rng.random(N)
        outs = [np.zeros(N) for _ in range(5)]  # the op zeroes its outputs first
        R.ref_acoustic_one_step_backward(*[P(o) for o in outs], P(gu), P(gp), P(gs),
                                         *[P(ins[k]) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")],
                                         C.c_double(dt), C.c_double(hx), C.c_double(hy),
                                         C.c_longlong(NX), C.c_longlong(NY), P(u), P(po), P(so))
        for k, v in ins.items():
            cases[f"{name}/{k}"] = v
        cases.update({f"{name}/u": u, f"{name}/phiout": po, f"{name}/psiout": so,
                      f"{name}/grad_u": gu, f"{name}/grad_phiout": gp, f"{name}/grad_psiout": gs,
                      f"{name}/grad_w": outs[0], f"{name}/grad_wold": outs[1],
                      f"{name}/grad_phi": outs[2], f"{name}/grad_psi": outs[3],
                      f"{name}/grad_c": outs[4], f"{name}/dims": np.array([NX, NY]),
                      f"{name}/scalars": np.array([dt, hx, hy])})
    np.savez_compressed(os.path.join(HERE, "acoustic_one_step.npz"), **cases)

    # MPI subdomain step: (n+2)^2 halo'd inputs, n*n c and outputs
    n = 8
    Np = (n + 2) ** 2
    ins = {k: rng.random(Np) for k in ("w", "wold", "phi", "psi", "sigma", "tau")}
    ins["c"] = rng.random(n * n)
    u, po, so = np.empty(n * n), np.empty(n * n), np.empty(n * n)
    R.ref_mpi_acoustic_one_step_forward(*[P(ins[k]) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")],
                                        C.c_double(0.05), C.c_double(0.2), C.c_double(0.3),
                                        C.c_longlong(n), C.c_longlong(n), P(u), P(po), P(so))
    np.savez_compressed(os.path.join(HERE, "mpi_acoustic_one_step.npz"), **ins, u=u, phiout=po,
                        psiout=so, n=np.array(n), scalars=np.array([0.05, 0.2, 0.3]))

    # index ops (1-based), duplicates included
    v = rng.random(50)
    ii = np.array([3, 1, 50, 7, 7, 12, 3], dtype=np.int64)
    out = np.empty(len(ii))
    R.ref_gather_forward(P(out), P(v), ii.ctypes.data_as(ip), C.c_int(len(ii)))
    upd = rng.random(len(ii))
    sa = np.empty(50)
    R.ref_scatter_add_forward(P(sa), P(v), ii.ctypes.data_as(ip), P(upd), C.c_int(50), C.c_int(len(ii)))
    snd = np.zeros(50)
    R.ref_scatter_nd_forward(P(snd), ii.ctypes.data_as(ip), P(upd), C.c_int(len(ii)))
    np.savez_compressed(os.path.join(HERE, "index_ops.npz"), v=v, ii=ii, upd=upd, gather=out,
                        scatter_add=sa, scatter_nd=snd)

    # wavelets / geometry shipped with the reference's examples
    import scipy.io as sio
    fx = {}
    for tag, f in (("marmousi2", "examples/nn_fwi/models/marmousi2-model-true.mat"),
                   ("BP", "examples/nn_fwi/models/BP-model-true.mat")):
        d = sio.loadmat(os.path.join(REF, f))
        src, rcv = d["source"][0], d["receiver"][0]
        fx[f"{tag}/vec"] = np.asarray(src[0]["vec"][0, 0], np.float64).ravel()
        fx[f"{tag}/src_ix"] = np.array([int(np.ravel(s["ix"][0, 0])[0]) for s in src], np.int64)
        fx[f"{tag}/src_iy"] = np.array([int(np.ravel(s["iy"][0, 0])[0]) for s in src], np.int64)
        fx[f"{tag}/rcv_ix"] = np.asarray(rcv[0]["ix"][0, 0], np.int64).ravel()
        fx[f"{tag}/rcv_iy"] = np.asarray(rcv[0]["iy"][0, 0], np.int64).ravel()
        fx[f"{tag}/rcv_type"] = np.asarray(rcv[0]["type"][0, 0], np.int64).ravel()
        fx[f"{tag}/nt_dt_f0_nx_ny"] = np.array([float(d["nt"][0, 0]), float(d["dt"][0, 0]),
                                                float(d["f0"][0, 0]), float(d["nx"][0, 0]),
                                                float(d["ny"][0, 0])])
    np.savez_compressed(os.path.join(HERE, "wavelet_fixtures.npz"), **fx)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    sys.exit(models() if sys.argv[1:] == ["models"] else main())
