"""GPU tests of the per-op drop-ins that no solver path exercises: AddSource forward and
backward (deps/CustomOps/SourceOps/AddSource.cpp:33-159), GetReceive (ReceiveOps/
GetReceive.cpp:10-46) and the FP32 variants of the index ops, the MPI subdomain step and the
one-step VJP.  Every call goes through the C-ABI of libadseismic_hip.so.

Oracles: the FP64 oracle's elastic step on a zero state (its source injection alone; the step
is pinned by the elastic probes and finite differences), its elastic adjoint's source gradient
(pinned by a finite difference on srcv), its GetReceive restatement, and the golden vectors of
the reference's compiled headers for the FP32 ops (FP32 rounding bars stated per test).
"""
import ctypes as C
import importlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NX, NY = 9, 7
N = (NX + 2) * (NY + 2)


def P(t):
    return C.c_void_p(t.data_ptr())


def H(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


def D(a, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


@pytest.fixture(scope="module")
def L(ads):
    ads.load_library()
    return importlib.import_module("adseismic_amd._lib")


# sources: every type, a duplicated (cell, type) pair, a cell on the padding ring, and two
# out-of-range types that the CPU op skips (AddSource.cpp:79-81)
SRCI = [5, 5, 3, 8, 1, 6, 4, 2, 7]
SRCJ = [4, 4, 6, 2, 3, NY + 2, 5, 2, 6]
STYP = [2, 2, 0, 1, 3, 4, 3, 5, -1]


def _oracle_sources(orc, srcv):
    """The oracle's elastic step on a zero state adds only the sources: (5, N) in the oracle's
    field order vx, vy, sxx, syy, sxy."""
    rho = np.full((NX + 2, NY + 2), 2000.0)
    lam, mu = np.full_like(rho, 8e9), np.full_like(rho, 4e9)
    a0, b0 = np.zeros(2 * NX), np.ones(2 * NX)
    c0, d0 = np.zeros(2 * NY), np.ones(2 * NY)
    f, _ = orc.elastic_step(NX, NY, 10.0, 10.0, 1e-3, rho, lam, mu, a0, b0, c0, d0,
                            np.zeros(5 * N), np.zeros(8 * N), SRCI, SRCJ, STYP, srcv)
    return f.reshape(5, N)


OP_ORDER = (2, 3, 4, 0, 1)  # the op's field order sxx, syy, sxy, vx, vy in oracle indices


def _add_source(L, lib, ins, outs, srcv):
    si, sip = H(SRCI)
    sj, sjp = H(SRCJ)
    st, stp = H(STYP)
    L.check(lib.ads_add_source_f64(*[P(t) for t in ins], sip, sjp, stp, len(SRCI), NX, NY,
                                   P(srcv), *[P(t) for t in outs], None))


def test_add_source_forward(orc, ads, L, cuda):
    lib = ads.load_library()
    rng = np.random.default_rng(1)
    srcv = rng.standard_normal(len(SRCI))
    want = _oracle_sources(orc, srcv)
    sv = D(srcv)
    # zero fields: bit-exact against the oracle's injection
    ins = [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(5)]
    outs = [torch.full((N,), np.nan, dtype=torch.float64, device="cuda") for _ in range(5)]
    _add_source(L, lib, ins, outs, sv)
    torch.cuda.synchronize()
    for q, o in zip(OP_ORDER, outs):
        np.testing.assert_array_equal(o.cpu().numpy(), want[q], err_msg=f"field {q}")
    # random fields: out = in + injection (copy semantics, AddSource.cpp:48-55), also in place
    f0 = rng.standard_normal((5, N))
    ins = [D(f0[q]) for q in OP_ORDER]
    outs = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(5)]
    _add_source(L, lib, ins, outs, sv)
    inplace = [t.clone() for t in ins]
    _add_source(L, lib, inplace, inplace, sv)
    torch.cuda.synchronize()
    for k, q in enumerate(OP_ORDER):
        np.testing.assert_allclose(outs[k].cpu().numpy(), f0[q] + want[q], rtol=1e-15, atol=1e-15)
        np.testing.assert_array_equal(inplace[k].cpu().numpy(), outs[k].cpu().numpy())
        np.testing.assert_array_equal(ins[k].cpu().numpy(), f0[q])  # inputs untouched


def test_add_source_rejects_a_source_outside_the_grid(ads, L, cuda):
    lib = ads.load_library()
    f = [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(5)]
    si, sip = H([NX + 3])
    sj, sjp = H([2])
    st, stp = H([0])
    rc = lib.ads_add_source_f64(*[P(t) for t in f], sip, sjp, stp, 1, NX, NY, P(D([1.0])),
                                *[P(t) for t in f], None)
    assert rc == L.ADS_ERR_INDEX
    assert b"outside the padded grid" in lib.ads_last_error()


def test_add_source_backward(orc, ads, L, cuda):
    """grad_srcv[s] = the cotangent of the source's field at its cell; the oracle's elastic
    adjoint gives the same for a one-step record (level 1 takes srcv row 0)."""
    lib = ads.load_library()
    rng = np.random.default_rng(2)
    rho = 2000.0 + rng.random((NX + 2, NY + 2))
    lam = 8e9 * (1 + 0.1 * rng.random((NX + 2, NY + 2)))
    mu = 4e9 * (1 + 0.1 * rng.random((NX + 2, NY + 2)))
    pml = orc.elastic_pml(NX, NY, 10.0, 10.0, 1e-3, 2, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2500.0)
    srcv = rng.standard_normal((1, len(SRCI)))
    hist, mem = orc.elastic_forward(NX, NY, 1, 10.0, 10.0, 1e-3, rho, lam, mu, pml, SRCI, SRCJ,
                                    STYP, srcv)
    g = np.zeros((5, 2, N))
    g[:, 1] = rng.standard_normal((5, N))
    *_, gsrc = orc.elastic_adjoint(NX, NY, 1, 10.0, 10.0, 1e-3, rho, lam, mu, pml, hist, mem, g,
                                   SRCI, SRCJ, STYP)
    gin = [D(g[q, 1]) for q in OP_ORDER]
    out = torch.full((len(SRCI),), np.nan, dtype=torch.float64, device="cuda")
    si, sip = H(SRCI)
    sj, sjp = H(SRCJ)
    st, stp = H(STYP)
    L.check(lib.ads_add_source_grad_f64(*[P(t) for t in gin], sip, sjp, stp, len(SRCI), NX, NY,
                                        P(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got, gsrc[0])
    assert got[0] == got[1]                    # duplicated (cell, type)
    assert got[7] == 0.0 and got[8] == 0.0     # out-of-range types get no cotangent


def test_get_receive(orc, ads, L, cuda):
    """GetReceive over histories of nt levels, layout (nrcv, nt); duplicated receivers, every
    type, a ring cell, and rows of out-of-range types left as they were."""
    lib = ads.load_library()
    rng = np.random.default_rng(3)
    nt = 6
    hist = rng.standard_normal((5, nt, NX + 2, NY + 2))
    rcvi = [2, 2, 5, NX + 2, 3, 7, 4, 6]
    rcvj = [3, 3, 4, 1, NY + 1, 2, 5, 6]
    rtyp = [0, 0, 1, 2, 3, 4, 7, 2]
    want = orc.elastic_receive(hist, rcvi, rcvj, rtyp, NX, NY)
    dev = [D(hist[q].reshape(nt, N)) for q in range(5)]   # vx, vy, sxx, syy, sxy
    out = torch.zeros(len(rcvi) * nt, dtype=torch.float64, device="cuda")
    ri, rip = H(rcvi)
    rj, rjp = H(rcvj)
    rt, rtp = H(rtyp)
    L.check(lib.ads_get_receive_f64(*[P(t) for t in dev], nt, rip, rjp, rtp, len(rcvi), NX, NY,
                                    P(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(len(rcvi), nt)
    np.testing.assert_array_equal(got, want)
    assert np.all(got[6] == 0.0)
    np.testing.assert_array_equal(got[0], got[1])


def test_index_ops_f32(ads, golden, L, cuda):
    """FP32 gather / scatter_nd select values (exactly the golden values rounded to FP32);
    scatter_add accumulates in FP32 in index order (within FP32 rounding of the golden)."""
    lib = ads.load_library()
    g = golden("index_ops.npz")
    v, upd = D(g["v"], torch.float32), D(g["upd"], torch.float32)
    ii = torch.as_tensor(g["ii"], device="cuda")
    n = len(g["ii"])
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    L.check(lib.ads_gather_f32(P(v), 50, P(ii), n, P(out), None))
    sa = torch.empty(50, dtype=torch.float32, device="cuda")
    L.check(lib.ads_scatter_add_f32(P(v), 50, P(ii), P(upd), n, P(sa), None))
    snd = torch.empty(50, dtype=torch.float32, device="cuda")
    L.check(lib.ads_scatter_nd_f32(P(ii), P(upd), n, 50, P(snd), None))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), g["gather"].astype(np.float32))
    np.testing.assert_array_equal(snd.cpu().numpy(), g["scatter_nd"].astype(np.float32))
    scale = np.abs(g["v"]).max() + np.abs(g["upd"]).sum()
    np.testing.assert_allclose(sa.cpu().numpy(), g["scatter_add"], rtol=0, atol=4e-7 * scale)


def test_mpi_one_step_f32(ads, golden, L, cuda):
    lib = ads.load_library()
    g = golden("mpi_acoustic_one_step.npz")
    n = int(g["n"])
    dt, hx, hy = (float(x) for x in g["scalars"])
    ins = [D(g[k], torch.float32) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")]
    outs = [torch.empty(n * n, dtype=torch.float32, device="cuda") for _ in range(3)]
    L.check(lib.ads_mpi_acoustic_one_step_f32(*[P(t) for t in ins], dt, hx, hy, n,
                                              *[P(t) for t in outs], None))
    torch.cuda.synchronize()
    for o, k in zip(outs, ("u", "phiout", "psiout")):
        ref = g[k]
        err = np.linalg.norm(o.cpu().numpy() - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err < 1e-6, (k, err)


@pytest.mark.parametrize("name", ["10x10", "7x12"])
def test_one_step_vjp_f32(ads, golden, cuda, name):
    """AcousticOneStepGrad in FP32 against the reference header's FP64 VJP (relative L2 1e-5,
    FP32 rounding of a handful of operations per cell)."""
    g = golden("acoustic_one_step.npz")
    NXg, NYg = (int(x) for x in g[f"{name}/dims"])
    dt, hx, hy = (float(x) for x in g[f"{name}/scalars"])
    ins = [D(g[f"{name}/{k}"], torch.float32) for k in ("w", "wold", "phi", "psi", "sigma", "tau", "c")]
    for t in ins:
        t.requires_grad_(True)
    u, po, so = ads.acoustic_one_step(*ins, dt, hx, hy, NXg, NYg)
    cot = [D(g[f"{name}/{k}"], torch.float32) for k in ("grad_u", "grad_phiout", "grad_psiout")]
    grads = torch.autograd.grad((u, po, so), [ins[0], ins[1], ins[2], ins[3], ins[6]], cot)
    for gr, k in zip(grads, ("grad_w", "grad_wold", "grad_phi", "grad_psi", "grad_c")):
        ref = g[f"{name}/{k}"]
        err = np.linalg.norm(gr.double().cpu().numpy() - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err < 1e-5, (k, err)
