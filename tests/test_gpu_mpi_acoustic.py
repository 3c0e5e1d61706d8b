"""GPU tests of the MPIAcoustic subdomain path (MPIAcoustic.jl) through the HIP library.

1. MPIAcousticPropagatorSolver on one block of a 2 x 2 split against the FP64 oracle run on
   the block's own (n+2)^2 grid: the block's PML profile from the oracle's pinned closed form
   at the block origin, c not squared, sources/receivers mapped from the global unpadded grid;
   histories and traces 1e-12, d loss / d c 1e-10 (relative L2), both phi/psi variants.
2. Decomposition invariance, the reference's own check (test/verify_forward.jl:32-90): a
   2 x 2 tiling of the MpiAcousticOneStep op with exact halo copies between steps equals the
   full-grid AcousticOneStep op bit for bit in FP64 (the two reference headers evaluate the
   same expression in the same order, and both HIP ops are bit-exact to their headers).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def T(x, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")


def _setup(ads, orc, kernel, rank=2, n=24, NSTEP=90):
    p = ads.MPIAcousticPropagatorParams(NX=2 * n, NY=2 * n, n=n, NSTEP=NSTEP, DELTAX=10.0,
                                        DELTAY=8.0, DELTAT=1e-3, NPOINTS_PML=6, vp_ref=2000.0,
                                        PropagatorKernel=kernel)
    ads.compute_PML_Params_mpi_acoustic(p, rank=rank, size=4)
    rng = np.random.default_rng(11 + rank)
    c = (1800.0 + 500.0 * rng.random((n, n))) ** 2          # the unsquared model of the block
    w = orc.ricker(NSTEP, 6.0, 14.0, 1e6)
    # global 1-based sources on the unpadded grid: two in block (2, 1), one elsewhere
    src = ads.MPIAcousticSource(p, [n + 6, 2 * n - 3, 3], [5, 12, 3],
                                np.stack([w, -0.5 * w, w], 1))
    ri = np.arange(1, 2 * n + 1)
    rcv = ads.MPIAcousticReceiver(p, ri, np.full(len(ri), 4))
    return p, c, src, rcv


def _oracle(orc, p, c, src, rcv, cobs=None, obs=None):
    """The block on its own (n+2)^2 grid: profile at the block origin, C = c (padded ring 0),
    1-based padded positions = block index + 1."""
    n = p.n
    ox, oy = p.origin()
    sx = orc.acoustic_pml_axis(n + 2, p.DELTAX, ox, p.NX, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    ty = orc.acoustic_pml_axis(n + 2, p.DELTAY, oy, p.NY, p.NPOINTS_PML, p.Rcoef, p.vp_ref)
    Sx = np.repeat(sx[:, None], n + 2, 1)
    Sy = np.repeat(ty[None, :], n + 2, 0)
    k = 0 if p.PropagatorKernel == 0 else 1
    C2 = np.pad(c if cobs is None else cobs, 1)
    u, ph, ps = orc.acoustic_forward(k, n, n, p.NSTEP, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy, C2,
                                     src.srci + 1, src.srcj + 1, src.srcv)
    tr = orc.acoustic_receive(u, rcv.rcvi + 1, rcv.rcvj + 1, n)
    if obs is None:
        return u, tr
    gu = np.zeros((p.NSTEP + 1, (n + 2) ** 2))
    np.add.at(gu, (slice(None), orc.flat_index(rcv.rcvi + 1, rcv.rcvj + 1, n)), 2.0 * (tr - obs))
    gC = orc.acoustic_adjoint(k, n, n, p.NSTEP, p.DELTAT, p.DELTAX, p.DELTAY, Sx, Sy, C2, u, ph,
                              ps, gu)
    return u, tr, float(np.sum((tr - obs) ** 2)), gC[1:n + 1, 1:n + 1]


@pytest.mark.parametrize("kernel", [0, 1])
def test_subdomain_solver_against_oracle(orc, ads, cuda, kernel):
    p, c, src, rcv = _setup(ads, orc, kernel)
    assert (p.II, p.JJ) == (2, 1) and len(src.srci) == 2 and len(rcv.rcvi) == p.n
    np.testing.assert_array_equal(src.srci, [6, p.n - 3])
    _, obs = _oracle(orc, p, c, src, rcv, cobs=c * 1.04)
    u_ref, tr_ref, L_ref, g_ref = _oracle(orc, p, c, src, rcv, obs=obs)
    ct = T(c).requires_grad_(True)
    ap = ads.MPIAcousticPropagatorSolver(p, src, ct)
    assert tuple(ap.u.shape) == (p.NSTEP + 1, p.n, p.n)
    assert rel(ap.u.detach().cpu(), u_ref[:, 1:p.n + 1, 1:p.n + 1]) < 1e-12
    rr = ads.MPISimulatedObservation_acoustic(ap, rcv)
    assert rel(rr.detach().cpu(), tr_ref) < 1e-12
    loss = torch.sum((rr - T(obs)) ** 2)
    loss.backward()
    assert loss.item() == pytest.approx(L_ref, rel=1e-10)
    assert rel(ct.grad.cpu(), g_ref) < 1e-10, rel(ct.grad.cpu(), g_ref)


def test_block_without_sources_or_receivers(orc, ads, cuda):
    p, c, _, _ = _setup(ads, orc, 0, rank=1, NSTEP=20)          # block (1, 2)
    src = ads.MPIAcousticSource(p, [3], [3], np.ones((p.NSTEP, 1)))
    rcv = ads.MPIAcousticReceiver(p, [3], [3])
    assert src.srci is None and rcv.rcvi is None
    ap = ads.MPIAcousticPropagatorSolver(p, src, T(c))
    assert not torch.any(ap.u != 0)
    assert tuple(ads.MPISimulatedObservation_acoustic(ap, rcv).shape) == (p.NSTEP + 1, 0)


def test_two_by_two_decomposition_is_bit_identical_to_the_full_grid(orc, ads, cuda):
    n, NSTEP = 10, 40
    NX = NY = 2 * n
    dt, hx, hy = 1e-3, 10.0, 10.0
    Sx, Sy = orc.acoustic_pml(NX, NY, hx, hy, 5, 1e-3, 2000.0)
    rng = np.random.default_rng(21)
    Cg = (1800.0 + 400.0 * rng.random((NX + 2, NY + 2))) ** 2
    w = orc.ricker(NSTEP, 4.0, 8.0, 1e6)
    src = (7, 13)                                                 # 0-based padded cell
    sig, tau, C = T(Sx), T(Sy), T(Cg)
    z = lambda: torch.zeros((NX + 2, NY + 2), dtype=torch.float64, device="cuda")
    full = [z(), z(), z(), z()]                                   # w, wold, phi, psi
    tile = [z(), z(), z(), z()]
    for k in range(2, NSTEP + 1):
        w_, wo, ph, ps = full
        u, po, so = ads.acoustic_one_step(w_, wo, ph, ps, sig, tau, C, dt, hx, hy, NX, NY)
        u, po, so = u.clone(), po.clone(), so.clone()
        u[src] += dt * dt * float(w[k - 1])
        full = [u, w_, po, so]
        w_, wo, ph, ps = tile
        U, PO, SO = z(), z(), z()
        for a in range(2):
            for b in range(2):
                blk = (slice(a * n, a * n + n + 2), slice(b * n, b * n + n + 2))
                inner = (slice(1 + a * n, 1 + a * n + n), slice(1 + b * n, 1 + b * n + n))
                ub, pb, sb = ads.mpi_acoustic_one_step(w_[blk], wo[blk], ph[blk], ps[blk],
                                                       sig[blk], tau[blk], C[inner], dt, hx, hy, n)
                U[inner], PO[inner], SO[inner] = ub, pb, sb
        U[src] += dt * dt * float(w[k - 1])
        tile = [U, w_, PO, SO]
        for q in (0, 2, 3):
            assert torch.equal(full[q], tile[q]), (k, q)
    assert torch.abs(full[0]).max().item() > 0
