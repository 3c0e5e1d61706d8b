"""CPU checks that pin the FP64 oracle (acoustic path) before it is trusted as the checker."""
import numpy as np
import pytest


def _golden_inputs(g, name):
    keys = ("w", "wold", "phi", "psi", "sigma", "tau", "c")
    NX, NY = (int(x) for x in g[f"{name}/dims"])
    dt, hx, hy = (float(x) for x in g[f"{name}/scalars"])
    return [g[f"{name}/{k}"] for k in keys], (dt, hx, hy, NX, NY)


@pytest.mark.parametrize("name", ["10x10", "7x12"])
def test_step_kernel1_equals_reference_header(orc, golden, name):
    """AcousticOneStepCpuForward compiled from the reference tree (oracle/_ref) wrote the
    golden outputs; the restatement must reproduce them bit for bit."""
    g = golden("acoustic_one_step.npz")
    ins, sc = _golden_inputs(g, name)
    u, po, so = orc.acoustic_step(1, *ins, *sc)
    np.testing.assert_array_equal(u, g[f"{name}/u"])
    np.testing.assert_array_equal(po, g[f"{name}/phiout"])
    np.testing.assert_array_equal(so, g[f"{name}/psiout"])


@pytest.mark.parametrize("name", ["10x10", "7x12"])
def test_step_vjp_kernel1_equals_reference_header(orc, golden, name):
    g = golden("acoustic_one_step.npz")
    ins, sc = _golden_inputs(g, name)
    outs = orc.acoustic_step_vjp(1, g[f"{name}/grad_u"], g[f"{name}/grad_phiout"],
                                 g[f"{name}/grad_psiout"], *ins, *sc)
    for o, k in zip(outs, ("grad_w", "grad_wold", "grad_phi", "grad_psi", "grad_c")):
        np.testing.assert_array_equal(o, g[f"{name}/{k}"], err_msg=k)


@pytest.mark.parametrize("name", ["10x10", "7x12"])
def test_step_vjp_gather_form_matches_reference_header(orc, golden, name):
    """The gather-form VJP (multi-threaded CPU baseline) gives the reference header's
    backward up to summation order."""
    g = golden("acoustic_one_step.npz")
    ins, sc = _golden_inputs(g, name)
    w, _, _, _, sigma, tau, c = ins
    prev = orc.num_threads()
    orc.set_num_threads(3)
    try:
        outs = orc.acoustic_step_vjp_gather(g[f"{name}/grad_u"], g[f"{name}/grad_phiout"],
                                            g[f"{name}/grad_psiout"], w, sigma, tau, c, *sc)
    finally:
        orc.set_num_threads(prev)
    for o, k in zip(outs, ("grad_w", "grad_wold", "grad_phi", "grad_psi", "grad_c")):
        ref = g[f"{name}/{k}"]
        np.testing.assert_allclose(o, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max(), err_msg=k)


def test_parallel_adjoint_sweep_equals_serial(orc):
    """The multi-threaded kernel-1 reverse sweep equals the serial one (scatter-form VJPs) on
    a PML grid with a source, up to summation order; kernel 0 is refused."""
    NX, NY, NSTEP, h, dt = 23, 19, 60, 10.0, 1e-3
    rng = np.random.default_rng(11)
    Sx, Sy = orc.acoustic_pml(NX, NY, h, h, 5, 1e-3, 2000.0)
    C2 = (1800.0 + 500.0 * rng.random((NX + 2, NY + 2))) ** 2
    srcv = rng.standard_normal((NSTEP, 1))
    u, ph, ps = orc.acoustic_forward(1, NX, NY, NSTEP, dt, h, h, Sx, Sy, C2, [7], [9], srcv)
    gu = rng.standard_normal((NSTEP + 1, (NX + 2) * (NY + 2)))
    ref = orc.acoustic_adjoint(1, NX, NY, NSTEP, dt, h, h, Sx, Sy, C2, u, ph, ps, gu)
    prev = orc.num_threads()
    orc.set_num_threads(4)
    try:
        par = orc.acoustic_adjoint_par(1, NX, NY, NSTEP, dt, h, h, Sx, Sy, C2, u, gu)
    finally:
        orc.set_num_threads(prev)
    assert np.abs(par - ref).max() <= 1e-12 * np.abs(ref).max()
    with pytest.raises(ValueError):
        orc.acoustic_adjoint_par(0, NX, NY, NSTEP, dt, h, h, Sx, Sy, C2, u, gu)


def test_kernel0_equals_kernel1_outside_pml(orc):
    """With sigma = tau = 0 and phi = psi = 0 both variants advance u identically and keep
    phi = psi = 0 (Core.jl:534-548 vs AcousticOneStepCpu.h:34-44)."""
    rng = np.random.default_rng(1)
    NX, NY = 9, 11
    N = (NX + 2) * (NY + 2)
    w, wold, c = rng.random(N), rng.random(N), rng.random(N)
    z = np.zeros(N)
    a = orc.acoustic_step(0, w, wold, z, z, z, z, c, 0.1, 0.3, 0.2, NX, NY)
    b = orc.acoustic_step(1, w, wold, z, z, z, z, c, 0.1, 0.3, 0.2, NX, NY)
    np.testing.assert_array_equal(a[0], b[0])
    assert not a[1].any() and not a[2].any() and not b[1].any()


def test_kernel0_and_kernel1_differ_in_pml(orc):
    """The variants must differ where sigma != tau (phi from u^n vs from w)."""
    rng = np.random.default_rng(2)
    NX, NY = 9, 11
    N = (NX + 2) * (NY + 2)
    args = [rng.random(N) for _ in range(7)]
    a = orc.acoustic_step(0, *args, 0.1, 0.3, 0.2, NX, NY)
    b = orc.acoustic_step(1, *args, 0.1, 0.3, 0.2, NX, NY)
    np.testing.assert_array_equal(a[0], b[0])
    assert np.abs(a[1] - b[1]).max() > 1e-3


@pytest.mark.parametrize("kernel", [0, 1])
def test_step_vjp_dot_product(orc, kernel):
    """<J dx, y> = <dx, J^T y> with J from central differences of the step."""
    rng = np.random.default_rng(3 + kernel)
    NX, NY = 8, 10
    N = (NX + 2) * (NY + 2)
    sig, tau = rng.random(N), rng.random(N)
    x0 = rng.random(5 * N)

    def F(x):
        w, wo, p, s, c = np.split(x, 5)
        return np.concatenate(orc.acoustic_step(kernel, w, wo, p, s, sig, tau, c, 0.1, 0.1, 0.1, NX, NY))

    dx = rng.standard_normal(5 * N)
    y = rng.standard_normal(3 * N)
    eps = 1e-6
    Jdx = (F(x0 + eps * dx) - F(x0 - eps * dx)) / (2 * eps)
    w, wo, p, s, c = np.split(x0, 5)
    gy = np.split(y, 3)
    o = orc.acoustic_step_vjp(kernel, *gy, w, wo, p, s, sig, tau, c, 0.1, 0.1, 0.1, NX, NY)
    lhs, rhs = Jdx @ y, dx @ np.concatenate(o)
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs)


def test_pml_profile_closed_form(orc):
    """pml_helper: at x = 0 d = L so sigma = xi (sin term vanishes); zero in the interior;
    mirror symmetric for a symmetric grid; xi = vp_ref/L ln(1/R)."""
    NX, h, npml, R, vp = 50, 10.0, 8, 1e-3, 2000.0
    s = orc.acoustic_pml_axis(NX + 2, h, 0.0, NX, npml, R, vp)
    xi = vp / (npml * h) * np.log(1 / R)
    assert s[0] == pytest.approx(xi, rel=1e-15)
    assert np.all(s[npml:NX + 2 - npml] == 0.0)
    assert np.all(s[:npml] > 0)
    np.testing.assert_allclose(s, s[::-1], rtol=1e-12, atol=1e-12)
    d = npml * h - 3 * h
    assert s[3] == pytest.approx(xi * (d / (npml * h) - np.sin(2 * np.pi * d / (npml * h)) / (2 * np.pi)))
    off = orc.acoustic_pml_axis(NX + 2, h, 0.0, NX, npml, R, vp, use_min=False, use_max=True)
    assert np.all(off[:npml] == 0) and np.all(off[-npml:] > 0)


def _setup(orc, NX=20, NY=24, NSTEP=60, seed=1):
    rng = np.random.default_rng(seed)
    Sx, Sy = orc.acoustic_pml(NX, NY, 10.0, 10.0, 5, 1e-3, 2000.0)
    c = 2000 + 300 * rng.random((NX + 2, NY + 2))
    srcv = orc.ricker(NSTEP, 8.0, 20.0, 1e6)[:, None]
    return rng, Sx, Sy, c, srcv


@pytest.mark.parametrize("kernel", [0, 1])
def test_misfit_gradient_finite_differences(orc, kernel):
    NX, NY, NSTEP = 20, 24, 60
    rng, Sx, Sy, c, srcv = _setup(orc, NX, NY, NSTEP)
    rcvi, rcvj = np.arange(1, NX + 1), np.full(NX, 2)
    u, _, _ = orc.acoustic_forward(kernel, NX, NY, NSTEP, 1e-3, 10.0, 10.0, Sx, Sy, (1.05 * c) ** 2,
                                   [NX // 2], [3], srcv)
    obs = orc.acoustic_receive(u, rcvi, rcvj, NY)
    L, g, _ = orc.acoustic_misfit_gradient(kernel, NX, NY, NSTEP, 1e-3, 10.0, 10.0, Sx, Sy, c,
                                           [NX // 2], [3], srcv, rcvi, rcvj, obs)
    v = rng.standard_normal(c.shape)
    eps = 1e-3
    f = lambda cc: orc.acoustic_misfit_gradient(kernel, NX, NY, NSTEP, 1e-3, 10.0, 10.0, Sx, Sy, cc,
                                                [NX // 2], [3], srcv, rcvi, rcvj, obs)[0]
    fd = (f(c + eps * v) - f(c - eps * v)) / (2 * eps)
    assert np.sum(g * v) == pytest.approx(fd, rel=1e-7)


def test_forward_linear_in_source(orc):
    NX, NY, NSTEP = 16, 18, 40
    rng, Sx, Sy, c, srcv = _setup(orc, NX, NY, NSTEP)
    f = lambda sv: orc.acoustic_forward(0, NX, NY, NSTEP, 1e-3, 10.0, 10.0, Sx, Sy, c ** 2,
                                        [5, 9], [4, 7], sv)[0]
    a = rng.standard_normal((NSTEP, 2))
    b = rng.standard_normal((NSTEP, 2))
    np.testing.assert_allclose(f(2 * a - 3 * b), 2 * f(a) - 3 * f(b), rtol=1e-9, atol=1e-12)
    assert not f(np.zeros((NSTEP, 2))).any()


def test_forward_causality_and_source_timing(orc):
    """A 5-point stencil moves information one cell per step: with the source injected at level
    2 (row 1 of srcv), cell (i, j) is exactly 0 until level 2 + |i-i0| + |j-j0|.  This pins the
    level/row alignment (level n takes row n-1) and the (srci-1)(NY+2)+(srcj-1) mapping."""
    NX, NY, NSTEP = 20, 20, 14
    Sx, Sy = np.zeros((NX + 2, NY + 2)), np.zeros((NX + 2, NY + 2))
    c2 = np.full((NX + 2, NY + 2), 2000.0 ** 2)
    srcv = np.zeros((NSTEP, 1))
    srcv[1, 0] = 1.0  # only row 1 -> level 2
    u, _, _ = orc.acoustic_forward(1, NX, NY, NSTEP, 1e-3, 10.0, 10.0, Sx, Sy, c2, [11], [9], srcv)
    i0, j0 = 10, 8  # 0-based position of 1-based (11, 9)
    assert u[2, i0, j0] == pytest.approx(1e-6)  # dt^2 * 1
    assert np.count_nonzero(u[2]) == 1
    ii, jj = np.meshgrid(np.arange(NX + 2), np.arange(NY + 2), indexing="ij")
    dist = np.abs(ii - i0) + np.abs(jj - j0)
    for n in range(2, NSTEP + 1):
        inside = dist <= n - 2
        assert not u[n][~inside].any(), n
        assert u[n][inside & (dist == n - 2) & (ii > 0) & (ii < NX + 1) & (jj > 0) & (jj < NY + 1)].all(), n


def test_source_time_alignment_rows(orc):
    """Row 0 of srcv is never used (level n >= 2 takes row n-1, Core.jl:593, 607-613)."""
    NX, NY, NSTEP = 10, 10, 6
    z = np.zeros((NX + 2, NY + 2))
    c2 = np.full((NX + 2, NY + 2), 1.0)
    sv = np.zeros((NSTEP, 1))
    sv[0, 0] = 123.0
    u, _, _ = orc.acoustic_forward(0, NX, NY, NSTEP, 1e-3, 10.0, 10.0, z, z, c2, [5], [5], sv)
    assert not u.any()


def test_ricker_fixture_matches_reference_data(orc, golden):
    """The wavelets in examples/nn_fwi/models/*.mat equal the frequency-domain Ricker
    restatement (examples/nn_fwi/Marmousi_model.ipynb cells 4 and 11)."""
    g = golden("wavelet_fixtures.npz")
    for tag in ("marmousi2", "BP"):
        nt, dt, f0, nx, ny = g[f"{tag}/nt_dt_f0_nx_ny"]
        vec = g[f"{tag}/vec"]
        mine = orc.ricker_freq(int(nt), f0, dt, 1e4)
        np.testing.assert_allclose(mine, vec, rtol=1e-12, atol=1e-9)
        assert int(np.argmax(vec)) == 128
        # geometry is 1-based on the padded grid: every position is inside
        for k in ("src_ix", "rcv_ix"):
            assert g[f"{tag}/{k}"].min() >= 1 and g[f"{tag}/{k}"].max() <= nx
        assert g[f"{tag}/rcv_iy"].min() >= 1 and g[f"{tag}/rcv_iy"].max() <= ny


def test_sample_ricker_closed_form(orc):
    """Ricker(param, a, shift, amp): peak amp*2/(sqrt(3a) pi^(1/4)) at k = shift, zeros at
    k = shift +- a (Utils.jl:188-204)."""
    a, shift, amp = 10.0, 50.0, 3.0
    r = orc.ricker(200, a, shift, amp)
    assert r[49] == pytest.approx(amp * 2 / (np.sqrt(3 * a) * np.pi ** 0.25))
    assert abs(r[59]) < 1e-15 and abs(r[39]) < 1e-15
    assert int(np.argmax(r)) == 49
