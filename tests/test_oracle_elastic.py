"""CPU checks that pin the FP64 oracle's elastic path (Core.jl:31-407 restated)."""
import numpy as np
import pytest


def _model(NX, NY, seed=0, vp0=2500.0, jitter=200.0):
    rng = np.random.default_rng(seed)
    vp = vp0 + jitter * rng.random((NX + 2, NY + 2))
    vs = vp / 1.73
    rho = 310 * vp ** 0.25
    return rho, rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2


def test_cpml_profile_spot_values(orc):
    """Outside the strips d = 0, alpha = 0 -> b = 1, a = 0; inside, 0 < b < 1 and a < 0;
    alpha grows into the PML (Core.jl:289) so b is smallest at the edge."""
    NX, h, dt, npml = 60, 10.0, 1e-3, 8
    a, b, K = orc.elastic_pml_axis(NX, h, 0.0, NX, npml, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2000.0, dt)
    inner = slice(npml + 1, NX - npml - 1)
    assert np.all(b[:, inner] == 1.0) and np.all(a[:, inner] == 0.0)
    assert np.all(K == 1.0)
    assert np.all(b[0, :npml] < 1.0) and np.all(a[0, :npml] < 0.0)
    assert b[0, 0] < b[0, npml // 2] < 1.0
    d0 = -(2.0 + 1) * 2000.0 * np.log(1e-3) / (2.0 * npml * h)
    al = 2 * np.pi * 5 * 1.0
    assert b[0, 0] == pytest.approx(np.exp(-(d0 + al) * dt), rel=1e-14)


@pytest.mark.parametrize("param", ["rho", "lam", "mu"])
def test_elastic_gradient_finite_differences(orc, param):
    NX, NY, NSTEP = 24, 20, 50
    hx = hy = 10.0
    dt = 1e-3
    rho, lam, mu = _model(NX, NY)
    pml = orc.elastic_pml(NX, NY, hx, hy, dt, 5, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2500.0)
    w = orc.ricker(NSTEP, 6.0, 15.0, 1e9)
    srci, srcj, st = [NX // 2, NX // 2, 7], [4, 4, 9], [2, 3, 1]
    srcv = np.stack([w, w, 0.3 * w], 1)
    rcvi = np.concatenate([np.arange(2, NX), np.arange(2, NX), [5]])
    rcvj = np.full(len(rcvi), 3)
    rt = np.concatenate([np.zeros(NX - 2), np.ones(NX - 2), [4]]).astype(np.int64)
    hist, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho * 1.02, lam * 0.98, mu * 1.03, pml,
                                  srci, srcj, st, srcv)
    obs = orc.elastic_receive(hist, rcvi, rcvj, rt, NX, NY)
    _, grads, _ = orc.elastic_misfit_gradient(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci,
                                              srcj, st, srcv, rcvi, rcvj, rt, obs)
    base = {"rho": rho, "lam": lam, "mu": mu}
    g = dict(zip(("rho", "lam", "mu"), grads))[param]
    v = np.random.default_rng(5).standard_normal(rho.shape) * base[param] * 1e-6

    def L(x):
        a = dict(base)
        a[param] = x
        return orc.elastic_misfit_gradient(NX, NY, NSTEP, hx, hy, dt, a["rho"], a["lam"], a["mu"],
                                           pml, srci, srcj, st, srcv, rcvi, rcvj, rt, obs)[0]

    fd = (L(base[param] + v) - L(base[param] - v)) / 2
    assert np.sum(g * v) == pytest.approx(fd, rel=1e-6)


def test_elastic_explosive_source_mirror_symmetry(orc):
    """Homogeneous medium, no PML, explosive source (sxx = syy) at 0-based cell (i0, j0): the
    staggered grid puts sxx at (i + 1/2, j), vx at (i, j), vy at (i + 1/2, j + 1/2), sxy at
    (i, j + 1/2).  Before the wave reaches the edges the fields obey the mirror symmetries
    about x = i0 + 1/2 and y = j0; this pins every index offset of fw1..fw4."""
    NX = NY = 40
    NSTEP = 14
    hx = hy = 10.0
    dt = 1e-3
    rho = np.full((NX + 2, NY + 2), 2000.0)
    lam = rho * (3000.0 ** 2 - 2 * 1700.0 ** 2)
    mu = rho * 1700.0 ** 2
    pml = orc.elastic_pml(NX, NY, hx, hy, dt, 5, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 3000.0,
                          use=(False, False, False, False))
    i0, j0 = 20, 20
    w = orc.ricker(NSTEP, 3.0, 5.0, 1e6)
    hist, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, [i0 + 1, i0 + 1],
                                  [j0 + 1, j0 + 1], [2, 3], np.stack([w, w], 1))
    vx, vy, sxx, syy, sxy = (hist[k, NSTEP] for k in range(5))
    assert np.abs(sxx).max() > 0
    k = np.arange(-8, 9)
    r = lambda a: np.abs(a).max()
    # x-mirror about i0 + 1/2 (sxx sits at i + 1/2): sxx(i0 + k) = sxx(i0 - k)
    Sx = sxx[i0 + k, j0 - 6:j0 + 7]
    assert r(Sx - sxx[i0 - k, j0 - 6:j0 + 7]) <= 1e-10 * r(sxx)
    # vx at integer i: vx(i) = -vx(2 i0 + 1 - i)
    Vx = vx[i0 + 1 + k, j0 - 6:j0 + 7]
    assert r(Vx + vx[i0 - k, j0 - 6:j0 + 7]) <= 1e-10 * r(vx)
    # y-mirror about j0: syy(j0 + k) = syy(j0 - k); vy at j + 1/2: vy(j) = -vy(2 j0 - 1 - j)
    assert r(syy[i0 - 6:i0 + 7, j0 + k] - syy[i0 - 6:i0 + 7, j0 - k]) <= 1e-10 * r(syy)
    assert r(vy[i0 - 6:i0 + 7, j0 + k] + vy[i0 - 6:i0 + 7, j0 - 1 - k]) <= 1e-10 * r(vy)
    # sxy at (i, j + 1/2): odd in both mirrors
    assert r(sxy[i0 + 1 + k, j0 - 6:j0 + 7] + sxy[i0 - k, j0 - 6:j0 + 7]) <= 1e-10 * r(sxy)


def test_elastic_p_wave_arrival(orc):
    """Homogeneous medium: the P wave from an explosive source reaches a receiver at distance r
    with the peak delay r / vp after the source peak (within two samples of the wavelet)."""
    NX, NY = 160, 60
    hx = hy = 10.0
    dt = 1e-3
    NSTEP = 420
    vp, vs = 3000.0, 1700.0
    rho = np.full((NX + 2, NY + 2), 2000.0)
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    pml = orc.elastic_pml(NX, NY, hx, hy, dt, 10, 2.0, 1.0, 2 * np.pi * 10, 1e-3, vp)
    shift = 60.0
    w = orc.ricker(NSTEP, 12.0, shift, 1e6)
    i0, j0 = 40, 30
    hist, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, [i0 + 1, i0 + 1],
                                  [j0 + 1, j0 + 1], [2, 3], np.stack([w, w], 1), keep_mem=False)
    dist_cells = 80
    tr = hist[3, :, i0 + dist_cells, j0]  # syy on the same row, 800 m away
    lag = np.argmax(np.abs(tr)) - (shift - 1)
    expected = dist_cells * hx / vp / dt
    assert abs(lag - expected) <= 12, (lag, expected)


# ------------------------------------------------------------------------------------------
# one-subdomain form of MPIElastic.jl (a block of n x n cells inside a 2-cell zero halo)
# ------------------------------------------------------------------------------------------

def _identity_profiles(n):
    z, o = np.zeros(2 * n), np.ones(2 * n)
    return z, o


def test_subdomain_form_equals_full_form_away_from_the_edges(orc):
    """With homogeneous parameters the averages of the full form are exact (lambda-bar = lambda,
    rho-bar = rho) and without CPML the two forms differ only in their update ranges, so a
    wavefield that has not reached the edges is bitwise the same (pins the subdomain update
    formulas and their index offsets against the already pinned full form)."""
    n = 30
    NX = NY = n + 2                     # padded (n+4)^2 either way
    shape = (NX + 2, NY + 2)
    rho = np.full(shape, 2048.0)
    lam = np.full(shape, 4.0e9)
    mu = np.full(shape, 2.0e9)
    a0, b0 = _identity_profiles(n)
    A0, B0 = _identity_profiles(NX)
    w = orc.ricker(3, 2.0, 2.0, 1e6)
    src = dict(srci=[17, 17, 16], srcj=[17, 17, 18], srctype=[2, 3, 0],
               srcv=np.stack([w, w, 0.5 * w], 1))
    h_sub, _ = orc.elastic_forward(NX, NY, 3, 10.0, 10.0, 1e-3, rho, lam, mu,
                                   (a0, b0, None, a0, b0, None), **src, sub=True)
    h_full, _ = orc.elastic_forward(NX, NY, 3, 10.0, 10.0, 1e-3, rho, lam, mu,
                                    (A0, B0, None, A0, B0, None), **src)
    assert np.abs(h_full[:, 3]).max() > 0
    np.testing.assert_array_equal(h_sub, h_full)


def test_subdomain_keeps_the_halo_zero_and_uses_block_profile_index(orc):
    """One step with a source in the block: the 2-cell halo stays 0, and a memory variable of
    a block cell is a[i-2] * D (profile index i-2, MPIElastic.jl:483-620) -- probed with a
    profile that is non-zero at one index only."""
    n = 12
    NX = NY = n + 2
    shape = (NX + 2, NY + 2)
    rng = np.random.default_rng(3)
    rho = 2000.0 + rng.random(shape)
    lam = 4e9 * (1 + 0.1 * rng.random(shape))
    mu = 2e9 * (1 + 0.1 * rng.random(shape))
    N = (NX + 2) * (NY + 2)
    ax = np.zeros(2 * n)
    bx = np.ones(2 * n)
    k = 5
    ax[k] = 0.25          # a_x (full grid) at block index k only
    a0, b0 = _identity_profiles(n)
    f = np.zeros(5 * N)
    f.reshape(5, NX + 2, NY + 2)[2, 2:NX, 2:NY] = rng.standard_normal((n, n))  # sxx in the block
    m = np.zeros(8 * N)
    f1, m1 = orc.elastic_step(NX, NY, 10.0, 10.0, 1e-3, rho, lam, mu, ax, bx, a0, b0, f.copy(), m,
                              [7], [7], [0], [0.0], sub=True)
    F = f1.reshape(5, NX + 2, NY + 2)
    Mv = m1.reshape(8, NX + 2, NY + 2)
    halo = np.ones((NX + 2, NY + 2), bool)
    halo[2:NX, 2:NY] = False
    assert np.all(F[:, halo] == 0) and np.all(Mv[:, halo] == 0)
    sxx = f.reshape(5, NX + 2, NY + 2)[2]
    i = k + 2
    for j in range(2, NY):
        dx = (27 * sxx[i, j] - 27 * sxx[i - 1, j] - sxx[i + 1, j] + sxx[i - 2, j]) / 240.0
        assert Mv[4, i, j] == pytest.approx(0.25 * dx, rel=1e-12)      # m5 of fw3
    assert np.all(Mv[4, 2:NX, 2:NY][np.arange(n) != k] == 0)


@pytest.mark.parametrize("param", ["rho", "lam", "mu"])
def test_subdomain_gradient_finite_differences(orc, param):
    n, NSTEP = 18, 40
    NX = NY = n + 2
    hx = hy = 10.0
    dt = 1e-3
    rho, lam, mu = _model(NX, NY, seed=7)
    pml = orc.elastic_pml(n, n, hx, hy, dt, 4, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2500.0)
    w = orc.ricker(NSTEP, 6.0, 15.0, 1e9)
    srci, srcj, st = [11, 11, 6], [5, 5, 12], [2, 3, 1]
    srcv = np.stack([w, w, 0.3 * w], 1)
    rcvi = np.concatenate([np.arange(3, n + 3), np.arange(3, n + 3)])   # block row 2, padded
    rcvj = np.full(len(rcvi), 4)
    rt = np.concatenate([np.zeros(n), np.ones(n)]).astype(np.int64)
    hist, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho * 1.02, lam * 0.98, mu * 1.03,
                                  pml, srci, srcj, st, srcv, sub=True)
    obs = orc.elastic_receive(hist, rcvi, rcvj, rt, NX, NY)
    _, grads, _ = orc.elastic_misfit_gradient(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci,
                                              srcj, st, srcv, rcvi, rcvj, rt, obs, sub=True)
    base = {"rho": rho, "lam": lam, "mu": mu}
    g = dict(zip(("rho", "lam", "mu"), grads))[param]
    halo = np.ones(rho.shape, bool)
    halo[2:NX, 2:NY] = False
    assert np.all(g[halo] == 0)      # no averaging: only block parameters are read
    v = np.random.default_rng(9).standard_normal(rho.shape) * base[param] * 1e-6

    def L(x):
        a = dict(base)
        a[param] = x
        return orc.elastic_misfit_gradient(NX, NY, NSTEP, hx, hy, dt, a["rho"], a["lam"], a["mu"],
                                           pml, srci, srcj, st, srcv, rcvi, rcvj, rt, obs,
                                           sub=True)[0]

    fd = (L(base[param] + v) - L(base[param] - v)) / 2
    assert np.sum(g * v) == pytest.approx(fd, rel=1e-6)
