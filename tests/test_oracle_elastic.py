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
