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


# ------------------------------------------------------------------------------------------
# one-step probes of the full form (Core.jl:96-213, 231-407).  Each probe is built so that a
# plausible slip in the oracle -- one CPML profile row or index swapped, one averaging
# neighbour moved, one update range off by a cell, the right-edge CPML origin misplaced, the
# sign of the memory recursion -- changes which cells move, not just a value.
# Field order f[0..4] = vx, vy, sxx, syy, sxy; memory m[0..7] = mem[1..8] of Core.jl.
# ------------------------------------------------------------------------------------------

# 0-based update ranges (i_lo, i_hi, j_lo, j_hi) read off the getid() index sets of
# fw1..fw4: fw1 ij = getid(2:NX, 3:NY+1) (Core.jl:100), fw2 getid(3:NX+1, 2:NY) (:134),
# fw3 getid(3:NX+1, 3:NY+1) (:164), fw4 getid(2:NX, 2:NY) (:191); getid is 1-based.
def _ranges(NX, NY):
    fw1 = (1, NX - 1, 2, NY)
    fw2 = (2, NX, 1, NY - 1)
    fw3 = (2, NX, 2, NY)
    fw4 = (1, NX - 1, 1, NY - 1)
    field = {0: fw3, 1: fw4, 2: fw1, 3: fw1, 4: fw2}            # vx, vy, sxx, syy, sxy
    mem = {0: fw1, 1: fw1, 2: fw2, 3: fw2, 4: fw3, 5: fw3, 6: fw4, 7: fw4}
    return field, mem


def _mask(NX, NY, r):
    m = np.zeros((NX + 2, NY + 2), bool)
    m[r[0]:r[1] + 1, r[2]:r[3] + 1] = True
    return m


def _probe_setup(NX, NY, seed):
    rng = np.random.default_rng(seed)
    rho, lam, mu = _model(NX, NY, seed=seed)
    N = (NX + 2) * (NY + 2)
    f = rng.standard_normal(5 * N) * np.repeat([1e-3, 1e-3, 1e5, 1e5, 1e5], N)
    m = rng.standard_normal(8 * N) * 1e-4
    return rho, lam, mu, f, m


def _ident(n):
    return np.zeros(2 * n), np.ones(2 * n)         # a = 0, b = 1: the CPML is the identity


def _step(orc, NX, NY, rho, lam, mu, ax, bx, ay, by, f, m):
    f1, m1 = orc.elastic_step(NX, NY, 10.0, 12.0, 1e-3, rho, lam, mu, ax, bx, ay, by, f.copy(),
                              m.copy(), [2], [2], [0], [0.0])
    sh = (NX + 2, NY + 2)
    return f1.reshape((5,) + sh), m1.reshape((8,) + sh)


def test_update_ranges_and_memory_ranges_are_the_reference_index_sets(orc):
    """Identity CPML, random fields and memory variables: after one step every field changes
    on every cell of its fw range and is bit-unchanged outside it; every memory variable equals
    its old value on its range (b = 1, a = 0) and is 0 outside (makevector scatters into
    zeros, Core.jl:697-699)."""
    NX, NY = 13, 11
    rho, lam, mu, f, m = _probe_setup(NX, NY, 11)
    ax, bx = _ident(NX)
    ay, by = _ident(NY)
    F, Mv = _step(orc, NX, NY, rho, lam, mu, ax, bx, ay, by, f, m)
    sh = (NX + 2, NY + 2)
    f0, m0 = f.reshape((5,) + sh), m.reshape((8,) + sh)
    fr, mr = _ranges(NX, NY)
    for k, r in fr.items():
        ins = _mask(NX, NY, r)
        assert np.all(F[k][~ins] == f0[k][~ins]), f"field {k} moved outside its range"
        assert np.all(F[k][ins] != f0[k][ins]), f"field {k} did not move inside its range"
    for k, r in mr.items():
        ins = _mask(NX, NY, r)
        assert np.all(Mv[k][ins] == m0[k][ins]), f"mem[{k + 1}] changed under b = 1, a = 0"
        assert np.all(Mv[k][~ins] == 0.0), f"mem[{k + 1}] not zero outside its range"


# which memory variables each CPML profile row drives (Core.jl:116-117, 146-147, 174-175,
# 205-206): bx/ax row 1 (full grid) -> mem[3], mem[5]; row 2 (half grid) -> mem[1], mem[7];
# by/ay row 1 -> mem[2], mem[6]; row 2 -> mem[4], mem[8]  (0-based memory indices below)
_PROFILE_MEMS = {("x", 0): (2, 4), ("x", 1): (0, 6), ("y", 0): (1, 5), ("y", 1): (3, 7)}


@pytest.mark.parametrize("axis,row", sorted(_PROFILE_MEMS))
@pytest.mark.parametrize("coef", ["a", "b"])
def test_each_cpml_profile_drives_its_own_memory_variables(orc, axis, row, coef):
    """One profile entry (row `row`, index k) is made non-identity, all others stay a = 0,
    b = 1.  Exactly the two memory variables the reference assigns to that row change, and only
    on the cells whose profile index i-1 (x) or j-1 (y) equals k (adbroadcast, Core.jl:675-682):
    b = 0.5 halves the old value there, a = 0.25 adds a quarter of the derivative there."""
    NX, NY = 14, 12
    k = 5
    rho, lam, mu, f, m = _probe_setup(NX, NY, 12)
    if coef == "a":
        m = np.zeros_like(m)
    ax, bx = _ident(NX)
    ay, by = _ident(NY)
    prof = {("x", "a"): ax, ("x", "b"): bx, ("y", "a"): ay, ("y", "b"): by}[(axis, coef)]
    n = NX if axis == "x" else NY
    prof[row * n + k] = 0.5 if coef == "b" else 0.25
    F, Mv = _step(orc, NX, NY, rho, lam, mu, ax, bx, ay, by, f, m)
    sh = (NX + 2, NY + 2)
    m0 = m.reshape((8,) + sh)
    _, mr = _ranges(NX, NY)
    line = np.zeros(sh, bool)
    if axis == "x":
        line[k + 1, :] = True
    else:
        line[:, k + 1] = True
    for q in range(8):
        ins = _mask(NX, NY, mr[q])
        moved = ins & (Mv[q] != m0[q])
        if q in _PROFILE_MEMS[(axis, row)]:
            np.testing.assert_array_equal(moved, ins & line, err_msg=f"mem[{q + 1}]")
            if coef == "b":
                np.testing.assert_array_equal(Mv[q][ins & line], 0.5 * m0[q][ins & line])
        else:
            assert not moved.any(), f"mem[{q + 1}] moved under a profile it does not use"


def _moved(orc, NX, NY, rho, lam, mu, f, m, pml):
    ax, bx, _, ay, by, _ = pml
    return _step(orc, NX, NY, rho, lam, mu, ax, bx, ay, by, f, m)


@pytest.mark.parametrize("param", ["lam", "mu", "rho"])
def test_parameter_averaging_neighbours(orc, param):
    """One cell (i0, j0) of lambda, mu or rho is perturbed; after one step the stresses and
    velocities differ exactly on the cells whose update reads that parameter value:
    lambda-bar, mu-bar of fw1 average (i, j) and (i+1, j) (Core.jl:109-110); mu-bar of fw2
    averages (i, j) and (i, j+1) (:141); fw3 divides by rho(i, j) (:181); rho-bar of fw4 is the
    mean of (i, j), (i+1, j), (i+1, j+1), (i, j+1) (:200).  fw1/fw2 do not read rho; fw3/fw4
    read lambda and mu only through the new stresses, so v is checked for rho only."""
    NX, NY = 15, 13
    i0, j0 = 7, 6
    rho, lam, mu, f, m = _probe_setup(NX, NY, 13)
    pml = orc.elastic_pml(NX, NY, 10.0, 12.0, 1e-3, 4, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2500.0)
    base = {"rho": rho, "lam": lam, "mu": mu}
    pert = {k: v.copy() for k, v in base.items()}
    pert[param][i0, j0] *= 1.001
    F0, _ = _moved(orc, NX, NY, base["rho"], base["lam"], base["mu"], f, m, pml)
    F1, _ = _moved(orc, NX, NY, pert["rho"], pert["lam"], pert["mu"], f, m, pml)

    def cells(*ij):
        s = np.zeros((NX + 2, NY + 2), bool)
        for a, b in ij:
            s[a, b] = True
        return s

    none = cells()
    ij_x = cells((i0, j0), (i0 - 1, j0))                       # (i, j) and (i+1, j) averages
    ij_y = cells((i0, j0), (i0, j0 - 1))                       # (i, j) and (i, j+1) averages
    ij_4 = cells((i0, j0), (i0 - 1, j0), (i0 - 1, j0 - 1), (i0, j0 - 1))
    expect = {"lam": {2: ij_x, 3: ij_x, 4: none},
              "mu": {2: ij_x, 3: ij_x, 4: ij_y},
              "rho": {2: none, 3: none, 4: none, 0: cells((i0, j0)), 1: ij_4}}[param]
    for q, want in expect.items():
        np.testing.assert_array_equal(F1[q] != F0[q], want, err_msg=f"{param}: field {q}")


def test_right_edge_cpml_origin_mirrors_the_left_edge(orc):
    """The right/top CPML origin is (NX - 0.5) h - thickness (Core.jl:256, 342) and profile
    entry k sits at x = k h (full) or k h + h/2 (half).  The staggered grid then mirrors the
    left edge onto the right with full and half grids exchanged: full[k] = half[NX-1-k] and
    half[k] = full[NX-1-k] for a, b.  Spot values: the last half-grid entry is at the outer
    edge (normalised abscissa 1, b = exp(-(d0 + alpha_max) dt)); the right strip's full grid
    starts at NX - npml, its half grid has abscissa 0 (a = 0, b = 1) at NX - 1 - npml."""
    for n, h in ((37, 10.0), (50, 7.5)):
        npml, npower, amax, R, vref, dt = 6, 2.0, 2 * np.pi * 6, 1e-3, 2400.0, 8e-4
        a, b, _ = orc.elastic_pml_axis(n, h, 0.0, n, npml, npower, 1.0, amax, R, vref, dt)
        for q in (a, b):
            np.testing.assert_allclose(q[0], q[1][::-1], rtol=1e-12, atol=1e-300)
            np.testing.assert_allclose(q[1], q[0][::-1], rtol=1e-12, atol=1e-300)
        d0 = -(npower + 1) * vref * np.log(R) / (2.0 * npml * h)
        assert b[1, n - 1] == pytest.approx(np.exp(-(d0 + amax) * dt), rel=1e-13)
        assert b[0, n - npml] < 1.0 and b[0, n - npml - 1] == 1.0 and a[0, n - npml - 1] == 0.0
        assert b[1, n - 1 - npml] == 1.0 and a[1, n - 1 - npml] == 0.0
        assert b[1, n - npml] < 1.0
        # the right strip alone (no left PML) is the mirror of the left strip alone
        al, bl, _ = orc.elastic_pml_axis(n, h, 0.0, n, npml, npower, 1.0, amax, R, vref, dt,
                                         use_max=False)
        ar, br, _ = orc.elastic_pml_axis(n, h, 0.0, n, npml, npower, 1.0, amax, R, vref, dt,
                                         use_min=False)
        np.testing.assert_allclose(bl[0], br[1][::-1], rtol=1e-12)
        assert np.all(bl[:, n // 2:] == 1.0) and np.all(br[:, :n // 2] == 1.0)


def test_energy_decays_in_the_cpml(orc):
    """Homogeneous medium, explosive source in the middle: with the CPML on all four sides the
    field energy falls to well below 1% of its peak once the waves have left the box; with the
    CPML off (zero-field outer ring, a reflecting box) it stays of the order of its peak.  A
    wrong sign or a swapped a/b in the memory recursion would not absorb (or would blow up)."""
    NX = NY = 60
    NSTEP = 700
    hx = hy = 10.0
    dt = 1e-3
    vp, vs = 3000.0, 1700.0
    rho = np.full((NX + 2, NY + 2), 2000.0)
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    w = orc.ricker(NSTEP, 6.0, 25.0, 1e6)
    src = ([31, 31], [31, 31], [2, 3], np.stack([w, w], 1))

    def energy(use):
        pml = orc.elastic_pml(NX, NY, hx, hy, dt, 12, 2.0, 1.0, 2 * np.pi * 10, 1e-3, vp, use=use)
        hist, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, *src,
                                      keep_mem=False)
        kin = 0.5 * rho[None] * (hist[0] ** 2 + hist[1] ** 2)
        strain = (hist[2] ** 2 + hist[3] ** 2 + 2 * hist[4] ** 2) / (4 * mu[None])
        return (kin + strain).sum(axis=(1, 2))

    e_pml = energy((True, True, True, True))
    e_box = energy((False, False, False, False))
    assert np.all(np.isfinite(e_pml))
    assert e_pml[-1] < 1e-2 * e_pml.max(), e_pml[-1] / e_pml.max()
    assert e_box[-1] > 0.2 * e_box.max(), e_box[-1] / e_box.max()


def test_source_gradient_finite_differences(orc):
    """The oracle's d misfit / d srcv (the AddSource backward of AddSource.cpp:91-159, composed
    through the time loop) against a central finite difference along a random srcv
    direction, with duplicated (cell, type) sources and every source type."""
    NX, NY, NSTEP = 22, 18, 40
    hx, hy, dt = 10.0, 10.0, 1e-3
    rho, lam, mu = _model(NX, NY, seed=4)
    pml = orc.elastic_pml(NX, NY, hx, hy, dt, 4, 2.0, 1.0, 2 * np.pi * 5, 1e-3, 2500.0)
    w = orc.ricker(NSTEP, 5.0, 12.0, 1e8)
    srci = [11, 11, 8, 14, 6, 11]
    srcj = [6, 6, 9, 5, 12, 6]
    st = [2, 2, 0, 1, 4, 3]
    srcv = np.stack([w, 0.5 * w, 1e-7 * w, -1e-7 * w, 0.3 * w, w], 1)
    rcvi = np.concatenate([np.arange(2, NX), np.arange(2, NX)])
    rcvj = np.full(len(rcvi), 3)
    rt = np.concatenate([np.zeros(NX - 2), np.ones(NX - 2)]).astype(np.int64)
    hist, mem = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci, srcj, st,
                                    srcv)
    rc = orc.elastic_receive(hist, rcvi, rcvj, rt, NX, NY)
    obs = 0.9 * rc
    res = rc - obs
    g = np.zeros_like(hist).reshape(5, NSTEP + 1, -1)
    idx = orc.flat_index(rcvi, rcvj, NY)
    for r in range(len(idx)):
        g[int(rt[r]), :, idx[r]] += 2.0 * res[r]
    *_, gsrc = orc.elastic_adjoint(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, hist, mem, g,
                                   srci, srcj, st)
    assert gsrc.shape == (NSTEP, len(srci))
    v = np.random.default_rng(6).standard_normal(srcv.shape) * np.abs(srcv).max(0) * 1e-4

    def L(sv):
        h2, _ = orc.elastic_forward(NX, NY, NSTEP, hx, hy, dt, rho, lam, mu, pml, srci, srcj, st,
                                    sv, keep_mem=False)
        return float(np.sum((orc.elastic_receive(h2, rcvi, rcvj, rt, NX, NY) - obs) ** 2))

    fd = (L(srcv + v) - L(srcv - v)) / 2
    assert np.sum(gsrc * v) == pytest.approx(fd, rel=1e-6)
    # the duplicated (cell, type) pair sees the same cotangent
    np.testing.assert_array_equal(gsrc[:, 0], gsrc[:, 1])
