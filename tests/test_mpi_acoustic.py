"""CPU checks of the MPIAcoustic subdomain host logic (MPIAcoustic.jl:3-199): the block grid
and rank -> (II, JJ) map, the block's PML profiles, and the global -> block index mapping of
sources and receivers."""
import numpy as np
import pytest


def _params(ads, **kw):
    d = dict(NX=60, NY=40, n=20, NSTEP=10, DELTAX=10.0, DELTAY=7.5, NPOINTS_PML=6, vp_ref=2200.0)
    d.update(kw)
    return ads.MPIAcousticPropagatorParams(**d)


def test_rank_to_block_and_validation(ads):
    seen = set()
    for r in range(6):
        p = ads.compute_PML_Params_mpi_acoustic(_params(ads), rank=r, size=6)
        assert (p.M, p.N) == (3, 2)
        assert p.II == r // 2 + 1 and p.JJ == r % 2 + 1          # MPIAcoustic.jl:160
        seen.add((p.II, p.JJ))
    assert len(seen) == 6
    with pytest.raises(ValueError):
        ads.compute_PML_Params_mpi_acoustic(_params(ads), rank=0, size=4)
    with pytest.raises(ValueError):
        ads.compute_PML_Params_mpi_acoustic(_params(ads, NX=50), rank=0)
    with pytest.raises(ValueError):
        ads.compute_PML_Params_mpi_acoustic(_params(ads), rank=6)


def test_block_profiles_are_slices_of_the_global_profiles(ads, orc):
    """Block (II, JJ) covers padded global cells (II-1) n .. (II-1) n + n + 1 in x (and the same
    in y, MPIAcoustic.jl:177-178), so its Sigma_x / Sigma_y are those rows / columns of the
    full grid's profiles, which the oracle's pinned closed form gives."""
    g = _params(ads)
    sx = orc.acoustic_pml_axis(g.NX + 2, g.DELTAX, 0.0, g.NX, g.NPOINTS_PML, g.Rcoef, g.vp_ref)
    ty = orc.acoustic_pml_axis(g.NY + 2, g.DELTAY, 0.0, g.NY, g.NPOINTS_PML, g.Rcoef, g.vp_ref)
    n = g.n
    for r in range(6):
        p = ads.compute_PML_Params_mpi_acoustic(_params(ads), rank=r)
        i0, j0 = (p.II - 1) * n, (p.JJ - 1) * n
        assert p.Σx.shape == p.Σy.shape == (n + 2, n + 2)
        np.testing.assert_allclose(p.Σx[:, 0], sx[i0:i0 + n + 2], rtol=1e-13, atol=1e-12)
        np.testing.assert_allclose(p.Σy[0, :], ty[j0:j0 + n + 2], rtol=1e-13, atol=1e-12)
        assert np.all(p.Σx == p.Σx[:, :1]) and np.all(p.Σy == p.Σy[:1, :])
    # the interior block of a 3 x 3 split has no PML at all
    p = ads.compute_PML_Params_mpi_acoustic(_params(ads, NX=60, NY=60), rank=4)
    assert (p.II, p.JJ) == (2, 2) and not p.Σx.any() and not p.Σy.any()


def test_sources_and_receivers_map_to_their_block(ads):
    """Global 1-based (i, j) on the unpadded grid: each position lands in exactly one block,
    at block index (i - (II-1) n, j - (JJ-1) n), with its srcv column (MPIAcoustic.jl:56-104)."""
    rng = np.random.default_rng(0)
    g = _params(ads)
    si = rng.integers(1, g.NX + 1, 40)
    sj = rng.integers(1, g.NY + 1, 40)
    si[:4] = [1, g.NX, 20, 21]                    # block edges
    sj[:4] = [1, g.NY, 20, 21]
    sv = rng.standard_normal((g.NSTEP, 40))
    hits = np.zeros(40, int)
    for r in range(6):
        p = ads.compute_PML_Params_mpi_acoustic(_params(ads), rank=r)
        s = ads.MPIAcousticSource(p, si, sj, sv)
        rv = ads.MPIAcousticReceiver(p, si, sj)
        if s.srci is None:
            assert rv.rcvi is None
            continue
        gi = s.srci + (p.II - 1) * p.n
        gj = s.srcj + (p.JJ - 1) * p.n
        assert s.srci.min() >= 1 and s.srci.max() <= p.n and s.srcj.min() >= 1 and s.srcj.max() <= p.n
        k = [int(np.nonzero((si == a) & (sj == b))[0][0]) for a, b in zip(gi, gj)]
        np.testing.assert_array_equal(s.srcv, sv[:, np.nonzero(
            ((p.II - 1) * p.n < si) & (si <= p.II * p.n) & ((p.JJ - 1) * p.n < sj) & (sj <= p.JJ * p.n))[0]])
        np.testing.assert_array_equal(rv.rcvi, s.srci)
        np.testing.assert_array_equal(rv.rcvj, s.srcj)
        for q in np.nonzero(((p.II - 1) * p.n < si) & (si <= p.II * p.n) &
                            ((p.JJ - 1) * p.n < sj) & (sj <= p.JJ * p.n))[0]:
            hits[q] += 1
        assert len(k) == len(s.srci)
    assert np.all(hits == 1)
