"""MAT loaders (src/IO.jl:4-161) on files with the layout of the reference's example models.

The files are rebuilt from committed fixtures (tests/matfixture.py), so what is pinned is the
loaders' reading of that layout: 1-based geometry kept as stored, padded-grid sizes turned
into NX = nx - 2, wavelets as (nt, nsrc) columns (transposed when stored as rows for the
elastic loader), and scalar struct fields turned into one-element vectors (safe_vec)."""
import numpy as np
import pytest

from matfixture import marmousi_fixture, write_mat


@pytest.fixture(scope="module")
def fx():
    return marmousi_fixture()


def test_load_params_acoustic_and_elastic(ads, fx, tmp_path):
    f = write_mat(str(tmp_path / "m.mat"), fx, fx["vp_true"])
    p = ads.load_params(f, vp_ref=1e3)
    assert (p.NX, p.NY, p.NSTEP) == (fx["nx"] - 2, fx["ny"] - 2, fx["nt"]) == (200, 66, 1678)
    assert p.DELTAX == fx["dx"] and p.DELTAY == fx["dy"] and p.DELTAT == fx["dt"]
    assert p.vp_ref == 1e3 and p.PropagatorKernel == 0
    pe = ads.load_params(f, "Elastic", NPOINTS_PML=20)
    assert isinstance(pe, ads.ElasticPropagatorParams) and pe.NPOINTS_PML == 20 and pe.NX == 200
    with pytest.raises(ValueError):
        ads.load_params(f, "Visco")


@pytest.mark.parametrize("cells", [True, False])
def test_load_acoustic_source_receiver(ads, fx, tmp_path, cells):
    f = write_mat(str(tmp_path / "m.mat"), fx, fx["vp_true"], cells=cells)
    src = ads.load_acoustic_source(f)
    rcv = ads.load_acoustic_receiver(f)
    assert len(src) == len(rcv) == 8
    for s in range(8):
        assert src[s].srci.tolist() == [fx["src_ix"][s]] and src[s].srcj.tolist() == [fx["src_iy"][s]]
        assert src[s].srcv.shape == (fx["nt"], 1)
        assert np.array_equal(src[s].srcv[:, 0], fx["vec"])
        assert np.array_equal(rcv[s].rcvi, fx["rcv_ix"]) and np.array_equal(rcv[s].rcvj, fx["rcv_iy"])
    # the fixture's own geometry: sources at depth 2 cells, receivers on one row (1-based)
    assert len(fx["rcv_ix"]) == 183


def test_wavelet_row_vector_and_elastic_transpose(ads, fx, tmp_path):
    f = write_mat(str(tmp_path / "r.mat"), fx, fx["vp_true"], vec_as_row=True)
    src = ads.load_acoustic_source(f)  # 1 x nt is reshaped to a column (IO.jl:148-150)
    assert src[0].srcv.shape == (fx["nt"], 1)
    fe = write_mat(str(tmp_path / "e.mat"), fx, fx["vp_true"], elastic=True, vec_as_row=True, nshots=3)
    es = ads.load_elastic_source(fe)
    er = ads.load_elastic_receiver(fe)
    assert len(es) == len(er) == 3
    assert es[1].srcv.shape == (fx["nt"], 1) and np.array_equal(es[1].srcv[:, 0], fx["vec"])
    assert es[1].srctype.tolist() == [2]
    assert np.array_equal(er[2].rcvtype, fx["rcv_type"])


def test_wrong_wavelet_length_is_an_error(ads, fx, tmp_path):
    bad = dict(fx, vec=fx["vec"][:-5])
    f = write_mat(str(tmp_path / "b.mat"), bad, fx["vp_true"])
    with pytest.raises(ValueError):
        ads.load_acoustic_source(f)


def test_safe_vec_and_single_struct(ads, fx, tmp_path):
    assert ads.safe_vec(3).tolist() == [3]
    assert ads.safe_vec(np.array([[1, 2], [3, 4]])).tolist() == [1, 3, 2, 4]  # column-major
    one = dict(fx, rcv_ix=fx["rcv_ix"][:1], rcv_iy=fx["rcv_iy"][:1], rcv_type=fx["rcv_type"][:1])
    f = write_mat(str(tmp_path / "one.mat"), one, fx["vp_true"], nshots=1)
    rcv = ads.load_acoustic_receiver(f)
    src = ads.load_acoustic_source(f)
    assert len(rcv) == 1 and rcv[0].rcvi.tolist() == [fx["rcv_ix"][0]]
    assert len(src) == 1 and src[0].srci.tolist() == [fx["src_ix"][0]]


def test_mpi_elastic_params_and_block_sources(ads, fx, tmp_path):
    import scipy.io as sio
    f = str(tmp_path / "mpi.mat")
    sio.savemat(f, {"nx": np.array([[400]]), "ny": np.array([[200]]), "nt": np.array([[100]]),
                    "dx": np.array([[10.0]]), "dy": np.array([[10.0]]), "dt": np.array([[1e-3]]),
                    "node_per_cpu": np.array([[100]]),
                    "source": np.array([[({"ix": np.array([[150.0]]), "iy": np.array([[120.0]]),
                                           "type": np.array([[2.0]]),
                                           "vec": np.ones((1, 100))},)]],
                                       dtype=[("s", object)])["s"]})
    p = ads.load_params(f, "MPIElastic")
    assert (p.NX, p.NY, p.n, p.M, p.N) == (400, 200, 100, 4, 2)
    ads.compute_PML_Params_mpi(p, rank=3)  # block (II, JJ) = (2, 2)
    src = ads.load_elastic_source(f, use_mpi=True, param=p)
    # global (150, 120) lies in block (2, 2): local 1-based index + 2 (MPIElastic.jl:85-86)
    assert src[0].srci.tolist() == [150 - 100 + 2] and src[0].srcj.tolist() == [120 - 100 + 2]
