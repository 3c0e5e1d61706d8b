"""CPU checks of the C-ABI library: it loads, exports every declared symbol, and its host-side
validation and coefficient tables behave (no device compute here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "adseismic.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ads_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(ads):
    lib = ads.load_library()
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_declarations(ads):
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    assert set(declared()) <= set(L.declared_symbols())


def test_version_and_error_text(ads):
    lib = ads.load_library()
    assert lib.ads_version() >= 100


def _params(ads, **kw):
    p = ads.AcousticPropagatorParams(NX=20, NY=30, NSTEP=50, NPOINTS_PML=5)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_plan_rejects_out_of_range_source(ads):
    p = _params(ads)
    src = ads.AcousticSource([23], [5], np.zeros((50, 1)))  # NX+2 = 22
    rcv = ads.AcousticReceiver([3], [3])
    with pytest.raises(ads.ADSError) as e:
        ads.AcousticPlan(p, [src], [rcv])
    assert e.value.code == 3 and "source" in str(e.value)


def test_plan_rejects_out_of_range_receiver(ads):
    p = _params(ads)
    src = ads.AcousticSource([3], [5], np.zeros((50, 1)))
    rcv = ads.AcousticReceiver([3, 0], [3, 3])  # 1-based: 0 is outside
    with pytest.raises(ads.ADSError) as e:
        ads.AcousticPlan(p, [src], [rcv])
    assert e.value.code == 3 and "receiver" in str(e.value)


def test_plan_rejects_bad_shapes_and_params(ads):
    src = ads.AcousticSource([3], [5], np.zeros((50, 1)))
    rcv = ads.AcousticReceiver([3], [3])
    with pytest.raises(ads.ADSError) as e:
        ads.AcousticPlan(_params(ads, NSTEP=0), [src], [rcv])
    assert e.value.code == 2
    with pytest.raises(ads.ADSError) as e:
        ads.AcousticPlan(_params(ads, PropagatorKernel=5), [src], [rcv])
    assert e.value.code == 1
    with pytest.raises(ads.ADSError) as e:
        ads.AcousticPlan(_params(ads), [src], [rcv], nstep=[60])
    assert e.value.code == 2


def test_set_shape_validates_before_touching_the_plan(ads):
    """Shapes are checked on the host: rows 1/2/4 (the forward also 8/16, its streaming
    rectangle), shots >= 1, batch 1/2 (2 only with 1-2 rows)."""
    import ctypes as C
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    lib = L.load()
    for bad in [(3, 1, 1, 1, 1, 1), (1, 0, 1, 1, 1, 1), (1, 1, 3, 1, 1, 1), (1, 1, 1, 4, 2, 2),
                (1, 1, 1, 8, 1, 1), (8, 2, 2, 1, 1, 1), (12, 1, 1, 1, 1, 1)]:
        rc = lib.ads_acoustic_plan_set_shape(None, (C.c_int * 6)(*bad))
        assert rc == 1, bad
        assert b"must" in lib.ads_last_error() or b"needs" in lib.ads_last_error()
    for good in [(1, 8, 2, 2, 8, 2), (8, 16, 1, 2, 16, 1), (16, 4, 1, 1, 4, 1)]:
        rc = lib.ads_acoustic_plan_set_shape(None, (C.c_int * 6)(*good))
        assert rc == 1 and b"null plan" in lib.ads_last_error()  # a valid shape still needs a plan


def test_cfl_check(ads):
    p = _params(ads, DELTAT=1e-3, DELTAX=10.0, DELTAY=10.0)
    ads.check_cfl(p, 2000.0)  # 0.28
    with pytest.raises(ads.ADSError) as e:
        ads.check_cfl(p, 8000.0)  # 1.13
    assert e.value.code == 4


def test_acoustic_pml_profiles_match_oracle(ads, orc):
    p = _params(ads, NX=40, NY=33, DELTAX=5.0, DELTAY=7.0, NPOINTS_PML=9, vp_ref=1800.0)
    ads.compute_PML_Params_(p)
    Sx, Sy = orc.acoustic_pml(40, 33, 5.0, 7.0, 9, 1e-3, 1800.0)
    np.testing.assert_array_equal(p.Σx, Sx)
    np.testing.assert_array_equal(p.Σy, Sy)
    # index sets follow getid2 (Core.jl:645-653): 1-based, row-major
    assert p.IJ[0] == 1 * (33 + 2) + 2 and p.IpJ[0] == p.IJ[0] + 35 and p.InJ[0] == p.IJ[0] - 35
    assert len(p.IJ) == 40 * 33


def test_elastic_pml_params_match_oracle(ads, orc):
    import importlib
    L = importlib.import_module("adseismic_amd._lib")
    ep = ads.ElasticPropagatorParams(NX=40, NY=30, DELTAX=10.0, DELTAY=8.0, DELTAT=1e-3,
                                     NPOINTS_PML=7, f0=6.0, vp_ref=2500.0, K_MAX_PML=1.5)
    pc = ep.to_c()
    outs = [np.zeros(2 * 40) for _ in range(3)] + [np.zeros(2 * 30) for _ in range(3)]
    L.check(ads.load_library().ads_elastic_pml_params(C.byref(pc), *[o.ctypes.data for o in outs]))
    ref = orc.elastic_pml(40, 30, 10.0, 8.0, 1e-3, 7, 2.0, 1.5, ep.ALPHA_MAX_PML, 1e-3, 2500.0)
    for o, r in zip(outs, ref):
        np.testing.assert_array_equal(o, r.reshape(-1))


def test_geometry_packing(ads):
    s1 = ads.AcousticSource([3, 4], [5, 6], np.zeros((10, 2)))
    s2 = ads.AcousticSource([7], [8], np.zeros((10, 1)))
    r1 = ads.AcousticReceiver([1, 2, 3], [1, 1, 1])
    r2 = ads.AcousticReceiver([9], [9])
    g = ads.Geometry([s1, s2], [r1, r2])
    assert list(g.src_ptr) == [0, 2, 3] and list(g.rcv_ptr) == [0, 3, 4]
    assert list(g.srci) == [3, 4, 7] and list(g.rcvj) == [1, 1, 1, 9]
    with pytest.raises(ValueError):
        ads.Geometry([s1], [r1, r2])
