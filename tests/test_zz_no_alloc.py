"""The allocation-free AddSource entry point (ads_add_source_noalloc_f64) against the
allocating one (ads_add_source_f64), through the C-ABI: bit-identical outputs on two calls in
a row, the second with fewer sources (the fixed per-device table keeps the first call's
entries beyond the second call's count, which the kernel must not read), and a clear error for
a source count above the fixed table size.  AddSource semantics: AddSource.cpp:33-87.
"""
import ctypes as C
import importlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NX, NY = 13, 11
N = (NX + 2) * (NY + 2)


def P(t):
    return C.c_void_p(t.data_ptr())


def H(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


@pytest.fixture(scope="module")
def L(ads):
    ads.load_library()
    return importlib.import_module("adseismic_amd._lib")


def _call(lib, L, fn, fields, outs, si, sj, st, sv):
    a, ap = H(si)
    b, bp = H(sj)
    c, cp = H(st)
    rc = getattr(lib, fn)(*[P(t) for t in fields], ap, bp, cp, len(si), NX, NY, P(sv),
                          *[P(t) for t in outs], None)
    L.check(rc)


def _sources(rng, n):
    si = rng.integers(1, NX + 3, n)
    sj = rng.integers(1, NY + 3, n)
    st = rng.integers(-1, 6, n)          # every field type plus types the op skips
    si[1], sj[1], st[1] = si[0], sj[0], st[0]  # a duplicated (cell, type) pair
    return si, sj, st


def test_noalloc_add_source_matches_allocating_entry_point(ads, L, cuda):
    lib = ads.load_library()
    rng = np.random.default_rng(7)
    for call, n in enumerate((300, 40)):  # 300 > one staging block of 256 entries
        si, sj, st = _sources(rng, n)
        srcv = torch.as_tensor(rng.standard_normal(n), dtype=torch.float64, device="cuda")
        f0 = [torch.as_tensor(rng.standard_normal(N), dtype=torch.float64, device="cuda")
              for _ in range(5)]
        want = [torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(5)]
        got = [torch.full((N,), np.nan, dtype=torch.float64, device="cuda") for _ in range(5)]
        _call(lib, L, "ads_add_source_f64", f0, want, si, sj, st, srcv)
        _call(lib, L, "ads_add_source_noalloc_f64", f0, got, si, sj, st, srcv)
        inplace = [t.clone() for t in f0]
        _call(lib, L, "ads_add_source_noalloc_f64", inplace, inplace, si, sj, st, srcv)
        torch.cuda.synchronize()
        changed = 0
        for q in range(5):
            w, g = want[q].cpu().numpy(), got[q].cpu().numpy()
            np.testing.assert_array_equal(g, w, err_msg=f"call {call}, field {q}")
            np.testing.assert_array_equal(inplace[q].cpu().numpy(), w)
            changed += int(np.sum(w != f0[q].cpu().numpy()))
        assert changed > 0, "no source landed"


def test_noalloc_add_source_rejects_too_many_sources(ads, L, cuda):
    lib = ads.load_library()
    n = 65536 + 1
    f = [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(5)]
    out = [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(5)]
    si, sip = H(np.full(n, 2))
    sj, sjp = H(np.full(n, 2))
    st, stp = H(np.zeros(n))
    sv = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = lib.ads_add_source_noalloc_f64(*[P(t) for t in f], sip, sjp, stp, n, NX, NY, P(sv),
                                        *[P(t) for t in out], None)
    assert rc == L.ADS_ERR_SHAPE
    assert b"exceed the fixed table" in lib.ads_last_error()
    # a bad index is rejected before any output is written
    bad_i, bip = H([2, NX + 3])
    bad_j, bjp = H([2, 2])
    bad_t, btp = H([0, 0])
    out2 = [torch.full((N,), 5.0, dtype=torch.float64, device="cuda") for _ in range(5)]
    rc = lib.ads_add_source_noalloc_f64(*[P(t) for t in f], bip, bjp, btp, 2, NX, NY, P(sv),
                                        *[P(t) for t in out2], None)
    assert rc == L.ADS_ERR_INDEX
    torch.cuda.synchronize()
    assert all(bool(torch.all(t == 5.0)) for t in out2)
