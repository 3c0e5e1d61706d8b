"""CPU checks of the checkpoint schedule (ads_checkpoint_schedule, host code of the C-ABI
library): replayed through a state machine, every schedule restores only states its slots
hold, saves only the current state, keeps at most S slots, runs each segment's adjoint once in
reverse order with that segment's levels in the ring, and rebuilds the fewest segments an
exhaustive search over every split finds (the binomial optimum)."""
import ctypes as C
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib(ads):
    return ads.load_library()


def ops_of(lib, M, S):
    L = importlib.import_module("adseismic_amd._lib")
    n = C.c_int64()
    rc = lib.ads_checkpoint_schedule(M, S, None, 0, C.byref(n))
    assert rc in (0, L.ADS_ERR_SHAPE)
    buf = np.zeros(3 * max(n.value, 1), dtype=np.int64)
    L.check(lib.ads_checkpoint_schedule(M, S, buf.ctypes.data_as(C.POINTER(C.c_int64)), n.value,
                                        C.byref(n)))
    return buf[:3 * n.value].reshape(-1, 3)


def min_rebuilds(n, s, memo={}):
    key = (n, s)
    if key not in memo:
        if n == 1:
            memo[key] = 1
        elif s == 0:
            memo[key] = n * (n + 1) // 2
        else:
            memo[key] = min(j + min_rebuilds(n - j, s - 1) + min_rebuilds(j, s) for j in range(1, n))
    return memo[key]


def replay(ops, M, S):
    """Checks the invariants; returns the segments advanced in the adjoint sweep."""
    slots = {}
    cur = M                      # the forward pass ends after the last segment
    ring = M - 1 if M else None  # segment whose levels the slot ring holds
    adj = []
    advanced = 0
    fwd_states = [(m, q) for k, m, q in ops if k == 4]
    assert [m for m, _ in fwd_states] == sorted(m for m, _ in fwd_states)
    for m, q in fwd_states:
        assert 0 < m < M and 0 <= q < S
        slots[q] = m
    for k, a, b in ops:
        if k == 4:
            continue
        if k == 0:                        # restore state a from slot b
            assert (b < 0 and a == 0) or slots.get(b) == a, (a, b, slots)
            cur, ring = a, None
        elif k == 1:                      # advance a -> b
            assert cur == a and a < b <= M
            advanced += b - a
            cur, ring = b, b - 1
        elif k == 2:                      # save state a in slot b
            assert cur == a and 0 <= b < S
            slots[b] = a
        elif k == 3:                      # adjoint of segment a
            assert ring == a, (a, ring)
            adj.append(a)
            ring = None
        else:
            raise AssertionError(k)
    assert adj == list(range(M - 1, -1, -1))
    assert len(set(q for q in slots)) <= max(S, 0)
    return advanced


@pytest.mark.parametrize("M", list(range(1, 31)) + [57, 100])
def test_schedule_is_valid_and_optimal(lib, M):
    for S in sorted(set([0, 1, 2, 3, 4, 6, 9, max(M - 2, 0), max(M - 1, 0), M + 3])):
        s_eff = min(S, max(M - 1, 0))
        ops = ops_of(lib, M, S)
        advanced = replay(ops, M, s_eff)
        if M:  # the forward pass made M of the advances the search counts
            assert advanced == min_rebuilds(M, s_eff) - M, (M, S, advanced)


def test_one_level_keeps_every_segment_start(lib):
    M = 9
    ops = ops_of(lib, M, -1)
    assert [tuple(r[1:]) for r in ops if r[0] == 4] == [(m, m - 1) for m in range(1, M)]
    assert sum(r[2] - r[1] for r in ops if r[0] == 1) == M - 1
