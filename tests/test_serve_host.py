"""Host-side tests of the serving engine's admission memory: the first-fit
arena that places each admitted request's x and y in the request arena."""
from __future__ import annotations

import random

import pytest

from hbdia_amd import HBDIAError, VectorArena


def test_arena_alignment_first_fit_and_exhaustion():
    a = VectorArena(1024, 64)
    assert a.allocate(1) == 0
    assert a.allocate(65) == 64          # rounded up to 128 units
    assert a.allocate(64) == 192
    assert a.info()["used"] == 64 + 128 + 64
    assert a.allocate(1024) == -1        # does not fit: refused, nothing changes
    assert a.info()["used"] == 256
    a.release(64)                        # a hole of 128 at 64
    assert a.allocate(100) == 64         # first fit reuses the hole
    assert a.allocate(10_000) == -1


def test_arena_coalesces_in_any_release_order():
    rng = random.Random(5)
    a = VectorArena(64 * 40, 64)
    offs = [a.allocate(64) for _ in range(40)]
    assert sorted(offs) == [64 * i for i in range(40)]
    assert a.allocate(1) == -1
    rng.shuffle(offs)
    for o in offs:
        a.release(o)
    info = a.info()
    assert info["used"] == 0 and info["free_blocks"] == 1 and info["largest_free"] == 64 * 40


def test_arena_against_a_brute_force_model():
    # random allocate / release against an explicit unit map with first-fit semantics
    rng = random.Random(11)
    cap, align = 64 * 64, 64
    a = VectorArena(cap, align)
    owner = [None] * (cap // align)  # per aligned unit: offset of the block holding it
    live = {}
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off = rng.choice(list(live))
            a.release(off)
            for u in range(off // align, (off + live.pop(off)) // align):
                owner[u] = None
        else:
            n = rng.randint(1, 8 * align)
            need = -(-n // align)
            want = -1
            for s in range(len(owner) - need + 1):  # first fit
                if all(owner[s + k] is None for k in range(need)):
                    want = s * align
                    break
            got = a.allocate(n)
            assert got == want
            if got >= 0:
                live[got] = need * align
                for k in range(need):
                    owner[got // align + k] = got
        assert a.info()["used"] == sum(live.values())


def test_arena_rejects_bad_calls():
    a = VectorArena(256, 64)
    with pytest.raises(HBDIAError):
        a.release(64)  # never allocated
    with pytest.raises(HBDIAError):
        a.allocate(0)
    with pytest.raises(HBDIAError):
        VectorArena(-1, 64)


def _arena_elems(budget: int, resident: int, elem: int) -> int:
    import ctypes

    from hbdia_amd.api import capi, check_capi
    out = ctypes.c_int64(0)
    check_capi(capi.hbdia_serve_arena_elems(budget, resident, elem, ctypes.byref(out)))
    return int(out.value)


def test_serve_arena_is_capped_below_int32_rows():
    # the serve kernel adds a request's slot shift to int32 pool rows: the
    # arena must stay below 2^31 elements whatever the HBM budget
    gib = 1 << 30
    assert _arena_elems(resident=gib, budget=gib + (1 << 20), elem=8) == (1 << 20) // 16
    assert _arena_elems(resident=0, budget=64 * gib, elem=8) == (1 << 31) - 64      # 64 GiB FP64 budget
    assert _arena_elems(resident=5 * gib, budget=280 * gib, elem=4) == (1 << 31) - 64
    assert _arena_elems(resident=0, budget=(1 << 31) * 16 - 1, elem=8) % 64 == 0
    with pytest.raises(HBDIAError):
        _arena_elems(resident=gib, budget=gib, elem=8)       # budget does not cover the matrices
    with pytest.raises(HBDIAError):
        _arena_elems(resident=gib, budget=gib + 1000, elem=8)  # room for fewer than 64 elements
