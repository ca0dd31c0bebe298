"""GPU tests of the continuous-batching server.

Every request's y must be the same, bit for bit, whether the request runs
alone or in a mix with others (admitted into shared launches, submitted from
several threads), and equal hbdiaSpMV on its matrix with y starting at zero
and the oracle.  Also: the plain form, device-resident vectors, admission
under a tight HBM budget and rejection of requests that can never fit.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA, DeviceArray, HBDIAError, HBDIAServer, HBDIAVector, gen_vector, hbdiaSpMV

pytestmark = pytest.mark.gpu


def pool(dtype=np.float64):
    ms = [HBDIA.stencil(8, 27, 0.01, seed=1, dtype=dtype), HBDIA.stencil(13, 7, 0.02, seed=2, dtype=dtype),
          HBDIA.stencil(20, 27, 0.005, seed=3, dtype=dtype), HBDIA.banded(3000, 5, 64, 0.01, seed=4, dtype=dtype),
          HBDIA.banded(777, 9, 64, 0.0, seed=5, dtype=dtype), HBDIA.stencil(16, 7, 0.0, seed=6, dtype=dtype)]
    for i, m in enumerate(ms):
        m.convertToHBDIAFormatOnHost([64, 128, 256][i % 3], 2)
    return ms


def xs_for(ms, dtype, variants=3):
    return [[gen_vector(m.getNumRows(), 100 * i + v, dtype=dtype) for v in range(variants)] for i, m in enumerate(ms)]


def expected(A: HBDIA, x: np.ndarray) -> np.ndarray:
    hb = A.getHBDIA()
    hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), A.getBlockWidth()
    return oracle.hbdia_spmv(hb, x.copy(), np.zeros_like(x), A.getNumRows())


def single(A: HBDIA, x: np.ndarray) -> np.ndarray:
    A.prepareForGPU()
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros_like(x))
    hbdiaSpMV(A, X, Y, False, True)
    A.synchronize()
    return Y.download()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_same_result_alone_and_in_the_mix(gpu, dtype):
    print(f"case: serve alone vs mix {np.dtype(dtype).name}", flush=True)
    ms = pool(dtype)
    xs = xs_for(ms, dtype)
    S = HBDIAServer(ms, max_jobs=64, max_in_flight=2)
    # alone: one request at a time
    alone = {}
    for m in range(len(ms)):
        for v in range(3):
            y = np.full(ms[m].getNumRows(), np.nan, dtype)
            S.wait(S.submit(m, xs[m][v], y))
            alone[m, v] = y
            np.testing.assert_array_equal(y, expected(ms[m], xs[m][v]), err_msg=f"matrix {m} vs oracle")
    for m in range(len(ms)):
        np.testing.assert_array_equal(alone[m, 0], single(ms[m], xs[m][0]), err_msg=f"matrix {m} vs hbdiaSpMV")
    # in the mix: a burst of requests from four threads at once
    rng = np.random.default_rng(7)
    picks = [(int(rng.integers(len(ms))), int(rng.integers(3))) for _ in range(240)]
    outs = [np.full(ms[m].getNumRows(), np.nan, dtype) for m, _ in picks]
    ids = [None] * len(picks)
    launches_before = S.info()["launches"]

    def submitter(k):
        for i in range(k, len(picks), 4):
            m, v = picks[i]
            ids[i] = S.submit(m, xs[m][v], outs[i])

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    S.drain()
    for i, (m, v) in enumerate(picks):
        np.testing.assert_array_equal(outs[i], alone[m, v], err_msg=f"request {i} (matrix {m}) differs in the mix")
    info = S.info()
    assert info["completed"] == 18 + len(picks)
    assert info["launches"] - launches_before < len(picks)  # requests shared launches
    assert info["largest_launch"] > 1
    for rid in ids:
        ts, ta, tc = S.times(rid)
        assert ts <= ta <= tc
    S.close()


def test_plain_form_and_device_vectors_agree(gpu):
    print("case: serve plain form and device vectors", flush=True)
    ms = pool()
    xs = xs_for(ms, np.float64, variants=1)
    P = HBDIAServer(ms, batching=False)
    B = HBDIAServer(ms)
    for m in range(len(ms)):
        want = expected(ms[m], xs[m][0])
        y1 = np.zeros(ms[m].getNumRows())
        P.wait(P.submit(m, xs[m][0], y1))
        np.testing.assert_array_equal(y1, want)
        X = DeviceArray.from_host(xs[m][0])
        Y = DeviceArray(ms[m].getNumRows(), np.float64, zero=True)
        B.wait(B.submit(m, X, Y, device=True))
        np.testing.assert_array_equal(Y.download(), want)
    assert P.info()["largest_launch"] == 1
    P.close()
    B.close()


def test_tight_budget_queues_and_rejects(gpu):
    print("case: serve tight budget", flush=True)
    ms = pool()
    xs = xs_for(ms, np.float64, variants=1)
    probe = HBDIAServer(ms)
    resident = probe.info()["resident_bytes"]
    probe.close()
    big = max(m.getNumRows() for m in ms)  # 8000 rows (20^3)
    # room for the x and y of two of the largest requests
    budget = resident + 2 * 2 * 8 * (big + 64)
    S = HBDIAServer(ms, hbm_budget=budget)
    assert S.info()["arena_bytes"] <= budget - resident
    outs, ids, want = [], [], []
    for i in range(60):
        m = i % len(ms)
        y = np.zeros(ms[m].getNumRows())
        outs.append(y)
        ids.append(S.submit(m, xs[m][0], y))
        want.append(m)
    S.drain()
    for y, m in zip(outs, want):
        np.testing.assert_array_equal(y, expected(ms[m], xs[m][0]))
    S.close()
    # a budget whose arena cannot hold the largest matrix's vectors: that request is refused
    small = resident + 2 * 8 * 4096
    T = HBDIAServer(ms, hbm_budget=small)
    with pytest.raises(HBDIAError, match="exceed the request arena"):
        T.submit(2, xs[2][0], np.zeros(ms[2].getNumRows()))
    y = np.zeros(ms[0].getNumRows())
    T.wait(T.submit(0, xs[0][0], y))  # a small one still runs
    np.testing.assert_array_equal(y, expected(ms[0], xs[0][0]))
    T.close()


def test_open_loop_driver_reports_latencies(gpu):
    print("case: serve open loop", flush=True)
    ms = pool()
    xs = xs_for(ms, np.float64, variants=2)
    S = HBDIAServer(ms)
    flat = [x for per in xs for x in per]
    ys = [np.zeros(m.getNumRows()) for m in ms]
    lat, wall = S.open_loop(2000.0, 200, 3, flat, 2, ys)
    assert lat.shape == (200,) and np.all(lat > 0) and wall > 0
    lat0, _ = S.open_loop(0.0, 100, 4, flat, 2, ys)  # burst
    assert np.all(lat0 > 0)
    S.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_serving_pool_of_configured_shapes(gpu, dtype):
    """A pool shaped like the serving workload's: 27-point stencils at 32^3, 48^3 and 64^3, a 7-point
    stencil at 40^3 and banded matrices of the batching set (up to 1e5 rows), widths 64 / 128.
    Every request's y equals the oracle's alone and in a 4-thread burst of 120 requests."""
    print(f"case: serve configured shapes {np.dtype(dtype).name}", flush=True)
    ms = [HBDIA.stencil(32, 27, 0.0, seed=11, dtype=dtype), HBDIA.stencil(48, 27, 0.005, seed=12, dtype=dtype),
          HBDIA.stencil(64, 27, 0.0, seed=13, dtype=dtype), HBDIA.stencil(40, 7, 0.0, seed=14, dtype=dtype)]
    ms += HBDIA.banded_batch(count=4, min_rows=20000, max_rows=100000, seed=15, dtype=dtype)
    for i, m in enumerate(ms):
        m.convertToHBDIAFormatOnHost([64, 128][i % 2], 2)
    xs = xs_for(ms, dtype, variants=2)
    want = {(m, v): expected(ms[m], xs[m][v]) for m in range(len(ms)) for v in range(2)}
    S = HBDIAServer(ms, max_jobs=32, max_in_flight=2)
    for (m, v), w in want.items():
        y = np.full(ms[m].getNumRows(), np.nan, dtype)
        S.wait(S.submit(m, xs[m][v], y))
        assert y.tobytes() == w.tobytes(), f"matrix {m} alone differs from the oracle"
    rng = np.random.default_rng(21)
    picks = [(int(rng.integers(len(ms))), int(rng.integers(2))) for _ in range(120)]
    outs = [np.full(ms[m].getNumRows(), np.nan, dtype) for m, _ in picks]

    def submitter(k):
        for i in range(k, len(picks), 4):
            m, v = picks[i]
            S.submit(m, xs[m][v], outs[i])

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    S.drain()
    for i, (m, v) in enumerate(picks):
        assert outs[i].tobytes() == want[m, v].tobytes(), f"request {i} (matrix {m}) differs in the mix"
    assert S.info()["largest_launch"] > 1
    S.close()


def test_submit_rejects_short_vectors(gpu):
    ms = pool()
    S = HBDIAServer(ms)
    n = S.num_rows(2)
    with pytest.raises(ValueError):
        S.submit(2, np.zeros(n - 1), np.zeros(n))
    with pytest.raises(ValueError):
        S.submit(2, np.zeros(n), np.zeros(n - 1))
    with pytest.raises(TypeError):
        S.submit(2, np.zeros(n, np.float32), np.zeros(n))
    S.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_vectors_in_place_in_the_mix(gpu, dtype):
    """Batched requests whose vectors are on the device are read and written in place, with y starting
    as NaN: the serve kernel writes every row (rows past a matrix's blocks as zero plus their fallback),
    so nothing is cleared.  Every request equals the oracle with y starting at zero, alone and in a
    burst from four threads, each request with its own y."""
    print(f"case: serve device vectors in place {np.dtype(dtype).name}", flush=True)
    ms = pool(dtype)
    # 300 x 300 with entries in the first 100 rows only: the blocks of rows 128.. hold no offsets,
    # and R's lone entry in row 250 is a fallback entry in such a block
    rows = np.array([0, 0, 1, 63, 64, 64, 99, 99, 250], np.int32)
    cols = np.array([0, 1, 1, 63, 64, 63, 99, 98, 7], np.int32)
    vals = np.array([2.0, 1.0, 3.0, 4.0, 5.0, 6.0, 1.5, -2.0, 0.5], dtype)
    R = HBDIA(rows, cols, vals, 300, 300, dtype=dtype)
    R.convertToHBDIAFormatOnHost(64, 2)
    assert R.info()["num_coo"] > 0
    ms.append(R)
    R2 = HBDIA(rows[:-1], cols[:-1], vals[:-1], 300, 300, dtype=dtype)
    R2.convertToHBDIAFormatOnHost(64, 1)
    assert R2.info()["num_coo"] == 0
    ms.append(R2)
    xs = xs_for(ms, dtype, variants=2)
    want = {(m, v): expected(ms[m], xs[m][v]) for m in range(len(ms)) for v in range(2)}
    S = HBDIAServer(ms, max_jobs=48, max_in_flight=2)
    dx = {(m, v): DeviceArray.from_host(xs[m][v]) for m in range(len(ms)) for v in range(2)}
    for (m, v), w in want.items():
        Y = DeviceArray.from_host(np.full(ms[m].getNumRows(), np.nan, dtype))
        S.wait(S.submit(m, dx[m, v], Y, device=True))
        assert Y.download().tobytes() == w.tobytes(), f"matrix {m} alone differs from the oracle"
    rng = np.random.default_rng(31)
    picks = [(int(rng.integers(len(ms))), int(rng.integers(2))) for _ in range(160)]
    outs = [DeviceArray.from_host(np.full(ms[m].getNumRows(), np.nan, dtype)) for m, _ in picks]

    def submitter(k):
        for i in range(k, len(picks), 4):
            m, v = picks[i]
            S.submit(m, dx[m, v], outs[i], device=True)

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    S.drain()
    for i, (m, v) in enumerate(picks):
        assert outs[i].download().tobytes() == want[m, v].tobytes(), f"request {i} (matrix {m}) differs in the mix"
    assert S.info()["largest_launch"] > 1
    S.close()


def test_host_and_device_requests_mixed_in_launches(gpu):
    """Requests with host vectors (copied through the arena) and with device vectors (used in place)
    submitted together from four threads share launches; each completes with its own y final and
    equal to the oracle, and the server's completion order stays the submission order."""
    print("case: serve host and device requests mixed", flush=True)
    ms = pool()
    xs = xs_for(ms, np.float64, variants=2)
    want = {(m, v): expected(ms[m], xs[m][v]) for m in range(len(ms)) for v in range(2)}
    dx = {(m, v): DeviceArray.from_host(xs[m][v]) for m in range(len(ms)) for v in range(2)}
    S = HBDIAServer(ms, max_jobs=24, max_in_flight=2)
    rng = np.random.default_rng(41)
    picks = [(int(rng.integers(len(ms))), int(rng.integers(2)), bool(rng.integers(2))) for _ in range(200)]
    outs = [DeviceArray.from_host(np.full(ms[m].getNumRows(), np.nan)) if dev else
            np.full(ms[m].getNumRows(), np.nan) for m, _, dev in picks]
    ids = [None] * len(picks)

    def submitter(k):
        for i in range(k, len(picks), 4):
            m, v, dev = picks[i]
            ids[i] = S.submit(m, dx[m, v] if dev else xs[m][v], outs[i], device=dev)

    th = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    S.drain()
    for i, (m, v, dev) in enumerate(picks):
        got = outs[i].download() if dev else outs[i]
        assert got.tobytes() == want[m, v].tobytes(), f"request {i} (matrix {m}, device={dev}) differs"
    done_at = sorted((S.times(r)[2], r) for r in ids)
    assert [r for _, r in done_at] == sorted(ids)  # completion in submission order
    assert S.info()["largest_launch"] > 1
    S.close()
