"""GPU parity of the HIP path against the oracle (CPU restatement of the
reference) on identical seeded inputs.

Expected agreement: bit-exact.  The blocked-DIA part sums each row over its
block's offsets in ascending order as FMAs (the reference kernel's order), and
the COO part sums each row's fallback entries sequentially in entry order (the
reference CPU path's order), then y = bdia + coo, so every rounding step is
the same as the oracle's.  The stated tolerance of the path, 1e-12 relative
per row in FP64, is therefore met with zero error; tests assert equality.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA, DeviceArray, ExecutionMode, Graph, HBDIAError, HBDIAVector, Stream, gen_vector, hbdiaSpMV
from hbdia_amd import kernels
from hbdia_amd.device import synchronize

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def oracle_y(A: HBDIA, x: np.ndarray, y0: np.ndarray) -> np.ndarray:
    hb = A.getHBDIA()
    hb["num_blocks"] = A.getNumBlocks()
    hb["block_width"] = A.getBlockWidth()
    return oracle.hbdia_spmv(hb, x.copy(), y0.copy(), A.getNumRows())


def run_spmv(A: HBDIA, x: np.ndarray, y0: np.ndarray, cpu: bool, gpu_flag: bool) -> np.ndarray:
    X = HBDIAVector(x)
    Y = HBDIAVector(y0)
    assert hbdiaSpMV(A, X, Y, cpu, gpu_flag)
    A.synchronize()
    return Y.download()


def test_config1_reference_example_path(gpu, golden):
    print("case: config1 8^3 bw32 thr2 execCOOCPU", flush=True)
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8, 0.0)
    A.convertToHBDIAFormat(32, 2)
    x = np.arange(1, 513, dtype=np.float64)
    y = run_spmv(A, x, np.zeros(512), True, False)
    np.testing.assert_array_equal(y, golden("config1_y.npz")["y"])
    assert y[0] == pytest.approx(-73.14, rel=1e-13) and y[511] == pytest.approx(11610.51, rel=1e-13)


def test_example_binary(gpu):
    print("case: simple_hbdia_example binary", flush=True)
    exe = ROOT / "hbdia_amd" / "lib" / "simple_hbdia_example"
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASS" in out.stdout


@pytest.mark.parametrize("name,n,noise,it,bw,thr,xk", [("noisy16_y.npz", 16, 0.1, 0, 32, 2, "i+1"),
                                                       ("noisy24_y.npz", 24, 0.02, 1, 16, 3, "sin")])
@pytest.mark.parametrize("cpu_flag", [False, True])
def test_golden_noisy(gpu, golden, name, n, noise, it, bw, thr, xk, cpu_flag):
    print(f"case: golden {name} cpu_coo={cpu_flag}", flush=True)
    A = HBDIA()
    A.create3DStencil27Point(n, n, n, noise, it)
    A.convertToHBDIAFormat(bw, thr)
    N = n ** 3
    x = np.arange(1, N + 1, dtype=np.float64) if xk == "i+1" else np.sin(np.arange(N, dtype=np.float64))
    y = run_spmv(A, x, np.zeros(N), cpu_flag, not cpu_flag)
    np.testing.assert_array_equal(y, golden(name)["y"])


CASES = [
    (16, 0.1, 0, 32, 2), (16, 0.3, 2, 64, 2), (12, 0.5, 3, 128, 1), (20, 0.05, 4, 48, 2),
    (10, 1.0, 5, 1, 1), (17, 0.2, 6, 7, 4), (9, 0.0, 0, 64, 2), (13, 2.0, 7, 256, 30), (11, 0.1, 8, 8, 2),
]


@pytest.mark.parametrize("n,noise,it,bw,thr", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("path", ["gpu_coo", "cpu_coo"])
def test_spmv_bit_exact_vs_oracle(gpu, n, noise, it, bw, thr, dtype, path):
    print(f"case: n={n} noise={noise} it={it} bw={bw} thr={thr} {np.dtype(dtype).name} {path}", flush=True)
    A = HBDIA(dtype=dtype)
    A.create3DStencil27Point(n, n, n, noise, it)
    A.convertToHBDIAFormat(bw, thr)
    N = n ** 3
    x = np.cos(np.arange(N) * 0.731).astype(dtype)
    y0 = np.full(N, 0.5, dtype=dtype)
    y = run_spmv(A, x, y0, path == "cpu_coo", path == "gpu_coo")
    np.testing.assert_array_equal(y, oracle_y(A, x, y0))


def test_both_flags_false_with_fallback_raises(gpu):
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8, 0.2, 0)
    A.convertToHBDIAFormat(32, 2)
    assert A.info()["num_coo"] > 0
    with pytest.raises(RuntimeError, match="Invalid COO execution flags"):
        hbdiaSpMV(A, HBDIAVector(np.ones(512)), HBDIAVector(np.zeros(512)), False, False)
    A.synchronize()


def test_no_fallback_needs_no_flags(gpu):
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8)
    A.convertToHBDIAFormat(32, 2)
    y = run_spmv(A, np.ones(512), np.zeros(512), False, False)
    np.testing.assert_allclose(y, np.bincount(A.getRowIndices(), weights=A.getValues(), minlength=512),
                               rtol=1e-13)


def test_gpu_coo_requires_gpu_mode(gpu):
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8, 0.2, 0)
    A.convertToHBDIAFormat(32, 2, execMode=ExecutionMode.CPU_COO)
    with pytest.raises(Exception, match="not on the device"):
        hbdiaSpMV(A, HBDIAVector(np.ones(512)), HBDIAVector(np.zeros(512)), False, True)
    y = run_spmv(A, np.ones(512), np.zeros(512), True, False)
    np.testing.assert_array_equal(y, oracle_y(A, np.ones(512), np.zeros(512)))


def test_repeated_calls_are_identical(gpu):
    A = HBDIA.stencil(20, 27, 0.01, seed=11)
    A.convertToHBDIAFormat(64, 2)
    N = 20 ** 3
    X = HBDIAVector(gen_vector(N, 3))
    Y = HBDIAVector(np.zeros(N))
    outs = []
    for _ in range(4):
        hbdiaSpMV(A, X, Y, False, True)
        A.synchronize()
        outs.append(Y.download())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("n,bw", [(32, 64), (33, 64), (24, 32)])
def test_stencil_scatter_workload(gpu, n, bw):
    print(f"case: stencil+scatter n={n} bw={bw}", flush=True)
    A = HBDIA.stencil(n, 27, 0.005, seed=0xC2)
    A.convertToHBDIAFormat(bw, 2)
    N = n ** 3
    x = gen_vector(N, 7)
    y = run_spmv(A, x, np.zeros(N), False, True)
    np.testing.assert_array_equal(y, oracle_y(A, x, np.zeros(N)))
    r, c, v = A.getCOO()
    ref = oracle.coo_spmv_longdouble(r, c, v, x, N)
    scale = oracle.coo_spmv_longdouble(r, c, np.abs(v), np.abs(x), N)
    assert np.all(np.abs(y - ref) <= 1e-14 * scale)


# ----------------------------------------------------------- raw C-ABI kernels
def _layout(values, offsets, block_start, N, bw, coo=None):
    if coo is None:
        return kernels.DeviceLayout(np.asarray(values), np.asarray(offsets, np.int32),
                                    np.asarray(block_start, np.int32), N, bw)
    return kernels.DeviceLayout(np.asarray(values), np.asarray(offsets, np.int32), np.asarray(block_start, np.int32),
                                N, bw, *coo)


@pytest.mark.parametrize("bw", [4, 64])
def test_raw_kernel_edge_semantics(gpu, bw):
    """Zero values hide NaN, out-of-range columns are skipped, rows of empty
    blocks and rows past the last block keep y (no fallback), and with a
    fallback every row receives y += coo."""
    rng = np.random.default_rng(bw)
    nb = 3
    N = nb * bw + bw // 2          # rows past the last block
    offs_per = [[-1, 0, 5], [], [-2 * bw, 0, 1]]
    block_start = np.cumsum([0] + [len(o) for o in offs_per]).astype(np.int32)
    offsets = np.array(sum(offs_per, []), np.int32)
    values = rng.standard_normal(len(offsets) * bw)
    values[::7] = 0.0
    x = rng.standard_normal(N)
    x[3] = np.nan
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": nb, "block_width": bw}
    y0 = rng.standard_normal(N)
    L = _layout(values, offsets, block_start, N, bw)
    X, Y = DeviceArray.from_host(x), DeviceArray.from_host(y0)
    L.spmv(X, Y)
    synchronize()
    want = oracle.bdia_spmv(hb, x, y0.copy(), N)
    np.testing.assert_array_equal(Y.download(), want)
    # with a fallback: entries in an empty block, past the last block and in normal rows
    coo_r = np.array(sorted([0, 0, bw + 1, bw + 1, nb * bw + 1, N - 1]), np.int32)
    coo_c = np.array([N - 1, 2, 0, 1, 4, 0], np.int32)
    coo_v = np.array([1.5, -2.0, 0.25, 4.0, 3.0, -1.0])
    L2 = _layout(values, offsets, block_start, N, bw, (coo_r, coo_c, coo_v))
    Y2 = DeviceArray.from_host(y0)
    L2.spmv(X, Y2)
    synchronize()
    hb2 = dict(hb, coo_rows=coo_r, coo_cols=coo_c, coo_values=coo_v)
    want2 = oracle.hbdia_spmv(hb2, x, y0.copy(), N)
    np.testing.assert_array_equal(Y2.download(), want2)


def test_raw_coo_kernel_and_group_pointer(gpu):
    rng = np.random.default_rng(1)
    N, nnz = 5000, 40000
    rows = np.sort(rng.integers(0, N, nnz)).astype(np.int32)
    rows[1000:1300] = rows[1000]            # a long row spanning several 64-entry batches
    rows = np.sort(rows)
    cols = rng.integers(0, N, nnz).astype(np.int32)
    vals = rng.standard_normal(nnz)
    x = rng.standard_normal(N)
    R, C, V = DeviceArray.from_host(rows), DeviceArray.from_host(cols), DeviceArray.from_host(vals)
    G = DeviceArray(kernels.num_groups(N) + 1, np.int32)
    kernels.coo_group_ptr(R, nnz, N, G)
    synchronize()
    np.testing.assert_array_equal(G.download(), np.searchsorted(rows, np.arange(0, kernels.num_groups(N) + 1) * 64))
    res0 = np.full(N, 9.0)
    RES = DeviceArray.from_host(res0)
    kernels.coo_spmv(R, C, V, G, DeviceArray.from_host(x), RES, N)
    synchronize()
    want = oracle.cpu_coo_spmv(rows, cols, vals, x, N, res0.copy())
    np.testing.assert_array_equal(RES.download(), want)


def test_vector_add(gpu):
    a = np.arange(1000.0)
    b = np.sqrt(np.arange(1000.0))
    A, B = DeviceArray.from_host(a), DeviceArray.from_host(b)
    kernels.vector_add(A, B, 1000)
    synchronize()
    np.testing.assert_array_equal(A.download(), a + b)


def test_unified_memory_modes(gpu):
    A = HBDIA()
    A.create3DStencil27Point(10, 10, 10, 0.1, 1)
    A.convertToHBDIAFormat(32, 2)
    N = 1000
    x = np.linspace(-1, 1, N)
    want = oracle_y(A, x, np.zeros(N))
    for flags in [(True, False, True, False), (True, True, False, False), (True, False, False, True)]:
        X = HBDIAVector(x, *flags)
        Y = HBDIAVector(np.zeros(N), *flags)
        for cpu_flag in (True, False):
            Y.upload(np.zeros(N))
            hbdiaSpMV(A, X, Y, cpu_flag, not cpu_flag)
            A.synchronize()
            np.testing.assert_array_equal(Y.download(), want, err_msg=str((flags, cpu_flag)))


def test_graph_capture_replays(gpu):
    A = HBDIA.stencil(24, 27, 0.005, seed=5)
    A.convertToHBDIAFormat(64, 2)
    N = 24 ** 3
    hb = A.getHBDIA()
    s = Stream()
    L = kernels.DeviceLayout(hb["values"], hb["offsets"], hb["block_start"], N, 64, hb["coo_rows"], hb["coo_cols"],
                             hb["coo_values"], stream=s)
    X = DeviceArray.from_host(gen_vector(N, 1))
    Y = DeviceArray(N, np.float64, zero=True)
    L.spmv(X, Y, s)
    s.synchronize()
    eager = Y.download()
    g = Graph.capture(s, lambda: L.spmv(X, Y, s))
    for rep in range(3):
        Y.upload(np.full(N, float(rep)))
        g.launch(s)
        s.synchronize()
        np.testing.assert_array_equal(Y.download(), eager)


# ------------------------------------------------------------ runs-of-three layout hint
def _runs3_case(N, bw, seed, with_coo):
    """A random layout whose blocks hold runs of three consecutive offsets (some partly or wholly
    outside [0, N)), with zero values over NaN x entries."""
    rng = np.random.default_rng(seed)
    nb = -(-N // bw)
    offs_per = []
    for _ in range(nb):
        k = int(rng.integers(0, 6))
        centres = 3 * np.sort(rng.choice(np.arange(-N // 2, N // 2), size=k, replace=False)) + 1
        offs_per.append([int(c + d) for c in centres for d in (-1, 0, 1)])
    block_start = np.zeros(nb + 1, np.int32)
    block_start[1:] = np.cumsum([len(o) for o in offs_per])
    offsets = np.array(sum(offs_per, []), np.int32)
    values = rng.standard_normal(len(offsets) * bw)
    values[rng.random(values.size) < 0.2] = 0.0
    x = rng.standard_normal(N)
    x[rng.random(N) < 0.02] = np.nan
    coo = None
    if with_coo:
        m = 3 * N // 2
        r = np.sort(rng.integers(0, N, m)).astype(np.int32)
        c = rng.integers(0, N, m).astype(np.int32)
        coo = (r, c, rng.standard_normal(m))
    return values, offsets, block_start, x, coo


@pytest.mark.parametrize("N,bw,with_coo,dtype", [(1000, 64, False, np.float64), (1000, 64, True, np.float64),
                                                 (4099, 128, True, np.float64), (64, 64, False, np.float64),
                                                 (777, 256, True, np.float64), (1000, 64, True, np.float32),
                                                 (4099, 128, False, np.float32), (777, 256, True, np.float32)])
def test_runs3_hint_is_bit_identical(gpu, N, bw, with_coo, dtype):
    """The runs-of-three paths (the two-row waves included, FP64 and FP32) equal the hint-free launch
    and the oracle bit for bit."""
    print(f"case: runs3 N={N} bw={bw} coo={with_coo} {np.dtype(dtype).name}", flush=True)
    values, offsets, block_start, x, coo = _runs3_case(N, bw, 31 + N + bw, with_coo)
    values, x = values.astype(dtype), x.astype(dtype)
    if coo is not None:
        coo = (coo[0], coo[1], coo[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, coo)
    assert L.runs3 is not None
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    want = oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N)
    np.testing.assert_array_equal(outs[1], want)


def _shared_runs3_case(N, bw, seed, coo):
    """Runs-of-three layouts whose consecutive blocks mostly share one offset list (as a stencil's
    do), so that the multi-row waves take their shared-list path, with every way out of it mixed in:
    a block whose list differs in one offset, one with an extra run, an empty block, a list longer
    than 27 offsets, and offsets whose x windows cross either end of x.  coo: None, "light" (about
    an eighth of an entry per row) or "heavy" (one and a half per row)."""
    rng = np.random.default_rng(seed)
    nb = -(-N // bw)
    group = max(1, 256 // bw)

    def runs(k):
        centres = 3 * np.sort(rng.choice(np.arange(-N // 6, N // 6), size=k, replace=False)) + 1
        if rng.random() < 0.3:  # windows across the ends of x
            centres[0], centres[-1] = -N + 2, N - 2
            centres = np.unique(centres)
        return [int(c + d) for c in centres for d in (-1, 0, 1)]

    offs_per = []
    for g0 in range(0, nb, group):
        base = runs(int(rng.integers(1, 11)))  # 3 .. 30 offsets
        for j in range(min(group, nb - g0)):
            o = list(base)
            u = rng.random()
            if u < 0.06:
                o = o + [o[-1] + 2, o[-1] + 3, o[-1] + 4]  # an extra run
            elif u < 0.12:
                o = o[:-3] + [v + 3 for v in o[-3:]]  # the last run moved
            elif u < 0.16:
                o = []
            elif u < 0.2:
                o = runs(max(1, len(base) // 3))
            offs_per.append(o)
    block_start = np.zeros(nb + 1, np.int32)
    block_start[1:] = np.cumsum([len(o) for o in offs_per])
    offsets = np.array(sum(offs_per, []), np.int32)
    values = rng.standard_normal(len(offsets) * bw)
    values[rng.random(values.size) < 0.2] = 0.0
    x = rng.standard_normal(N)
    x[rng.random(N) < 0.02] = np.nan
    x[np.isnan(x) & (rng.random(N) < 0.5)] = 0.0
    cooc = None
    if coo is not None:
        m = N // 8 if coo == "light" else 3 * N // 2
        r = np.sort(rng.integers(0, N, m)).astype(np.int32)
        c = rng.integers(0, N, m).astype(np.int32)
        cooc = (r, c, rng.standard_normal(m))
    return values, offsets, block_start, x, cooc


@pytest.mark.parametrize("N,bw,coo,dtype", [(5000, 64, "light", np.float32), (5000, 64, "heavy", np.float32),
                                            (5000, 64, None, np.float32), (4099, 128, "light", np.float32),
                                            (3000, 256, None, np.float32), (1000, 512, "heavy", np.float32),
                                            (300, 64, "light", np.float32), (65, 64, None, np.float32),
                                            (5000, 64, "light", np.float64), (4099, 128, "heavy", np.float64)])
def test_shared_offset_lists_multi_row_waves(gpu, N, bw, coo, dtype):
    """Consecutive blocks with one offset list: the FP32 four-row and FP64 two-row waves (shared-list
    path, per-block path, one-pass and per-group fallback sums) equal the hint-free launch and the
    oracle bit for bit."""
    print(f"case: shared runs3 N={N} bw={bw} coo={coo} {np.dtype(dtype).name}", flush=True)
    values, offsets, block_start, x, cooc = _shared_runs3_case(N, bw, 7 * N + bw, coo)
    values, x = values.astype(dtype), x.astype(dtype)
    if cooc is not None:
        cooc = (cooc[0], cooc[1], cooc[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, cooc)
    assert L.runs3 is not None
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if cooc is not None:
        hb.update(coo_rows=cooc[0], coo_cols=cooc[1], coo_values=cooc[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    want = oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N)
    assert outs[0].tobytes() == want.tobytes()
    assert outs[1].tobytes() == want.tobytes()


@pytest.mark.parametrize("N,cover,bw,with_coo,dtype", [(1000, 700, 64, True, np.float64), (1000, 700, 64, False, np.float64),
                                                       (4096, 3000, 32, True, np.float64), (2000, 1100, 128, True, np.float64),
                                                       (1000, 700, 64, True, np.float32), (4096, 3000, 32, False, np.float32)])
def test_rows_past_the_last_block_keep_y(gpu, N, cover, bw, with_coo, dtype):
    """Blocks that cover only the first `cover` rows (as for a matrix with fewer columns than rows):
    on the runs-of-three paths (two-row waves included) the rows past the last block keep y, plus
    their fallback sums when there is a fallback; equal to the hint-free launch and the oracle."""
    print(f"case: rows past the last block N={N} cover={cover} bw={bw} coo={with_coo} {np.dtype(dtype).name}",
          flush=True)
    values, offsets, block_start, _, _ = _runs3_case(cover, bw, 17 + N + bw, False)
    rng = np.random.default_rng(N + cover)
    x = rng.standard_normal(N).astype(dtype)
    values = values.astype(dtype)
    coo = None
    if with_coo:
        r = np.sort(rng.integers(0, N, N)).astype(np.int32)
        coo = (r, rng.integers(0, N, N).astype(np.int32), rng.standard_normal(N).astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, coo)
    assert L.runs3 is not None
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, dtype))
    want = oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N)
    np.testing.assert_array_equal(outs[1], want)
    nb = len(block_start) - 1
    if coo is None:
        np.testing.assert_array_equal(outs[1][nb * bw:], y0[nb * bw:])


def _dense_fallback(N, seed):
    """Fallback entries, row-sorted, entry order within a row random: every third 64-row group is
    dense (0-25 entries per row, so far more than 64 per group, some rows empty), the others
    sparse (at most one entry per row, so that two groups often fit one 64-entry pass); duplicate
    entries and zero values included."""
    rng = np.random.default_rng(seed)
    rows = []
    for g in range(-(-N // 64)):
        r = np.arange(64 * g, min(64 * g + 64, N))
        per = rng.integers(0, 26, len(r)) if g % 3 == 0 else (rng.random(len(r)) < 0.15).astype(np.int64)
        rows.append(np.repeat(r, per))
    r = np.concatenate(rows).astype(np.int32)
    c = rng.integers(0, N, len(r)).astype(np.int32)
    c[1::17] = c[0::17][:len(c[1::17])]  # some duplicate columns
    v = rng.standard_normal(len(r))
    v[rng.random(len(r)) < 0.1] = 0.0
    return r, c, v


@pytest.mark.parametrize("N,bw,dtype", [(4099, 64, np.float64), (4096, 32, np.float64), (2000, 128, np.float64),
                                        (4099, 64, np.float32)])
def test_dense_fallback_groups_are_bit_identical(gpu, N, bw, dtype):
    """Groups with more than 64 fallback entries (each lane sums its own row's segment) next to
    sparse ones (two groups in one pass on the two-row kernels): results equal the hint-free launch
    and the oracle bit for bit."""
    print(f"case: dense fallback N={N} bw={bw} {np.dtype(dtype).name}", flush=True)
    values, offsets, block_start, x, _ = _runs3_case(N, bw, 5 + N + bw, False)
    coo = _dense_fallback(N, N + bw)
    counts = np.bincount(coo[0] // 64, minlength=-(-N // 64))
    assert counts.max() > 64 and np.any((counts[0::2][:len(counts[1::2])] + counts[1::2]) <= 64)
    values, x = values.astype(dtype), x.astype(dtype)
    coo = (coo[0], coo[1], coo[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, coo)
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw, "coo_rows": coo[0], "coo_cols": coo[1], "coo_values": coo[2]}
    np.testing.assert_array_equal(outs[1], oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N))


@pytest.mark.parametrize("N,bw,with_coo,k,dtype", [(1000, 64, True, 3, np.float64), (1000, 64, False, 8, np.float64),
                                                   (4099, 128, True, 5, np.float64), (777, 256, True, 16, np.float64),
                                                   (1000, 64, True, 8, np.float32), (777, 256, False, 5, np.float32),
                                                   (1000, 64, False, 12, np.float64), (4099, 128, False, 17, np.float64),
                                                   (777, 256, False, 24, np.float64)])
def test_runs3_hint_multi_rhs_is_bit_identical(gpu, N, bw, with_coo, k, dtype):
    """The multi-RHS kernel's runs-of-three path: every column equals the hint-free multi-RHS result,
    the single-vector kernel and the oracle."""
    print(f"case: runs3 multi-RHS N={N} bw={bw} coo={with_coo} k={k} {np.dtype(dtype).name}", flush=True)
    values, offsets, block_start, x0, coo = _runs3_case(N, bw, 7 + N + bw + k, with_coo)
    values = values.astype(dtype)
    if coo is not None:
        coo = (coo[0], coo[1], coo[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, coo)
    rng = np.random.default_rng(k)
    Xh = np.vstack([x0] + [rng.standard_normal(N) for _ in range(k - 1)]).astype(dtype)
    Y0 = np.vstack([np.linspace(-1.0, 1.0, N) + j for j in range(k)]).astype(dtype)
    X = DeviceArray.from_host(Xh.ravel())
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(Y0.ravel())
        L.spmm(X, Y, k, hints=hints)
        synchronize()
        outs.append(Y.download().reshape(k, N))
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    for j in range(k):
        np.testing.assert_array_equal(outs[1][j], oracle.hbdia_spmv(hb, Xh[j].copy(), Y0[j].copy(), N),
                                      err_msg=f"column {j}")


@pytest.mark.parametrize("k", [5, 8, 12, 32])
def test_multi_rhs_zero_and_nonfinite_sums_bit_exact(gpu, k):
    """The FP64 multi-RHS kernel sums without per-term skips and redoes a wave with them when a stored
    sum is zero, infinite or NaN.  Columns whose sums underflow to -0, are +0, pass through an
    infinity or NaN, or stay finite and non-zero all equal the oracle and the hint-free launch bit for
    bit, signs of zero included."""
    print(f"case: multi-RHS zero and non-finite sums k={k}", flush=True)
    n = 16
    A = HBDIA.stencil(n, 27, 0.0, seed=3)
    A.convertToHBDIAFormat(64, 2)
    hb = A.getHBDIA()
    N = n ** 3
    values = np.abs(hb["values"]) * 1e-200  # products with x of -1e-200 underflow to -0
    L = kernels.DeviceLayout(values, hb["offsets"], hb["block_start"], N, 64)
    assert L.runs3 is not None
    cols = [gen_vector(N, 5), np.full(N, -1e-200), np.full(N, -0.0), np.zeros(N), gen_vector(N, 6),
            gen_vector(N, 7), np.full(N, 1e-200), gen_vector(N, 8)]
    cols[4][N // 3] = np.nan
    cols[5][[7, N // 2]] = [np.inf, -np.inf]
    cols[7][: N // 2] = -1e-200
    Xh = np.vstack([cols[j % len(cols)] for j in range(k)])
    Y0 = np.full((k, N), 0.5)
    X = DeviceArray.from_host(Xh.ravel())
    outs = []
    for hints in (True, False):
        Y = DeviceArray.from_host(Y0.ravel())
        L.spmm(X, Y, k, hints=hints)
        synchronize()
        outs.append(Y.download().reshape(k, N))
    assert outs[0].tobytes() == outs[1].tobytes()
    hbo = {"values": values, "offsets": hb["offsets"], "block_start": hb["block_start"],
           "num_blocks": A.getNumBlocks(), "block_width": 64, "coo_rows": np.zeros(0, np.int32),
           "coo_cols": np.zeros(0, np.int32), "coo_values": np.zeros(0)}
    for j in range(k):
        want = oracle.hbdia_spmv(hbo, Xh[j].copy(), Y0[j].copy(), N)
        assert outs[0][j].tobytes() == want.tobytes(), f"column {j}"
    assert np.signbit(outs[0][1]).all() and (outs[0][1] == 0).all()  # the -0 sums survived


def _width32_case(N, seed, with_coo, share, group=2, kmax=5):
    """Block width 32: blocks after the first of each `group` (2: pairs 2k, 2k+1; 4: quads 4k..4k+3)
    that with probability `share` repeat the previous block's offset list (then one wave runs them
    together), otherwise their own; lists of up to kmax - 1 runs of three or single offsets (some out
    of range) and empty blocks; zero values over NaN x entries."""
    rng = np.random.default_rng(seed)
    bw = 32
    nb = -(-N // bw)
    offs_per = []
    for b in range(nb):
        if b % group != 0 and rng.random() < share:
            offs_per.append(list(offs_per[-1]))
            continue
        k = int(rng.integers(0, kmax))
        if rng.random() < 0.6:
            centres = 3 * np.sort(rng.choice(np.arange(-N // 2, N // 2), size=k, replace=False)) + 1
            offs_per.append([int(c + d) for c in centres for d in (-1, 0, 1)])
        else:
            offs_per.append(sorted(int(o) for o in rng.choice(np.arange(-N, N), size=k, replace=False)))
    block_start = np.zeros(nb + 1, np.int32)
    block_start[1:] = np.cumsum([len(o) for o in offs_per])
    offsets = np.array(sum(offs_per, []), np.int32)
    values = rng.standard_normal(len(offsets) * bw)
    values[rng.random(values.size) < 0.2] = 0.0
    x = rng.standard_normal(N)
    x[rng.random(N) < 0.02] = np.nan
    coo = None
    if with_coo:
        m = N
        r = np.sort(rng.integers(0, N, m)).astype(np.int32)
        coo = (r, rng.integers(0, N, m).astype(np.int32), rng.standard_normal(m))
    return values, offsets, block_start, x, coo


@pytest.mark.parametrize("N,with_coo,share,group,kmax,seed", [
    (1000, False, 0.8, 2, 5, None), (1000, True, 0.8, 2, 5, None), (4096, True, 1.0, 2, 5, None),
    (777, False, 0.5, 2, 5, None), (64, True, 1.0, 2, 5, None), (96, False, 1.0, 2, 5, None),
    # quads of blocks with one offset list: the two-row width-32 waves, their partial last wave and
    # their fallback to 64-row halves; kmax 13 reaches lists of more than 27 offsets; seed 0 gives
    # the small cases one quad of runs of three
    (1000, True, 0.9, 4, 5, None), (4096, False, 1.0, 4, 5, None), (4096, True, 1.0, 4, 13, None),
    (777, True, 0.7, 4, 13, None), (130, True, 1.0, 4, 5, 0), (224, False, 1.0, 4, 5, 0), (128, True, 1.0, 4, 5, 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_width32_shared_offsets_are_bit_identical(gpu, N, with_coo, share, group, kmax, seed, dtype):
    """Width-32 waves whose blocks share their offsets run the uniform sums (with the runs-of-three
    hint where it applies; four blocks with one list run two rows per lane): results equal the
    hint-free launch and the oracle."""
    print(f"case: width32 N={N} coo={with_coo} share={share} group={group} kmax={kmax}", flush=True)
    if seed is None:
        seed = 90 + N + (0 if group == 2 else 7 * group + kmax)
    values, offsets, block_start, x, coo = _width32_case(N, seed, with_coo, share, group, kmax)
    values, x = values.astype(dtype), x.astype(dtype)
    if coo is not None:
        coo = (coo[0], coo[1], coo[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, 32, coo)
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": 32}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    np.testing.assert_array_equal(outs[1], oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N))


def test_reference_configuration_at_width32(gpu):
    """The reference's own setup (create3DStencil27Point, block width 32, threshold 8) through
    hbdiaSpMV: the runs-of-three hint is built for width 32 and the result equals the oracle."""
    A = HBDIA()
    A.create3DStencil27Point(32, 32, 32, 0.0)
    A.convertToHBDIAFormat(32, 8)
    assert A.info()["runs3_blocks"] == A.getNumBlocks()  # each block is one grid line: whole runs
    N = 32 ** 3
    x = np.arange(N, dtype=np.float64)
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros(N))
    hbdiaSpMV(A, X, Y, False, True)
    A.synchronize()
    hb = A.getHBDIA()
    hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), 32
    np.testing.assert_array_equal(Y.download(), oracle.hbdia_spmv(hb, x.copy(), np.zeros(N), N))


def test_stencil27_gets_the_runs3_hint_and_stays_exact(gpu):
    A = HBDIA.stencil(20, 27, 0.01, seed=77)
    A.convertToHBDIAFormat(64, 2)
    # most blocks qualify; a block where two scatter entries reached the threshold on one offset does not
    assert A.info()["runs3_blocks"] >= 0.9 * A.getNumBlocks()
    B = HBDIA.stencil(20, 7, 0.0, seed=78)
    B.convertToHBDIAFormat(64, 2)
    assert B.info()["runs3_blocks"] == 0
    N = 20 ** 3
    x = gen_vector(N, 3)
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros(N))
    hbdiaSpMV(A, X, Y, False, True)
    A.synchronize()
    hb = A.getHBDIA()
    hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), 64
    np.testing.assert_array_equal(Y.download(), oracle.hbdia_spmv(hb, x, np.zeros(N), N))


@pytest.mark.parametrize("seed", [1, 2])
def test_runs3_hint_on_mixed_layouts(gpu, seed):
    # blocks whose runs are broken keep the per-offset path; the others take the shifted loads
    print(f"case: runs3 hint on mixed layout seed={seed}", flush=True)
    N, bw = 3000, 64
    values, offsets, block_start, x, coo = _runs3_case(N, bw, 500 + seed, True)
    rng = np.random.default_rng(seed)
    keep = np.ones(len(offsets), bool)
    keep[rng.choice(len(offsets), size=len(offsets) // 7, replace=False)] = False
    nb = len(block_start) - 1
    blk_of = np.repeat(np.arange(nb), np.diff(block_start))
    offsets2 = offsets[keep]
    block_start2 = np.zeros(nb + 1, np.int32)
    block_start2[1:] = np.cumsum(np.bincount(blk_of[keep], minlength=nb))
    values2 = values.reshape(-1, bw)[keep].ravel()
    L = _layout(values2, offsets2, block_start2, N, bw, coo)
    bits = kernels.runs3_bits(offsets2, block_start2, bw)
    assert bits is not None and 0 < int(sum(bin(int(w)).count("1") for w in bits)) < nb
    y0 = np.linspace(-1.0, 1.0, N)
    X = DeviceArray.from_host(x)
    outs = []
    for hints in (False, True):
        Y = DeviceArray.from_host(y0)
        L.spmv(X, Y, hints=hints)
        synchronize()
        outs.append(Y.download())
    np.testing.assert_array_equal(outs[1], outs[0])
    hb = {"values": values2, "offsets": offsets2, "block_start": block_start2, "num_blocks": nb, "block_width": bw,
          "coo_rows": coo[0], "coo_cols": coo[1], "coo_values": coo[2]}
    np.testing.assert_array_equal(outs[1], oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N))


@pytest.mark.parametrize("dtype,tiny", [(np.float64, 1e-200), (np.float32, 1e-30)])
@pytest.mark.parametrize("offsets", [(0, 1), (-1, 0, 1), (-1, 0, 1, 5)])
def test_negative_zero_sum_survives_skipped_entries(gpu, dtype, tiny, offsets):
    """A row sum that underflows to -0 keeps its sign across later zero-valued entries, which the
    reference skips (src/Operations/HBDIASpMV.cu:50-53): the first offset's product -tiny * tiny rounds
    to -0, the other offsets hold explicit zeros.  Covers the plain, runs-of-three and two-row paths."""
    print(f"case: -0 across skipped entries {np.dtype(dtype).name} offsets {offsets}", flush=True)
    n = 512
    rows, cols, vals = [], [], []
    for r in range(n):
        for k, o in enumerate(offsets):
            if 0 <= r + o < n:
                rows.append(r)
                cols.append(r + o)
                vals.append(-tiny if k == 0 else 0.0)
    A = HBDIA(np.array(rows), np.array(cols), np.array(vals, dtype=dtype), n, n)
    A.convertToHBDIAFormat(64, 1)
    assert A.info()["num_coo"] == 0
    x = np.full(n, tiny, dtype)
    want = oracle_y(A, x, np.zeros(n, dtype))
    assert np.signbit(want).sum() > n // 2  # the case is exercised: most rows end at -0
    y = run_spmv(A, x, np.zeros(n, dtype), False, True)
    assert y.tobytes() == want.tobytes()


@pytest.mark.parametrize("symmetric", [False, True])
def test_matrix_market_file_through_the_gpu_path(gpu, tmp_path, symmetric):
    """loadMTX -> convertToHBDIAFormat -> hbdiaSpMV on a Matrix Market file written here (the reference's
    route for SuiteSparse inputs, run_HBDIA_test.cpp:19-83): bit-exact against the oracle on the loaded
    COO, and within the stated tolerance of a long-double SpMV of the file's entries."""
    print(f"case: Matrix Market 14^3 stencil + noise, symmetric={symmetric}", flush=True)
    r, c, v, N = oracle.stencil27(14, 14, 14, 0.05, 2)
    lines = []
    if symmetric:  # the lower triangle, pattern only (value 1 each), mirrored by the loader
        keep = r >= c
        lines = [f"{a + 1} {b + 1}" for a, b in zip(r[keep], c[keep])]
        header = "%%MatrixMarket matrix coordinate pattern symmetric\n"
    else:
        lines = [f"{a + 1} {b + 1} {val!r}" for a, b, val in zip(r, c, v.tolist())]
        header = "%%MatrixMarket matrix coordinate real general\n"
    f = tmp_path / "m.mtx"
    f.write_text(header + f"{N} {N} {len(lines)}\n" + "\n".join(lines) + "\n")
    A = HBDIA()
    assert A.loadMTX(str(f))
    R, C, V = A.getCOO()
    A.convertToHBDIAFormat(64, 2)
    x = np.random.default_rng(9).uniform(-1, 1, N)
    y = run_spmv(A, x, np.zeros(N), False, True)
    np.testing.assert_array_equal(y, oracle_y(A, x, np.zeros(N)))
    ref = oracle.coo_spmv_longdouble(R, C, V, x, N)
    scale = oracle.coo_spmv_longdouble(R, C, np.abs(V), np.abs(x), N)
    assert np.max(np.abs(y - ref) / np.maximum(scale, 1e-300)) < 1e-12
    if symmetric:
        assert np.all(V == 1.0) and len(V) > len(lines)


def test_rectangular_matrix_needs_x_over_every_column(gpu):
    """A 500 x 900 matrix whose fallback reaches columns past the rows: an x of
    500 entries is refused before any launch (the fallback gathers x[col] for
    col < 900); with 900 entries the result is the oracle's."""
    print("case: rectangular 500x900, x length check", flush=True)
    nr, nc = 500, 900
    rows, cols = [], []
    for i in range(nr):
        for c in sorted({max(i - 1, 0), i, min(i + 1, nr - 1), nr + (7 * i) % (nc - nr)}):
            rows.append(i)
            cols.append(c)
    vals = np.random.default_rng(12).uniform(-1, 1, len(rows))
    A = HBDIA(np.array(rows), np.array(cols), vals, nr, nc)
    A.convertToHBDIAFormat(64, 2)
    assert A.info()["num_coo"] >= nr
    with pytest.raises(ValueError, match="columns"):
        hbdiaSpMV(A, HBDIAVector(np.ones(nr)), HBDIAVector(np.zeros(nr)), False, True)
    x = np.random.default_rng(13).uniform(-1, 1, nc)
    y = run_spmv(A, x, np.zeros(nr), False, True)
    assert y.tobytes() == oracle_y(A, x, np.zeros(nr)).tobytes()
    X = DeviceArray.from_host(np.ones(nr))
    Y = DeviceArray(nr, np.float64, zero=True)
    with pytest.raises(ValueError):
        from hbdia_amd import hbdiaSpMVMulti
        hbdiaSpMVMulti(A, X, Y, 1)


@pytest.mark.parametrize("N,bw,with_coo,dtype", [(1000, 64, False, np.float64), (1000, 64, True, np.float64),
                                                 (4099, 128, True, np.float64), (1000, 64, True, np.float32),
                                                 (4096, 32, True, np.float64), (4096, 32, False, np.float32),
                                                 (1003, 5, True, np.float64), (64, 64, True, np.float64)])
def test_exactly_sized_offsets_array(gpu, N, bw, with_coo, dtype):
    """The single-vector entry points of the reference's launch_bdia_spmv_kernel signature (no
    layout hints) read nothing past block_start[num_blocks]: with the offsets uploaded at their exact
    size -- no padding -- the uniform, width-32 and per-lane paths give the oracle's y bit for bit."""
    hints = False
    print(f"case: exact offsets N={N} bw={bw} coo={with_coo} {np.dtype(dtype).name} hints={hints}", flush=True)
    if bw == 32:
        values, offsets, block_start, x, coo = _width32_case(N, 77, with_coo, share=1.0, group=4)
    else:
        values, offsets, block_start, x, coo = _runs3_case(N, bw, 97 + N + bw, with_coo)
    values, x = values.astype(dtype), x.astype(dtype)
    if coo is not None:
        coo = (coo[0], coo[1], coo[2].astype(dtype))
    args = (np.asarray(values), np.asarray(offsets, np.int32), np.asarray(block_start, np.int32), N, bw)
    L = kernels.DeviceLayout(*args, *(coo or ()), pad_offsets=False)
    assert L.offsets.size == max(len(offsets), 1)
    y0 = np.linspace(-1.0, 1.0, N).astype(dtype)
    X = DeviceArray.from_host(x)
    Y = DeviceArray.from_host(y0)
    L.spmv(X, Y, hints=hints)
    synchronize()
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    want = oracle.hbdia_spmv(hb, x.copy(), y0.copy(), N)
    assert Y.download().tobytes() == want.tobytes()


def _mixed_runs_case(N, bw, seed, with_coo):
    """Like _runs3_case, but each block's sorted offsets are groups of one, two or three consecutive
    values (groups may touch, making longer runs), some outside [-N, N)."""
    rng = np.random.default_rng(seed)
    nb = -(-N // bw)
    offs_per = []
    for _ in range(nb):
        k = int(rng.integers(0, 8))
        starts = np.sort(rng.choice(np.arange(-N - 5, N + 5, 4), size=k, replace=False))
        offs = sorted({int(s0 + d) for s0 in starts for d in range(int(rng.integers(1, 4)))})
        offs_per.append(offs)
    block_start = np.zeros(nb + 1, np.int32)
    block_start[1:] = np.cumsum([len(o) for o in offs_per])
    offsets = np.array(sum(offs_per, []), np.int32)
    values = rng.standard_normal(len(offsets) * bw)
    values[rng.random(values.size) < 0.2] = 0.0
    x = rng.standard_normal(N)
    x[rng.random(N) < 0.02] = np.nan
    coo = None
    if with_coo:
        m = 3 * N // 2
        r = np.sort(rng.integers(0, N, m)).astype(np.int32)
        coo = (r, rng.integers(0, N, m).astype(np.int32), rng.standard_normal(m))
    return values, offsets, block_start, x, coo


@pytest.mark.parametrize("N,bw,with_coo,k,dtype", [(1000, 64, False, 2, np.float64), (1000, 64, True, 4, np.float64),
                                                   (4099, 128, True, 8, np.float64), (777, 256, True, 32, np.float64),
                                                   (1000, 64, True, 64, np.float64), (1000, 64, True, 6, np.float64),
                                                   (1000, 64, True, 4, np.float32), (4099, 128, False, 32, np.float32),
                                                   (777, 64, True, 128, np.float32), (64, 64, True, 12, np.float32)])
def test_row_major_multi_rhs_is_bit_identical(gpu, N, bw, with_coo, k, dtype):
    """The row-major (interleaved) multi-RHS kernel on random layouts -- runs of three and lone
    offsets, columns outside [0, N), zero values over NaN x entries, a fallback in some cases,
    FP64 and FP32, k up to 512 bytes per row: every column equals the oracle's single-vector result."""
    print(f"case: row-major multi-RHS N={N} bw={bw} coo={with_coo} k={k} {np.dtype(dtype).name}", flush=True)
    values, offsets, block_start, x0, coo = _mixed_runs_case(N, bw, 11 + N + bw + k, with_coo)
    rng = np.random.default_rng(N + k)
    values = values.astype(dtype)
    if coo is not None:
        coo = (coo[0], coo[1], coo[2].astype(dtype))
    L = _layout(values, offsets, block_start, N, bw, coo)
    Xh = np.vstack([x0] + [rng.standard_normal(N) for _ in range(k - 1)]).astype(dtype)  # (k, N)
    Y0 = np.vstack([np.linspace(-1.0, 1.0, N) + j for j in range(k)]).astype(dtype)
    X = DeviceArray.from_host(np.ascontiguousarray(Xh.T).ravel())
    Y = DeviceArray.from_host(np.ascontiguousarray(Y0.T).ravel())
    L.spmm_rm(X, Y, k)
    synchronize()
    out = Y.download().reshape(N, k).T
    hb = {"values": values, "offsets": offsets, "block_start": block_start, "num_blocks": len(block_start) - 1,
          "block_width": bw}
    if coo is not None:
        hb.update(coo_rows=coo[0], coo_cols=coo[1], coo_values=coo[2])
    else:
        hb.update(coo_rows=np.zeros(0, np.int32), coo_cols=np.zeros(0, np.int32), coo_values=np.zeros(0, values.dtype))
    for j in range(k):
        want = oracle.hbdia_spmv(hb, Xh[j].copy(), Y0[j].copy(), N)
        assert out[j].tobytes() == want.tobytes(), f"column {j}"


@pytest.mark.parametrize("k,dtype", [(2, np.float64), (8, np.float64), (32, np.float64), (4, np.float32),
                                     (32, np.float32)])
def test_row_major_multi_rhs_stencil_and_api(gpu, k, dtype):
    """hbdiaSpMVMulti(layout="row") on a 20^3 stencil with 1 % scatter (block width 64) and on a
    matrix whose blocks cover only part of the rows: every column equals the oracle; rows without a
    block keep Y; unsupported widths are refused."""
    print(f"case: row-major multi-RHS stencil k={k} {np.dtype(dtype).name}", flush=True)
    from hbdia_amd import hbdiaSpMVMulti
    n = 20
    A = HBDIA.stencil(n, 27, 0.01, seed=21, dtype=dtype)
    A.convertToHBDIAFormat(64, 2)
    N = n ** 3
    rng = np.random.default_rng(k)
    Xh = rng.uniform(-1, 1, (N, k)).astype(dtype)
    Y0 = np.full((N, k), 0.25, dtype)
    X, Y = DeviceArray.from_host(Xh.ravel()), DeviceArray.from_host(Y0.ravel())
    hbdiaSpMVMulti(A, X, Y, k, layout="row")
    A.synchronize()
    out = Y.download().reshape(N, k)
    for j in range(k):
        assert out[:, j].tobytes() == oracle_y(A, np.ascontiguousarray(Xh[:, j]), Y0[:, j].copy()).tobytes()
    # blocks over the first 100 of 200 rows only, no fallback: rows past them keep Y
    rows = np.array([0, 0, 1, 63, 64, 64], np.int32)
    cols = np.array([0, 1, 1, 63, 64, 63], np.int32)
    B = HBDIA(rows, cols, np.array([2.0, 1.0, 3.0, 4.0, 5.0, 6.0], dtype), 200, 100)
    B.convertToHBDIAFormat(64, 1)
    Xb = rng.uniform(-1, 1, (200, k)).astype(dtype)
    Yb0 = np.full((200, k), 7.0, dtype)
    XB, YB = DeviceArray.from_host(Xb.ravel()), DeviceArray.from_host(Yb0.ravel())
    hbdiaSpMVMulti(B, XB, YB, k, layout="row")
    B.synchronize()
    ob = YB.download().reshape(200, k)
    for j in range(k):
        assert ob[:, j].tobytes() == oracle_y(B, np.ascontiguousarray(Xb[:, j]), Yb0[:, j].copy()).tobytes()
    assert np.all(ob[128:] == 7.0)
    bad = 3 if dtype == np.float64 else 6  # a row of 24 bytes: not a multiple of 16
    with pytest.raises((HBDIAError, ValueError)):
        hbdiaSpMVMulti(A, DeviceArray(N * bad, dtype), DeviceArray(N * bad, dtype), bad, layout="row")
