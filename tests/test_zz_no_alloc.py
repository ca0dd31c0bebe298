"""The allocation-free reduced-precision entry point, through the C ABI.

hbdia_hbdia_spmv_mixed_noalloc keeps the per-group FP64 masks in a
library-owned device array instead of a workspace argument.  Its y must equal
hbdia_hbdia_spmv_mixed_ex's bit for bit (same kernels, same arithmetic) on two
calls in a row, the second on a smaller matrix, so a stale mask from the first
call would show; sizes past its maximum are refused before any launch.
"""
from __future__ import annotations

import numpy as np
import pytest

from hbdia_amd import HBDIA, DeviceArray, Stream, gen_vector, kernels
from hbdia_amd._native import check_hip, hip

pytestmark = pytest.mark.gpu

VDTYPE = {"fp32": 1, "bf16": 2}
NOALLOC_MAX_ROWS = 1 << 24


class Case:
    """A converted matrix with its low-precision value copies, runs-of-three bits, x and a workspace."""

    def __init__(self, n: int, precision: str, seed: int, stream: Stream):
        A = HBDIA.stencil(n, 27, 0.005, seed=seed)
        A.convertToHBDIAFormat(64, 2)
        self.A, self.N, self.vd = A, n ** 3, VDTYPE[precision]
        self.p = A.device_ptrs()
        info = A.info()
        et = np.uint32 if precision == "fp32" else np.uint16
        self.vlp = DeviceArray(max(info["num_values"], 1), et)
        self.clp = DeviceArray(max(info["num_coo"], 1), et)
        check_hip(hip.hbdia_convert_values(self.p["values"], self.vlp.ptr, info["num_values"], self.vd, stream.handle))
        check_hip(hip.hbdia_convert_values(self.p["coo_values"], self.clp.ptr, info["num_coo"], self.vd, stream.handle))
        hb = A.getHBDIA()
        bits = kernels.runs3_bits(hb["offsets"], hb["block_start"], 64)
        assert bits is not None
        self.runs3 = DeviceArray.from_host(bits)
        self.X = DeviceArray.from_host(gen_vector(self.N, seed + 1))
        self.ws = DeviceArray(int(hip.hbdia_mixed_workspace_bytes(self.N)) // 8, np.uint64, zero=True)
        stream.synchronize()

    def args(self, Y: DeviceArray):
        p = self.p
        return (self.vd, self.vlp.ptr, p["values"], p["offsets"], p["block_start"], self.A.getNumBlocks(), 64,
                p["coo_rows"], p["coo_cols"], self.clp.ptr, p["coo_values"], p["coo_group_ptr"], self.X.ptr, Y.ptr,
                self.N, 1e4)

    def with_workspace(self, stream: Stream, hints: bool) -> np.ndarray:
        Y = DeviceArray(self.N, np.float64, zero=True)
        check_hip(hip.hbdia_hbdia_spmv_mixed_ex(*self.args(Y), self.ws.ptr,
                                                self.runs3.ptr if hints else None, stream.handle))
        stream.synchronize()
        return Y.download()

    def no_alloc(self, stream: Stream, hints: bool) -> np.ndarray:
        Y = DeviceArray(self.N, np.float64, zero=True)
        check_hip(hip.hbdia_hbdia_spmv_mixed_noalloc(*self.args(Y), self.runs3.ptr if hints else None,
                                                     stream.handle))
        stream.synchronize()
        return Y.download()


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("hints", [True, False])
def test_no_alloc_equals_workspace_entry_point(gpu, precision, hints):
    print(f"case: no-alloc mixed {precision} hints={hints}: 40^3 then 24^3", flush=True)
    s = Stream()
    big, small = Case(40, precision, 0xA1, s), Case(24, precision, 0xA2, s)
    for c in (big, small):  # the second call has fewer rows than the first
        want = c.with_workspace(s, hints)
        got = c.no_alloc(s, hints)
        assert got.tobytes() == want.tobytes(), f"{c.N} rows: the allocation-free result differs"
    # and once more on the large matrix after the small one
    assert big.no_alloc(s, hints).tobytes() == big.with_workspace(s, hints).tobytes()


def test_no_alloc_refuses_rows_past_its_maximum(gpu):
    print("case: no-alloc size limit", flush=True)
    s = Stream()
    c = Case(8, "bf16", 0xA3, s)
    Y = DeviceArray(c.N, np.float64, zero=True)
    args = list(c.args(Y))
    args[14] = NOALLOC_MAX_ROWS + 1  # num_rows: refused on the host before any launch
    rc = hip.hbdia_hbdia_spmv_mixed_noalloc(*args, None, s.handle)
    assert rc == 1  # HBDIA_ERROR_INVALID_ARGUMENT
    assert b"HBDIA_NOALLOC_MAX_ROWS" in hip.hbdia_last_error()
