"""ctypes bindings of the two native libraries of hbdia_amd.

libhbdia_hip.so exports the device C ABI of include/hbdia_hip.h; libhbdia.so
exports the C handle API of include/hbdia_capi.h over the C++ classes.  Both
are built in-tree by hbdia_amd.build.  There is no fallback: a missing library
raises NativeLibraryMissing on import.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"


class NativeLibraryMissing(RuntimeError):
    """A native library of hbdia_amd has not been built."""


def _load(name: str) -> ctypes.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python hbdia_amd/build.py` (or __graft_entry__.build())")
    return ctypes.CDLL(str(path), mode=os.RTLD_NOW | os.RTLD_GLOBAL)


hip = _load("libhbdia_hip.so")
capi = _load("libhbdia.so")

_i = c_int
_p = c_void_p
_pi32 = POINTER(c_int32)
_pi64 = POINTER(c_int64)

_HIP_SIGS = {
    "hbdia_last_error": (c_char_p, []),
    "hbdia_get_device_count": (_i, [POINTER(c_int)]),
    "hbdia_set_device": (_i, [c_int]),
    "hbdia_get_device": (_i, [POINTER(c_int)]),
    "hbdia_device_synchronize": (_i, []),
    "hbdia_mem_info": (_i, [POINTER(c_size_t), POINTER(c_size_t)]),
    "hbdia_device_name": (_i, [ctypes.c_char_p, c_size_t]),
    "hbdia_device_compute_units": (_i, [POINTER(c_int)]),
    "hbdia_malloc": (_i, [POINTER(c_void_p), c_size_t]),
    "hbdia_free": (_i, [_p]),
    "hbdia_malloc_host": (_i, [POINTER(c_void_p), c_size_t]),
    "hbdia_free_host": (_i, [_p]),
    "hbdia_malloc_managed": (_i, [POINTER(c_void_p), c_size_t]),
    "hbdia_memcpy_h2d": (_i, [_p, _p, c_size_t, _p]),
    "hbdia_memcpy_d2h": (_i, [_p, _p, c_size_t, _p]),
    "hbdia_memcpy_d2d": (_i, [_p, _p, c_size_t, _p]),
    "hbdia_memcpy": (_i, [_p, _p, c_size_t]),
    "hbdia_memset": (_i, [_p, c_int, c_size_t, _p]),
    "hbdia_stream_create": (_i, [POINTER(c_void_p)]),
    "hbdia_stream_destroy": (_i, [_p]),
    "hbdia_stream_synchronize": (_i, [_p]),
    "hbdia_event_create": (_i, [POINTER(c_void_p)]),
    "hbdia_event_query": (_i, [_p, POINTER(c_int)]),
    "hbdia_event_destroy": (_i, [_p]),
    "hbdia_event_record": (_i, [_p, _p]),
    "hbdia_event_synchronize": (_i, [_p]),
    "hbdia_stream_wait_event": (_i, [_p, _p]),
    "hbdia_event_elapsed_ms": (_i, [POINTER(c_float), _p, _p]),
    "hbdia_graph_begin_capture": (_i, [_p]),
    "hbdia_graph_end_capture": (_i, [_p, POINTER(c_void_p)]),
    "hbdia_graph_launch": (_i, [_p, _p]),
    "hbdia_graph_destroy": (_i, [_p]),
    "hbdia_bdia_spmv_f64": (_i, [_p, _p, _p, _p, _p, c_int64, c_int32, c_int32, _p]),
    "hbdia_bdia_spmv_f32": (_i, [_p, _p, _p, _p, _p, c_int64, c_int32, c_int32, _p]),
    "hbdia_hbdia_spmv_f64": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int64, _p]),
    "hbdia_hbdia_spmv_f32": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int64, _p]),
    "hbdia_bdia_spmv_ex_f64": (_i, [_p, _p, _p, _p, _p, c_int64, c_int32, c_int32, _p, _p]),
    "hbdia_bdia_spmv_ex_f32": (_i, [_p, _p, _p, _p, _p, c_int64, c_int32, c_int32, _p, _p]),
    "hbdia_hbdia_spmv_ex_f64": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int64, _p, _p]),
    "hbdia_hbdia_spmv_ex_f32": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int64, _p, _p]),
    "hbdia_hbdia_spmm_f64": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, c_int64, _p, c_int64, c_int32,
                                  c_int64, _p]),
    "hbdia_hbdia_spmm_f32": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, c_int64, _p, c_int64, c_int32,
                                  c_int64, _p]),
    "hbdia_convert_values": (_i, [_p, _p, c_int64, c_int, _p]),
    "hbdia_mixed_workspace_bytes": (c_size_t, [c_int64]),
    "hbdia_batch_make_tiles": (_i, [_p, c_int32, _p, _p, _p, c_int32, _p, c_int64]),
    "hbdia_batch_spmv_f64": (_i, [_p, _p, _p, _p, _p]),
    "hbdia_serve_spmv_f64": (_i, [_p, _p, c_int32, c_int64, c_int32, c_int32, _p]),
    "hbdia_serve_spmv_f32": (_i, [_p, _p, c_int32, c_int64, c_int32, c_int32, _p]),
    "hbdia_batch_spmv_f32": (_i, [_p, _p, _p, _p, _p]),
    "hbdia_batch_occupancy": (_i, [_p, c_int32, c_int32, POINTER(c_int)]),
    "hbdia_hbdia_spmm_rm_f64": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int32, c_int64, _p]),
    "hbdia_hbdia_spmm_rm_f32": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, c_int32, c_int64, _p]),
    "hbdia_hbdia_spmm_ex_f64": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, c_int64, _p, c_int64, c_int32,
                                     c_int64, _p, _p]),
    "hbdia_hbdia_spmm_ex_f32": (_i, [_p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, c_int64, _p, c_int64, c_int32,
                                     c_int64, _p, _p]),
    "hbdia_hbdia_spmv_mixed_ex": (_i, [c_int, _p, _p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, _p, c_int64,
                                       c_double, _p, _p, _p]),
    "hbdia_hbdia_spmv_mixed_noalloc": (_i, [c_int, _p, _p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, _p,
                                            c_int64, c_double, _p, _p]),
    "hbdia_hbdia_spmv_mixed": (_i, [c_int, _p, _p, _p, _p, c_int32, c_int32, _p, _p, _p, _p, _p, _p, _p, c_int64,
                                    c_double, _p, _p]),
    "hbdia_coo_spmv_f64": (_i, [_p, _p, _p, _p, _p, _p, c_int64, _p]),
    "hbdia_coo_spmv_f32": (_i, [_p, _p, _p, _p, _p, _p, c_int64, _p]),
    "hbdia_coo_group_ptr": (_i, [_p, c_int64, c_int64, _p, _p]),
    "hbdia_vector_add_f64": (_i, [_p, _p, c_int64, _p]),
    "hbdia_vector_add_f32": (_i, [_p, _p, c_int64, _p]),
    "hbdia_convert_coo": (_i, [_p, _p, _p]),
    "hbdia_convert_release": (_i, [_p]),
    "hbdia_convert_trim": (_i, []),
}

_CAPI_SIGS = {
    "hbdia_capi_last_error": (c_char_p, []),
    "hbdia_set_num_threads": (_i, [c_int]),
    "hbdia_get_num_threads": (_i, []),
    "hbdia_matrix_create": (_i, [POINTER(c_void_p), c_int]),
    "hbdia_matrix_create_coo": (_i, [POINTER(c_void_p), c_int, _p, _p, _p, c_int64, c_int32, c_int32]),
    "hbdia_matrix_destroy": (_i, [_p]),
    "hbdia_matrix_stencil27": (_i, [_p, c_int, c_int, c_int, c_double, c_int]),
    "hbdia_matrix_convert": (_i, [_p, c_int, c_int, c_int, c_int, c_int]),
    "hbdia_matrix_convert_host": (_i, [_p, c_int, c_int, c_int, c_int, c_int]),
    "hbdia_matrix_convert_device": (_i, [_p, c_int, c_int, c_int, c_int, c_int]),
    "hbdia_matrix_convert_device_ex": (_i, [_p, c_int, c_int, c_int, c_int, c_int, c_int]),
    "hbdia_matrix_remove_duplicates": (_i, [_p]),
    "hbdia_matrix_delete_hbdia": (_i, [_p]),
    "hbdia_matrix_prepare_gpu": (_i, [_p]),
    "hbdia_matrix_synchronize": (_i, [_p]),
    "hbdia_matrix_info": (_i, [_p, _p]),
    "hbdia_matrix_storage_report": (_i, [_p, _p, _p, _p, c_int64]),
    "hbdia_matrix_storage_report_csv": (_i, [_p, c_char_p, _p, c_size_t, _p]),
    "hbdia_matrix_get_coo": (_i, [_p, _p, _p, _p]),
    "hbdia_matrix_get_hbdia": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "hbdia_matrix_device_ptrs": (_i, [_p, _p]),
    "hbdia_vector_create": (_i, [POINTER(c_void_p), c_int, _p, c_int64, c_int, c_int, c_int, c_int]),
    "hbdia_vector_destroy": (_i, [_p]),
    "hbdia_vector_upload": (_i, [_p, _p, c_int64]),
    "hbdia_vector_download": (_i, [_p, _p, c_int64]),
    "hbdia_vector_download_coo": (_i, [_p, _p, c_int64]),
    "hbdia_vector_device_ptrs": (_i, [_p, POINTER(c_void_p), POINTER(c_void_p)]),
    "hbdia_spmv": (_i, [_p, _p, _p, c_int, c_int]),
    "hbdia_spmv_time_synced": (_i, [_p, _p, _p, c_int, c_int, c_int, POINTER(c_double)]),
    "hbdia_matrix_rank_local": (_i, [POINTER(c_void_p), _p, c_int, c_int, c_int, c_int, c_int, c_int]),
    "hbdia_matrix_build_local_coo": (_i, [_p]),
    "hbdia_matrix_local_info": (_i, [_p, _p]),
    "hbdia_matrix_get_local": (_i, [_p, _p, _p, _p, _p, _p]),
    "hbdia_matrix_get_vector_offsets": (_i, [_p, _p]),
    "hbdia_vector_create_halo": (_i, [POINTER(c_void_p), c_int, _p, c_int64, _p]),
    "hbdia_vector_fill_halo": (_i, [_p, _p, _p, c_int64]),
    "hbdia_vector_verify_halo": (_i, [_p, _p, POINTER(c_int)]),
    "hbdia_spmv_multi": (_i, [_p, _p, c_int64, _p, c_int64, c_int, c_int]),
    "hbdia_spmv_multi_rm": (_i, [_p, _p, _p, c_int, c_int]),
    "hbdia_mixed_create": (_i, [POINTER(c_void_p), _p, c_int, c_double]),
    "hbdia_mixed_destroy": (_i, [_p]),
    "hbdia_mixed_spmv": (_i, [_p, _p, _p]),
    "hbdia_mixed_spmv_raw": (_i, [_p, _p, _p]),
    "hbdia_mixed_fp64_rows": (_i, [_p, POINTER(c_int64)]),
    "hbdia_mixed_error_bound": (_i, [_p, POINTER(c_double), POINTER(c_double)]),
    "hbdia_mixed_fp64_masks": (_i, [_p, _p, c_int64]),
    "hbdia_gen_stencil": (_i, [POINTER(c_void_p), c_int, c_int, c_int, c_double, c_uint64]),
    "hbdia_gen_banded": (_i, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_double, c_uint64]),
    "hbdia_gen_vector": (_i, [c_int, c_int64, c_uint64, c_double, c_double, _p]),
    "hbdia_matrix_load_mtx": (_i, [_p, ctypes.c_char_p, c_int, c_int, POINTER(c_int)]),
    "hbdia_server_create": (_i, [POINTER(c_void_p), POINTER(c_void_p), c_int, c_int64, c_int, c_int, c_int, c_int,
                                 c_int]),
    "hbdia_server_destroy": (_i, [_p]),
    "hbdia_server_submit": (_i, [_p, c_int, _p, _p, c_int, POINTER(c_int64)]),
    "hbdia_server_wait": (_i, [_p, c_int64]),
    "hbdia_server_drain": (_i, [_p]),
    "hbdia_server_times": (_i, [_p, c_int64, POINTER(c_int64)]),
    "hbdia_server_info": (_i, [_p, POINTER(c_int64)]),
    "hbdia_server_num_rows": (_i, [_p, c_int, POINTER(c_int64)]),
    "hbdia_server_open_loop": (_i, [_p, c_double, c_int64, c_uint64, POINTER(c_void_p), c_int, POINTER(c_void_p),
                                    c_int, POINTER(c_double), POINTER(c_double)]),
    "hbdia_arena_create": (_i, [POINTER(c_void_p), c_int64, c_int64]),
    "hbdia_arena_destroy": (_i, [_p]),
    "hbdia_arena_allocate": (_i, [_p, c_int64, POINTER(c_int64)]),
    "hbdia_arena_release": (_i, [_p, c_int64]),
    "hbdia_arena_info": (_i, [_p, POINTER(c_int64)]),
    "hbdia_serve_arena_elems": (_i, [c_int64, c_int64, c_int64, POINTER(c_int64)]),
    "hbdia_gen_banded_batch": (_i, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_double,
                                    c_uint64]),
    "hbdia_banded_batch_shape": (_i, [c_int, c_int, c_int, c_int, c_int, c_uint64, POINTER(c_int32),
                                      POINTER(c_int32), POINTER(c_uint64)]),
    "hbdia_batch_create": (_i, [POINTER(c_void_p), POINTER(c_void_p), c_int]),
    "hbdia_batch_autotune": (_i, [POINTER(c_void_p), POINTER(c_void_p), c_int, POINTER(c_int), c_int, c_int, c_int,
                                  POINTER(c_double), c_int, POINTER(c_int)]),
    "hbdia_batch_destroy": (_i, [_p]),
    "hbdia_batch_info": (_i, [_p, POINTER(c_int64)]),
    "hbdia_batch_row_offsets": (_i, [_p, POINTER(c_int64)]),
    "hbdia_batch_set_schedule": (_i, [_p, c_int, c_int, c_int]),
    "hbdia_batch_resident_workgroups": (_i, [_p, c_int, POINTER(c_int)]),
    "hbdia_batch_spmv": (_i, [_p, _p, _p]),
    "hbdia_batch_spmv_per_matrix": (_i, [_p, _p, _p]),
    "hbdia_batch_capture_graph": (_i, [_p, _p, _p]),
    "hbdia_batch_launch_graph": (_i, [_p]),
    "hbdia_batch_synchronize": (_i, [_p]),
    "hbdia_batch_stream": (_i, [_p, POINTER(c_void_p)]),
}


def _declare(lib: ctypes.CDLL, sigs: dict) -> None:
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_declare(hip, _HIP_SIGS)
_declare(capi, _CAPI_SIGS)

HIP_FUNCTIONS = tuple(_HIP_SIGS)
CAPI_FUNCTIONS = tuple(_CAPI_SIGS)


class HBDIAError(RuntimeError):
    """A native call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


def check_hip(status: int) -> None:
    if status != 0:
        raise HBDIAError(status, hip.hbdia_last_error().decode(errors="replace"))


def check_capi(status: int) -> None:
    if status != 0:
        raise HBDIAError(status, capi.hbdia_capi_last_error().decode(errors="replace"))


def gpu_available() -> bool:
    """True when a HIP device is visible (no kernel is launched)."""
    n = c_int(0)
    if hip.hbdia_get_device_count(ctypes.byref(n)) != 0:
        return False
    return n.value > 0
