"""ctypes declarations of the C ABI in include/set_intersection.h (libsi.so).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, c_char, c_char_p, c_double, c_float, c_int, c_int32, c_uint32,
                    c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsi.so")
GEN_PATH = os.path.join(HERE, "libsigen.so")

# status codes
SI_OK = 0
STATUS_NAMES = {0: "SI_OK", -1: "SI_ERR_ARG", -2: "SI_ERR_HIP", -3: "SI_ERR_NOMEM", -4: "SI_ERR_RANGE",
                -5: "SI_ERR_UNSORTED", -6: "SI_ERR_CAPACITY", -7: "SI_ERR_DEVICE"}

# techniques
TECH_IP, TECH_OBS, TECH_HI, TECH_BITMAP_NAIVE, TECH_BITMAP_DYNAMIC, TECH_MM = range(6)
TECHNIQUES = {"ip": TECH_IP, "obs": TECH_OBS, "hi": TECH_HI, "bitmap_naive": TECH_BITMAP_NAIVE,
              "bitmap_dynamic": TECH_BITMAP_DYNAMIC, "mm": TECH_MM}
MM_I8, MM_F32, MM_POPCOUNT, MM_FP4 = 0, 1, 2, 3
SI_TUNING_VERSION = 2


class si_params(Structure):
    _fields_ = [("blocks", c_uint32), ("threads", c_uint32), ("warp", c_uint32), ("path", c_uint32),
                ("buckets", c_uint32), ("bucket_size", c_uint32), ("reserved", c_uint32 * 4)]


class si_tile(Structure):
    _fields_ = [("id", c_uint32), ("start", c_uint32), ("end", c_uint32)]


class si_device_info(Structure):
    _fields_ = [("name", c_char * 64), ("multiprocessor_count", c_int32), ("max_threads_per_block", c_int32),
                ("warp_size", c_int32), ("lds_per_block", c_int32), ("total_global_mem", c_uint64),
                ("free_global_mem", c_uint64), ("clock_khz", c_int32), ("memory_clock_khz", c_int32),
                ("memory_bus_width", c_int32)]


class si_tuning(Structure):
    _fields_ = [("version", c_uint32), ("bitmap_min", c_uint32), ("ratio_search", c_uint32),
                ("chunk_elems", c_uint32), ("chunk_elems_bitmap", c_uint32), ("probe_wg_per_cu", c_uint32),
                ("use_graph", c_uint32), ("edge_bitmap_halves", c_uint32),
                ("reserved", c_uint32 * 8)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


P = c_void_p  # opaque handles
U32P = POINTER(c_uint32)
U64P = POINTER(c_uint64)

# name -> (restype, argtypes); every function declared in the public header
SIGNATURES = {
    "si_last_error": (c_char_p, []),
    "si_open": (c_int, [c_int, POINTER(P)]),
    "si_close": (c_int, [P]),
    "si_device_info_get": (c_int, [P, POINTER(si_device_info)]),
    "si_sync": (c_int, [P]),
    "si_event_record": (c_int, [P, POINTER(c_int32)]),
    "si_event_elapsed": (c_int, [P, c_int32, c_int32, POINTER(c_float)]),
    "si_event_reset": (c_int, [P]),
    "si_kernel_timing": (c_int, [P, c_int32]),
    "si_kernel_stats": (c_int, [P, c_char_p, POINTER(c_double), U64P]),
    "si_kernel_stats_reset": (c_int, [P]),
    "si_buffer_alloc": (c_int, [P, c_uint64, POINTER(P)]),
    "si_buffer_free": (c_int, [P]),
    "si_buffer_device_ptr": (c_int, [P, POINTER(c_void_p)]),
    "si_buffer_download": (c_int, [P, c_uint64, c_void_p, c_uint64]),
    "si_buffer_upload": (c_int, [P, c_uint64, c_void_p, c_uint64]),
    "si_buffer_zero": (c_int, [P]),
    "si_buffer_copy": (c_int, [P, P, c_uint64]),
    "si_host_alloc": (c_int, [P, c_uint64, POINTER(c_void_p)]),
    "si_host_free": (c_int, [P, c_void_p]),
    "si_buffer_download_async": (c_int, [P, c_uint64, c_void_p, c_uint64, POINTER(c_int32), POINTER(c_int32)]),
    "si_event_wait": (c_int, [P, c_int32]),
    "si_dataset_upload": (c_int, [P, c_uint32, c_uint32, c_uint64, U32P, U32P, c_int32, POINTER(P)]),
    "si_dataset_free": (c_int, [P]),
    "si_dataset_info": (c_int, [P, U32P, U32P, U64P]),
    "si_dataset_upload_events": (c_int, [P, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    "si_pair_count": (c_int, [P, c_uint32, c_uint32, c_int32, POINTER(si_params), U64P]),
    "si_pair_count_async": (c_int, [P, c_uint32, c_uint32, c_int32, POINTER(si_params), P, c_uint64]),
    "si_pair_count_noalloc": (c_int, [P, c_uint32, c_uint32, c_int32, POINTER(si_params), P, c_uint64]),
    "si_pair_materialize": (c_int, [P, c_uint32, c_uint32, U32P, c_uint64, U64P]),
    "si_pair_materialize_async": (c_int, [P, c_uint32, c_uint32, P, P]),
    "si_run_count": (c_int, [P, c_int32, POINTER(si_params), si_tile, si_tile, c_uint32, U32P]),
    "si_run_count_async": (c_int, [P, c_int32, POINTER(si_params), si_tile, si_tile, c_uint32, P]),
    "si_all_pairs_bitmap": (c_int, [P, POINTER(si_params), U32P]),
    "si_all_pairs_bitmap_async": (c_int, [P, POINTER(si_params), P]),
    "si_mm_counts": (c_int, [P, si_tile, si_tile, c_int32, c_uint32, U32P]),
    "si_mm_counts_async": (c_int, [P, si_tile, si_tile, c_int32, c_uint32, P]),
    "si_batched_default_tuning": (c_int, [P, POINTER(si_tuning)]),
    "si_batched_count": (c_int, [P, POINTER(si_tuning), U32P]),
    "si_batched_count_async": (c_int, [P, POINTER(si_tuning), P]),
    "si_batched_autotune": (c_int, [P, c_double, POINTER(si_tuning)]),
    "si_batched_autotune_seeded": (c_int, [P, POINTER(si_tuning), c_double, POINTER(si_tuning)]),
    "si_edge_plan_create": (c_int, [P, U32P, U32P, c_uint64, POINTER(si_tuning), POINTER(P)]),
    "si_edge_plan_free": (c_int, [P]),
    "si_edge_count_async": (c_int, [P, P, P]),
    "si_edge_count_plain_async": (c_int, [P, P, P]),
    "si_edge_autotune": (c_int, [P, U32P, U32P, c_uint64, c_double, POINTER(si_tuning)]),
    "si_edge_count": (c_int, [P, U32P, U32P, c_uint64, POINTER(si_tuning), U32P, U64P]),
}

GEN_SIGNATURES = {
    "sigen_last_error": (c_char_p, []),
    "sigen_uniform_sorted": (c_int, [c_uint64, c_uint64, c_uint64, c_uint64, U32P]),
    "sigen_zipf_collection": (c_void_p, [c_uint32, c_double, c_uint32, c_uint32, c_uint64, c_uint64]),
    "sigen_sized_collection": (c_void_p, [c_uint32, U32P, c_uint64, c_uint64]),
    "sigen_density_collection": (c_void_p, [c_uint32, c_uint32, c_double, c_double, c_uint64]),
    "sigen_rmat_collection": (c_void_p, [c_uint32, c_uint32, c_double, c_double, c_double, c_uint64]),
    "sigen_coll_k": (c_uint32, [c_void_p]),
    "sigen_coll_universe": (c_uint32, [c_void_p]),
    "sigen_coll_total": (c_uint64, [c_void_p]),
    "sigen_coll_sizes": (None, [c_void_p, U32P]),
    "sigen_coll_elements": (None, [c_void_p, U32P]),
    "sigen_coll_free": (None, [c_void_p]),
}

_lib = None
_gen = None


class LibraryMissing(RuntimeError):
    pass


def _declare(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """libsi.so with declared signatures; raises LibraryMissing if not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built (run __graft_entry__.build() or "
                                 f"make -C set_intersection_amd)")
        _lib = _declare(ctypes.CDLL(LIB_PATH), SIGNATURES)
    return _lib


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise LibraryMissing(f"{GEN_PATH} is not built")
        _gen = _declare(ctypes.CDLL(GEN_PATH), GEN_SIGNATURES)
    return _gen
