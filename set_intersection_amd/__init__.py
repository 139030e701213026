"""MI355X (gfx950) set intersection: Python binding of libsi.so.

The product is the C ABI (include/set_intersection.h) and the per-technique
executables under set_intersection_amd/bin; this module is a thin ctypes
binding used by the tests and bench.py.  Every call goes to the HIP kernels;
a missing library raises LibraryMissing and there is no CPU fallback.
"""
import ctypes
import weakref
from ctypes import byref, c_double, c_float, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _abi
from ._abi import (MM_F32, MM_FP4, MM_I8, MM_POPCOUNT, TECH_BITMAP_DYNAMIC, TECH_BITMAP_NAIVE, TECH_HI, TECH_IP,
                   TECH_MM, TECH_OBS, TECHNIQUES, LibraryMissing, si_params, si_tile, si_tuning)
from .datasets import Collection, read_bin, write_bin

__all__ = ["Context", "Dataset", "Buffer", "SIError", "LibraryMissing", "Collection", "read_bin", "write_bin",
           "TECHNIQUES", "TECH_IP", "TECH_OBS", "TECH_HI", "TECH_BITMAP_NAIVE", "TECH_BITMAP_DYNAMIC",
           "TECH_MM", "MM_I8", "MM_F32", "MM_POPCOUNT", "MM_FP4", "params", "tile", "run_slots"]


class SIError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{_abi.STATUS_NAMES.get(status, status)}: {message}")


def _check(status):
    if status != _abi.SI_OK:
        raise SIError(status, _abi.lib().si_last_error().decode(errors="replace"))


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(c_uint32))


def params(blocks=0, threads=0, warp=False, path=False, buckets=0, bucket_size=0):
    p = si_params()
    p.blocks, p.threads, p.warp, p.path = blocks, threads, int(bool(warp)), int(bool(path))
    p.buckets, p.bucket_size = buckets, bucket_size
    return p


def tile(id_, start, end):
    t = si_tile()
    t.id, t.start, t.end = id_, start, end
    return t


def run_slots(A, B, n):
    return n * (n - 1) // 2 if A.id == B.id else n * n


class Buffer:
    """Device buffer owned by the library."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.h = c_void_p()
        _check(_abi.lib().si_buffer_alloc(ctx.h, self.nbytes, byref(self.h)))
        ctx._children.add(self)

    def download(self, dtype=np.uint32, count=None, offset=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        _check(_abi.lib().si_buffer_download(self.h, offset, out.ctypes.data, out.nbytes))
        return out

    def download_async(self, host, offset=0):
        """Enqueues a copy of host.nbytes bytes at `offset` into the writable
        contiguous array `host` on the context's copy stream, after the work
        enqueued so far (si_buffer_download_async).  Returns the ids of the
        events bracketing the copy; host is final after ctx.wait(end)."""
        if not (host.flags.c_contiguous and host.flags.writeable):
            raise ValueError("host must be a writable contiguous array")
        e0, e1 = c_int32(), c_int32()
        _check(_abi.lib().si_buffer_download_async(self.h, offset, host.ctypes.data, host.nbytes, byref(e0),
                                                    byref(e1)))
        return e0.value, e1.value

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        _check(_abi.lib().si_buffer_upload(self.h, offset, arr.ctypes.data, arr.nbytes))

    def zero(self):
        _check(_abi.lib().si_buffer_zero(self.h))

    def copy_from(self, src, nbytes):
        _check(_abi.lib().si_buffer_copy(self.h, src.h, nbytes))

    def free(self):
        if self.h and self.ctx.h:
            _abi.lib().si_buffer_free(self.h)
        self.h = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class EdgePlan:
    """Edges grouped for the binned dispatcher (si_edge_plan_create)."""

    def __init__(self, ds, u, v, tuning=None):
        self.ds = ds
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        self.m = int(u.size)
        self.h = c_void_p()
        _check(_abi.lib().si_edge_plan_create(ds.h, _u32p(u), _u32p(v), u.size, byref(tuning) if tuning else None,
                                              byref(self.h)))
        ds.ctx._children.add(self)

    def count_async(self, counts, total):
        _check(_abi.lib().si_edge_count_async(self.h, counts.h, total.h))

    def count_plain_async(self, counts, total):
        """The plain per-edge baseline (si_edge_count_plain_async)."""
        _check(_abi.lib().si_edge_count_plain_async(self.h, counts.h, total.h))

    def free(self):
        if self.h and self.ds.h and self.ds.ctx.h:
            _abi.lib().si_edge_plan_free(self.h)
        self.h = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU and one HIP stream (si_open)."""

    def __init__(self, device=0):
        self.h = c_void_p()
        self._children = weakref.WeakSet()
        self._host = {}  # host_array data address -> si_host_alloc pointer
        _check(_abi.lib().si_open(device, byref(self.h)))

    def info(self):
        i = _abi.si_device_info()
        _check(_abi.lib().si_device_info_get(self.h, byref(i)))
        return {f: (getattr(i, f).decode() if f == "name" else getattr(i, f)) for f, _ in i._fields_}

    def sync(self):
        _check(_abi.lib().si_sync(self.h))

    def record(self):
        eid = c_int32()
        _check(_abi.lib().si_event_record(self.h, byref(eid)))
        return eid.value

    def elapsed_ms(self, e0, e1):
        ms = c_float()
        _check(_abi.lib().si_event_elapsed(self.h, e0, e1, byref(ms)))
        return ms.value

    def reset_events(self):
        _check(_abi.lib().si_event_reset(self.h))

    def wait(self, event_id):
        """Blocks until the event has completed (si_event_wait)."""
        _check(_abi.lib().si_event_wait(self.h, event_id))

    def host_array(self, count, dtype=np.uint32):
        """A numpy array over page-locked host memory (si_host_alloc).  It is
        valid until host_free(array) or close()."""
        dtype = np.dtype(dtype)
        nbytes = max(int(count) * dtype.itemsize, 1)
        p = c_void_p()
        _check(_abi.lib().si_host_alloc(self.h, nbytes, byref(p)))
        raw = (ctypes.c_uint8 * nbytes).from_address(p.value)
        arr = np.frombuffer(raw, dtype=np.uint8)[: int(count) * dtype.itemsize].view(dtype)
        self._host[arr.ctypes.data] = p.value
        return arr

    def host_free(self, arr):
        p = self._host.pop(arr.ctypes.data)
        _check(_abi.lib().si_host_free(self.h, c_void_p(p)))

    def kernel_timing(self, enable=True):
        _check(_abi.lib().si_kernel_timing(self.h, int(enable)))

    def kernel_stats(self, name):
        ms, n = c_double(), c_uint64()
        _check(_abi.lib().si_kernel_stats(self.h, name.encode(), byref(ms), byref(n)))
        return ms.value, n.value

    def kernel_stats_reset(self):
        _check(_abi.lib().si_kernel_stats_reset(self.h))

    def buffer(self, nbytes):
        return Buffer(self, nbytes)

    def upload(self, coll, validate=True):
        return Dataset(self, coll, validate)

    def close(self):
        """Frees every buffer, dataset and plan created from this context, then
        the context itself."""
        if self.h:
            for kind in (EdgePlan, Dataset, Buffer):
                for child in list(self._children):
                    if isinstance(child, kind):
                        child.free()
            _abi.lib().si_close(self.h)
            self._host.clear()
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Dataset:
    """A collection resident in HBM (si_dataset_upload)."""

    def __init__(self, ctx, coll, validate=True):
        self.ctx = ctx
        self.coll = coll
        self.h = c_void_p()
        sizes = np.ascontiguousarray(coll.sizes, dtype=np.uint32)
        elems = np.ascontiguousarray(coll.elements, dtype=np.uint32)
        _check(_abi.lib().si_dataset_upload(ctx.h, len(sizes), int(coll.universe), int(elems.size), _u32p(sizes),
                                            _u32p(elems), int(bool(validate)), byref(self.h)))
        self.k = len(sizes)
        ctx._children.add(self)

    # ---- SISI
    def pair_count(self, a, b, technique=TECH_IP, p=None):
        out = c_uint64()
        _check(_abi.lib().si_pair_count(self.h, a, b, technique, byref(p) if p else None, byref(out)))
        return out.value

    def pair_count_async(self, a, b, out, offset=0, technique=TECH_IP, p=None):
        _check(_abi.lib().si_pair_count_async(self.h, a, b, technique, byref(p) if p else None, out.h, offset))

    def pair_count_noalloc(self, a, b, out, offset=0, technique=TECH_IP, p=None):
        """si_pair_count_noalloc: as pair_count_async, with the kernels'
        scratch in the library's static device array (no allocation)."""
        _check(_abi.lib().si_pair_count_noalloc(self.h, a, b, technique, byref(p) if p else None, out.h, offset))

    def pair_materialize(self, a, b):
        cap = int(min(self.coll.sizes[a], self.coll.sizes[b]))
        out = np.empty(max(cap, 1), dtype=np.uint32)
        n = c_uint64()
        _check(_abi.lib().si_pair_materialize(self.h, a, b, _u32p(out), cap, byref(n)))
        return out[: n.value].copy()

    def pair_materialize_async(self, a, b, out, count_out):
        _check(_abi.lib().si_pair_materialize_async(self.h, a, b, out.h, count_out.h))

    # ---- MISI
    def run_count(self, technique, A, B, n, p=None):
        counts = np.zeros(max(run_slots(A, B, n), 1), dtype=np.uint32)
        _check(_abi.lib().si_run_count(self.h, technique, byref(p) if p else None, A, B, n, _u32p(counts)))
        return counts[: run_slots(A, B, n)]

    def run_count_async(self, technique, A, B, n, counts, p=None):
        _check(_abi.lib().si_run_count_async(self.h, technique, byref(p) if p else None, A, B, n, counts.h))

    def all_pairs_bitmap_async(self, counts, p=None):
        """si_all_pairs_bitmap_async: every pair's count into `counts`
        (uint32, triangular order) by bitmap AND + popcount."""
        _check(_abi.lib().si_all_pairs_bitmap_async(self.h, byref(p) if p else None, counts.h))

    def all_pairs_bitmap(self, p=None):
        k = self.k
        counts = np.zeros(max(k * (k - 1) // 2, 1), dtype=np.uint32)
        _check(_abi.lib().si_all_pairs_bitmap(self.h, byref(p) if p else None, _u32p(counts)))
        return counts[: k * (k - 1) // 2]

    def mm_counts(self, A, B, variant=MM_I8, ld=None):
        ld = ld or (B.end - B.start)
        out = np.zeros(max((A.end - A.start) * ld, 1), dtype=np.uint32)
        _check(_abi.lib().si_mm_counts(self.h, A, B, variant, ld, _u32p(out)))
        return out[: (A.end - A.start) * ld].reshape(A.end - A.start, ld)

    def mm_counts_async(self, A, B, counts, variant=MM_I8, ld=None):
        ld = ld or (B.end - B.start)
        _check(_abi.lib().si_mm_counts_async(self.h, A, B, variant, ld, counts.h))

    def default_tuning(self):
        t = si_tuning()
        _check(_abi.lib().si_batched_default_tuning(self.h, byref(t)))
        return t

    def batched_count(self, tuning=None):
        k = self.k
        counts = np.zeros(max(k * (k - 1) // 2, 1), dtype=np.uint32)
        _check(_abi.lib().si_batched_count(self.h, byref(tuning) if tuning else None, _u32p(counts)))
        return counts[: k * (k - 1) // 2]

    def batched_count_async(self, counts, tuning=None):
        _check(_abi.lib().si_batched_count_async(self.h, byref(tuning) if tuning else None, counts.h))

    def autotune(self, budget_ms=2000.0, seed=None):
        """si_batched_autotune, or si_batched_autotune_seeded from `seed`."""
        t = si_tuning()
        if seed is None:
            _check(_abi.lib().si_batched_autotune(self.h, budget_ms, byref(t)))
        else:
            _check(_abi.lib().si_batched_autotune_seeded(self.h, byref(seed), budget_ms, byref(t)))
        return t

    def edge_count(self, u, v, tuning=None):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        counts = np.zeros(max(u.size, 1), dtype=np.uint32)
        total = c_uint64()
        _check(_abi.lib().si_edge_count(self.h, _u32p(u), _u32p(v), u.size, byref(tuning) if tuning else None,
                                        _u32p(counts), byref(total)))
        return counts[: u.size], total.value

    def edge_autotune(self, u, v, budget_ms=5000.0):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        t = si_tuning()
        _check(_abi.lib().si_edge_autotune(self.h, _u32p(u), _u32p(v), u.size, budget_ms, byref(t)))
        return t

    def edge_plan(self, u, v, tuning=None):
        return EdgePlan(self, u, v, tuning)

    def upload_ms(self):
        """(transfer, offsets scan) milliseconds of the upload, from the
        events si_dataset_upload recorded (si_dataset_upload_events)."""
        ev = [c_int32(), c_int32(), c_int32()]
        _check(_abi.lib().si_dataset_upload_events(self.h, *[byref(e) for e in ev]))
        return self.ctx.elapsed_ms(ev[0].value, ev[1].value), self.ctx.elapsed_ms(ev[1].value, ev[2].value)

    def free(self):
        if self.h:
            if self.ctx.h:
                _abi.lib().si_dataset_free(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
