"""CPU oracle (test infrastructure only).

ctypes binding of oracle/liboracle.so, the CPU restatement of the reference's
set-intersection algorithms (see the header of oracle/oracle.cpp for the
file:line each function follows).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use it; the product never does.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

METHODS = {"std": 0, "serial": 1, "bitset": 2, "simd": 3}

_lib = None
U32P = POINTER(c_uint32)
U64P = POINTER(c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", HERE, LIB_PATH], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name in ("oracle_std_count", "oracle_serial_count", "oracle_simd_count"):
            getattr(L, name).restype = c_uint64
            getattr(L, name).argtypes = [U32P, c_uint64, U32P, c_uint64]
        L.oracle_std_materialize.restype = c_uint64
        L.oracle_std_materialize.argtypes = [U32P, c_uint64, U32P, c_uint64, U32P]
        L.oracle_bitset_count.restype = c_uint64
        L.oracle_bitset_count.argtypes = [U32P, c_uint64, U32P, c_uint64, c_uint64]
        L.oracle_triangular_index.restype = c_uint64
        L.oracle_triangular_index.argtypes = [c_uint64, c_uint64, c_uint64]
        L.oracle_all_pairs.restype = c_double
        L.oracle_all_pairs.argtypes = [c_int, c_uint32, c_uint32, U64P, U32P, c_uint32, U32P]
        L.oracle_pairs.restype = c_double
        L.oracle_pairs.argtypes = [c_int, c_uint32, U64P, U32P, U32P, U32P, c_uint64, c_uint32, U32P]
        L.oracle_edges.restype = c_uint64
        L.oracle_edges.argtypes = [U64P, U32P, U32P, U32P, c_uint64, c_uint32, U32P]
        L.oracle_tile_matrix.restype = None
        L.oracle_tile_matrix.argtypes = [U64P, U32P, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, U32P]
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(U32P)


def _p64(a):
    return a.ctypes.data_as(U64P)


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def count(a, b, method="std", universe=None):
    a, b = _arr(a), _arr(b)
    L = lib()
    if method == "std":
        return int(L.oracle_std_count(_p32(a), a.size, _p32(b), b.size))
    if method == "serial":
        return int(L.oracle_serial_count(_p32(a), a.size, _p32(b), b.size))
    if method == "simd":
        return int(L.oracle_simd_count(_p32(a), a.size, _p32(b), b.size))
    if method == "bitset":
        u = universe if universe is not None else 0
        return int(L.oracle_bitset_count(_p32(a), a.size, _p32(b), b.size, u))
    raise KeyError(method)


def materialize(a, b):
    a, b = _arr(a), _arr(b)
    out = np.empty(max(min(a.size, b.size), 1), dtype=np.uint32)
    n = lib().oracle_std_materialize(_p32(a), a.size, _p32(b), b.size, _p32(out))
    return out[:n].copy()


def triangular_index(n, i, j):
    return int(lib().oracle_triangular_index(n, i, j))


def all_pairs(coll, method="std", threads=None):
    """Triangular-order counts of all pairs; returns (counts, seconds)."""
    threads = threads or min(16, os.cpu_count() or 1)
    k = coll.k
    counts = np.zeros(max(k * (k - 1) // 2, 1), dtype=np.uint32)
    offs = np.ascontiguousarray(coll.offsets, dtype=np.uint64)
    el = _arr(coll.elements)
    secs = lib().oracle_all_pairs(METHODS[method], k, coll.universe, _p64(offs), _p32(el), threads, _p32(counts))
    return counts[: k * (k - 1) // 2], secs


def pairs(coll, pa, pb, method="std", threads=None):
    threads = threads or min(16, os.cpu_count() or 1)
    pa, pb = _arr(pa), _arr(pb)
    counts = np.zeros(max(pa.size, 1), dtype=np.uint32)
    offs = np.ascontiguousarray(coll.offsets, dtype=np.uint64)
    el = _arr(coll.elements)
    secs = lib().oracle_pairs(METHODS[method], coll.universe, _p64(offs), _p32(el), _p32(pa), _p32(pb), pa.size,
                              threads, _p32(counts))
    return counts[: pa.size], secs


def edges(coll, u, v, threads=None):
    threads = threads or min(16, os.cpu_count() or 1)
    u, v = _arr(u), _arr(v)
    counts = np.zeros(max(u.size, 1), dtype=np.uint32)
    offs = np.ascontiguousarray(coll.offsets, dtype=np.uint64)
    total = lib().oracle_edges(_p64(offs), _p32(_arr(coll.elements)), _p32(u), _p32(v), u.size, threads,
                               _p32(counts))
    return counts[: u.size], int(total)


def tile_matrix(coll, a0, a1, b0, b1, threads=None):
    threads = threads or min(16, os.cpu_count() or 1)
    ld = b1 - b0
    C = np.zeros(max((a1 - a0) * ld, 1), dtype=np.uint32)
    offs = np.ascontiguousarray(coll.offsets, dtype=np.uint64)
    lib().oracle_tile_matrix(_p64(offs), _p32(_arr(coll.elements)), a0, a1, b0, b1, ld, threads, _p32(C))
    return C[: (a1 - a0) * ld].reshape(a1 - a0, ld)
