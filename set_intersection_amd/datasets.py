"""Collections in the reference's .bin layout and the synthetic workloads.

.bin layout (include/io.hpp:25-54 of the reference): u32 k, u32 universe,
u64 totalElements, u32 sizes[k], u32 elements[total], little-endian, packed.
"""
import ctypes
import os

import numpy as np

from . import _abi

HEADER = np.dtype([("k", "<u4"), ("universe", "<u4"), ("total", "<u8")])


class Collection:
    """k sorted sets stored back to back (sizes[k], elements[total])."""

    def __init__(self, sizes, elements, universe=None):
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        self.elements = np.ascontiguousarray(elements, dtype=np.uint32)
        if int(self.sizes.sum(dtype=np.uint64)) != self.elements.size:
            raise ValueError("sum(sizes) != len(elements)")
        self.offsets = np.zeros(self.sizes.size + 1, dtype=np.uint64)
        np.cumsum(self.sizes, dtype=np.uint64, out=self.offsets[1:])
        if universe is None:
            universe = int(self.elements.max()) if self.elements.size else 0
        self.universe = int(universe)

    @property
    def k(self):
        return int(self.sizes.size)

    def set(self, i):
        return self.elements[int(self.offsets[i]):int(self.offsets[i + 1])]

    @classmethod
    def from_sets(cls, sets, universe=None):
        sets = [np.asarray(s, dtype=np.uint32) for s in sets]
        sizes = np.array([s.size for s in sets], dtype=np.uint32)
        elems = np.concatenate(sets) if sets else np.zeros(0, np.uint32)
        if universe is None:
            universe = max((int(s[-1]) for s in sets if s.size), default=0)
        return cls(sizes, elems, universe)


def read_bin(path):
    """Reads a .bin file; raises ValueError on a truncated/inconsistent file."""
    with open(path, "rb") as f:
        raw = f.read(16)
        if len(raw) < 16:
            raise ValueError(f"truncated dataset file {path!r}")
        h = np.frombuffer(raw, dtype=HEADER)[0]
        k, universe, total = int(h["k"]), int(h["universe"]), int(h["total"])
        sizes = np.fromfile(f, dtype="<u4", count=k)
        if sizes.size != k:
            raise ValueError(f"truncated dataset file {path!r} (sizes)")
        if int(sizes.sum(dtype=np.uint64)) != total:
            raise ValueError(f"dataset {path!r}: sum of sizes != totalElements")
        elems = np.fromfile(f, dtype="<u4", count=total)
        if elems.size != total:
            raise ValueError(f"truncated dataset file {path!r} (elements)")
    return Collection(sizes, elems, universe)


def write_bin(coll, path):
    h = np.zeros(1, dtype=HEADER)
    h["k"], h["universe"], h["total"] = coll.k, coll.universe, coll.elements.size
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(coll.sizes.astype("<u4").tobytes())
        f.write(coll.elements.astype("<u4").tobytes())


# ----------------------------------------------------------- generators --
def _take(h):
    g = _abi.gen()
    if not h:
        raise RuntimeError(g.sigen_last_error().decode())
    try:
        k = g.sigen_coll_k(h)
        total = g.sigen_coll_total(h)
        sizes = np.empty(k, dtype=np.uint32)
        elems = np.empty(total, dtype=np.uint32)
        g.sigen_coll_sizes(h, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        g.sigen_coll_elements(h, elems.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return Collection(sizes, elems, g.sigen_coll_universe(h))
    finally:
        g.sigen_coll_free(h)


def uniform_sorted(n, N, seed, lo=0):
    """n distinct values, uniform over [lo, lo+N), sorted (exact size)."""
    out = np.empty(int(n), dtype=np.uint32)
    g = _abi.gen()
    if g.sigen_uniform_sorted(int(n), int(N), int(seed), int(lo),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))) != 0:
        raise RuntimeError(g.sigen_last_error().decode())
    return out


def sisi_pair(na, nb, universe, seed):
    """One pair of sorted unique sets drawn uniformly from [0, universe)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(2) as ex:  # the generator releases the GIL
        fa = ex.submit(uniform_sorted, na, universe, seed)
        fb = ex.submit(uniform_sorted, nb, universe, seed + 1)
        a, b = fa.result(), fb.result()
    return Collection.from_sets([a, b], universe=universe - 1)


def zipf_collection(k, s, min_size, max_size, universe, seed):
    return _take(_abi.gen().sigen_zipf_collection(k, s, min_size, max_size, universe, seed))


def sized_collection(sizes, universe, seed):
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    return _take(_abi.gen().sigen_sized_collection(sizes.size, sizes.ctypes.data_as(
        ctypes.POINTER(ctypes.c_uint32)), universe, seed))


def density_collection(k, universe, pmin, pmax, seed):
    return _take(_abi.gen().sigen_density_collection(k, universe, pmin, pmax, seed))


def rmat_collection(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    return _take(_abi.gen().sigen_rmat_collection(scale, edge_factor, a, b, c, seed))


def edge_list(coll):
    """Edges (u, v) with u < v of an adjacency collection, in CSR order."""
    u = np.repeat(np.arange(coll.k, dtype=np.uint32), coll.sizes)
    v = coll.elements
    keep = u < v
    return u[keep].copy(), v[keep].copy()


# the workloads the benchmarks are quoted on
WORKLOADS = {
    "sisi_2^20_in_2^24": dict(kind="sisi", na=1 << 20, nb=1 << 20, universe=1 << 24),
    "sisi_2^28_in_2^32": dict(kind="sisi", na=1 << 28, nb=1 << 28, universe=1 << 32),
    "sisi_2^27x2^16_in_2^32": dict(kind="sisi", na=1 << 27, nb=1 << 16, universe=1 << 32),
    "misi_zipf1.2_16384": dict(kind="zipf", k=16384, s=1.2, min_size=16, max_size=65536, universe=1 << 20),
    "mm_8192x2^16_density1-10": dict(kind="density", k=8192, universe=1 << 16, pmin=0.01, pmax=0.10),
    "rmat20_ef16": dict(kind="rmat", scale=20, edge_factor=16),
}

BASE_SEED = 20261203


def make_workload(name, seed=BASE_SEED):
    w = WORKLOADS[name]
    if w["kind"] == "sisi":
        return sisi_pair(w["na"], w["nb"], w["universe"], seed)
    if w["kind"] == "zipf":
        return zipf_collection(w["k"], w["s"], w["min_size"], w["max_size"], w["universe"], seed)
    if w["kind"] == "density":
        return density_collection(w["k"], w["universe"], w["pmin"], w["pmax"], seed)
    if w["kind"] == "rmat":
        return rmat_collection(w["scale"], w["edge_factor"], seed=seed)
    raise KeyError(name)
