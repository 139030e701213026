"""Shared helpers of the GPU parity tests."""
import numpy as np

import oracle
import set_intersection_amd as si


def run_layout_counts(ds, coll, technique, partition, p=None):
    """All pair counts through the reference's multi-run driver layout
    (findTilePairs runs, triangular index for self-join runs, quadratic
    otherwise, numOfSets = k for a single run else partition)."""
    k = coll.k
    tiles = [(t, t * partition, min((t + 1) * partition, k)) for t in range((k + partition - 1) // partition)]
    single = len(tiles) == 1
    n = k if single else partition
    res = {}
    for x in range(len(tiles)):
        for y in range(x, len(tiles)):
            A, B = si.tile(*tiles[x]), si.tile(*tiles[y])
            c = ds.run_count(technique, A, B, n, p)
            self_join = A.id == B.id
            for i in range(A.start, A.end):
                for j in range(i + 1 if self_join else B.start, B.end):
                    li, lj = i - A.id * n, j - B.id * n
                    idx = oracle.triangular_index(n, li, lj) if self_join else li * n + lj
                    res[(i, j)] = int(c[idx])
    return res


def tri_start(k, i):
    """triangular_index(k, i, i + 1): where row i's pairs (i, j > i) begin."""
    return k * (k - 1) // 2 - (k - i) * (k - i - 1) // 2


def run_layout_triangular(ds, coll, technique, partition, p=None):
    """All pair counts through the multi-run driver layout (as
    run_layout_counts), reassembled into the single-run triangular order
    counts[triangular_index(k, i, j)] with row slices instead of per-pair
    Python loops, for collections of thousands of sets."""
    k = coll.k
    tiles = [(t, t * partition, min((t + 1) * partition, k)) for t in range((k + partition - 1) // partition)]
    n = k if len(tiles) == 1 else partition
    out = np.zeros(k * (k - 1) // 2, dtype=np.uint32)
    for x in range(len(tiles)):
        for y in range(x, len(tiles)):
            A, B = si.tile(*tiles[x]), si.tile(*tiles[y])
            c = ds.run_count(technique, A, B, n, p)
            for i in range(A.start, A.end):
                li = i - A.id * n
                if A.id == B.id:
                    ln = B.end - i - 1
                    src = tri_start(n, li)
                    dst = tri_start(k, i)
                else:
                    ln = B.end - B.start
                    src = li * n
                    dst = tri_start(k, i) + B.start - i - 1
                if ln > 0:
                    out[dst:dst + ln] = c[src:src + ln]
    return out


def check_all_pairs(got, coll):
    counts, _ = oracle.all_pairs(coll)
    k = coll.k
    assert len(got) == k * (k - 1) // 2
    bad = [(i, j, c, int(counts[oracle.triangular_index(k, i, j)])) for (i, j), c in got.items()
           if c != counts[oracle.triangular_index(k, i, j)]]
    assert not bad, bad[:10]


def edge_case_pairs():
    """(a, b) pairs covering empty/singleton sets, signedness and range ends."""
    top = np.array([0, 1, 2**31 - 1, 2**31, 2**31 + 1, 2**32 - 2, 2**32 - 1], dtype=np.uint32)
    empty = np.zeros(0, dtype=np.uint32)
    return [(top, top), (top, empty), (empty, empty), (top, top[[0, 3, 6]]),
            (np.array([5], np.uint32), np.array([5], np.uint32)),
            (np.array([5], np.uint32), np.array([6], np.uint32)),
            (np.arange(0, 20000, 2, dtype=np.uint32), np.arange(1, 20000, 2, dtype=np.uint32)),
            (np.arange(0, 9001, dtype=np.uint32), np.arange(9000, 30000, dtype=np.uint32)),
            (np.arange(0, 100000, 3, dtype=np.uint32), np.arange(0, 100000, 7, dtype=np.uint32))]
