"""GPU parity of the batched dispatcher, the bitmap-GEMM path and the
edge-wise counts against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets

pytestmark = pytest.mark.gpu


def tuning(ds, **kw):
    t = ds.default_tuning()
    for k, v in kw.items():
        setattr(t, k, v)
    return t


@pytest.mark.parametrize("graph", [0, 1])
def test_batched_zipf_default(ctx, graph):
    coll = datasets.zipf_collection(600, 1.2, 16, 65536, 1 << 20, 21)
    ds = ctx.upload(coll)
    exp, _ = oracle.all_pairs(coll)
    t = tuning(ds, use_graph=graph)
    assert np.array_equal(ds.batched_count(t), exp)
    assert np.array_equal(ds.batched_count(t), exp)  # replay


@pytest.mark.parametrize("bitmap_min,chunk,chunk_bm", [(1, 64, 64), (64, 1000, 5000), (100000, 1 << 16, 1 << 20)])
def test_batched_bins_and_chunks(ctx, bitmap_min, chunk, chunk_bm):
    """Force every row into the bitmap bin (bitmap_min=1) or the hash bin,
    with tiny chunks so sets are cut by wave boundaries."""
    coll = datasets.zipf_collection(300, 1.1, 1, 9000, 1 << 18, 5)
    ds = ctx.upload(coll)
    exp, _ = oracle.all_pairs(coll)
    t = tuning(ds, bitmap_min=bitmap_min, chunk_elems=chunk, chunk_elems_bitmap=chunk_bm, use_graph=0)
    assert np.array_equal(ds.batched_count(t), exp)


@pytest.mark.parametrize("ratio", [1, 4, 1 << 30])
def test_batched_big_rows_search_and_merge(ctx, ratio):
    """Universe 2^32 (no LDS bitmap) with rows larger than an LDS hash table:
    the SEARCH and MERGE bins."""
    sizes = [60000, 30000, 20000, 9000, 5000, 4097, 300, 20, 3, 1, 0, 0]
    coll = datasets.sized_collection(sizes, 1 << 20, 8)
    # spread the values over the full uint32 range, keep them sorted and unique
    coll = si.Collection(coll.sizes, coll.elements.astype(np.uint64) * 4093 + 7, universe=2**32 - 1)
    ds = ctx.upload(coll)
    exp, _ = oracle.all_pairs(coll)
    t = tuning(ds, ratio_search=ratio, use_graph=0, chunk_elems=4096)
    assert np.array_equal(ds.batched_count(t), exp)


@pytest.mark.parametrize("graph", [0, 1])
def test_batched_bitmap_rows_many_tiny_partners(ctx, graph):
    """Default tuning: two large bitmap rows whose partners are thousands of
    sets of 1-3 elements, so one wave's chunk spans more than 128 sets (the
    set bounds loaded beside a chunk cover 128; the rest are loaded in the
    counting loop), plus partial chunks at every wave's end."""
    rng = np.random.default_rng(11)
    sizes = [20000, 15000, 600, 40] + list(rng.integers(1, 4, 3000))
    coll = datasets.sized_collection(sizes, 1 << 16, 12)
    ds = ctx.upload(coll)
    exp, _ = oracle.all_pairs(coll)
    t = tuning(ds, use_graph=graph)
    assert np.array_equal(ds.batched_count(t), exp)
    assert np.array_equal(ds.batched_count(t), exp)


def test_batched_empty_and_tiny(ctx):
    coll = si.Collection.from_sets([[], [1, 2, 3], [], [2], [3, 4, 2**32 - 1], [2**32 - 1], []])
    ds = ctx.upload(coll)
    exp, _ = oracle.all_pairs(coll)
    assert np.array_equal(ds.batched_count(), exp)
    one = ctx.upload(si.Collection.from_sets([[1, 2]]))
    assert one.batched_count().size == 0


def test_batched_autotune(ctx):
    coll = datasets.zipf_collection(500, 1.2, 16, 65536, 1 << 20, 2)
    ds = ctx.upload(coll)
    t = ds.autotune(300.0)
    assert t.version == si._abi.SI_TUNING_VERSION
    exp, _ = oracle.all_pairs(coll)
    assert np.array_equal(ds.batched_count(t), exp)


# ------------------------------------------------------------ bitmap GEMM --
@pytest.mark.parametrize("variant", [si.MM_I8, si.MM_F32, si.MM_POPCOUNT, si.MM_FP4])
def test_mm_tiles(ctx, variant):
    coll = datasets.density_collection(300, 1 << 12, 0.01, 0.3, 11)
    ds = ctx.upload(coll)
    # self tile (mirrored upper tiles), off-diagonal tile pair, ragged tiles
    for (a0, a1), (b0, b1) in [((0, 300), (0, 300)), ((0, 130), (130, 300)), ((7, 140), (200, 261)),
                               ((250, 300), (0, 3))]:
        A, B = si.tile(0, a0, a1), si.tile(1, b0, b1)
        got = ds.mm_counts(A, B, variant)
        exp = oracle.tile_matrix(coll, a0, a1, b0, b1)
        assert np.array_equal(got, exp), ((a0, a1), (b0, b1))


def test_mm_fp32_range(ctx):
    """The FP32-accumulated variants refuse a universe above 2^24 (counts could
    round); the int8 path takes it."""
    coll = datasets.density_collection(4, 1 << 12, 0.1, 0.2, 5)
    coll.universe = (1 << 24) + 5
    ds = ctx.upload(coll)
    A = si.tile(0, 0, 4)
    for v in (si.MM_FP4, si.MM_F32):
        with pytest.raises(si.SIError):
            ds.mm_counts(A, A, v)
    assert np.array_equal(ds.mm_counts(A, A, si.MM_I8), oracle.tile_matrix(coll, 0, 4, 0, 4))


def test_mm_ld_padding(ctx):
    coll = datasets.density_collection(70, 1 << 10, 0.05, 0.5, 3)
    ds = ctx.upload(coll)
    A, B = si.tile(0, 0, 40), si.tile(1, 40, 70)
    got = ds.mm_counts(A, B, si.MM_I8, ld=50)
    exp = oracle.tile_matrix(coll, 0, 40, 40, 70)
    assert np.array_equal(got[:, :30], exp)


@pytest.mark.parametrize("variant", [si.MM_I8, si.MM_F32, si.MM_FP4])
def test_mm_c4_shape_slice(ctx, variant):
    """A 1024-set slice of the 8192 x 2^16 density workload (full K)."""
    coll = datasets.density_collection(1024, 1 << 16, 0.01, 0.10, 20261203)
    ds = ctx.upload(coll)
    A = si.tile(0, 0, 1024)
    got = ds.mm_counts(A, A, variant)
    exp = oracle.tile_matrix(coll, 0, 1024, 0, 1024)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("variant", [si.MM_I8, si.MM_F32, si.MM_FP4])
def test_mm_split_tail(ctx, variant):
    """6000 sets: 1128 self tiles, more than one round of resident workgroups,
    so the last partial round is split along K and accumulated atomically.
    Checked against the independent AND/popcount kernel everywhere and the
    oracle on sampled rows."""
    coll = datasets.density_collection(6000, 1 << 12, 0.01, 0.10, 77)
    ds = ctx.upload(coll)
    A = si.tile(0, 0, 6000)
    got = ds.mm_counts(A, A, variant)
    ref = ds.mm_counts(A, A, si.MM_POPCOUNT)
    assert np.array_equal(got, ref)
    for r0 in (0, 2990, 5936):
        assert np.array_equal(got[r0:r0 + 64], oracle.tile_matrix(coll, r0, r0 + 64, 0, 6000)), r0


# ------------------------------------------------------------------ edges --
def test_edges_rmat_small(ctx):
    coll = datasets.rmat_collection(12, 16, seed=4)
    u, v = datasets.edge_list(coll)
    ds = ctx.upload(coll)
    got, total = ds.edge_count(u, v)
    exp, etot = oracle.edges(coll, u, v)
    assert np.array_equal(got, exp)
    assert total == etot and total % 3 == 0


@pytest.mark.parametrize("bitmap_min,ratio", [(1, 32), (1 << 30, 32), (1 << 30, 1)])
def test_edges_bins(ctx, bitmap_min, ratio):
    coll = datasets.rmat_collection(11, 24, seed=9)
    u, v = datasets.edge_list(coll)
    ds = ctx.upload(coll)
    t = tuning(ds, bitmap_min=bitmap_min, ratio_search=ratio, chunk_elems=512)
    got, total = ds.edge_count(u, v, t)
    exp, etot = oracle.edges(coll, u, v)
    assert np.array_equal(got, exp) and total == etot


def _halves_collection():
    """Lists over [0, 2^20) that straddle the universe midpoint 2^19 in every
    way: all below, all above, elements at 2^19 - 1, 2^19 and 2^20 - 1, long
    rows cut into several items, and an empty list."""
    rng = np.random.default_rng(11)
    H = 1 << 19
    sets = []
    for i, n in enumerate([5000, 3000, 2000, 1200, 900, 700, 600, 520, 300, 64, 40, 7, 1, 0]):
        if i % 4 == 1:
            x = rng.integers(0, H, n)                # lower half only
        elif i % 4 == 2:
            x = rng.integers(H, 2 * H, n)            # upper half only
        else:
            x = rng.integers(0, 2 * H, n)
        sets.append(np.unique(x).astype(np.uint32))
    sets[0] = np.unique(np.concatenate([sets[0], [H - 1, H, 2 * H - 1, 0]])).astype(np.uint32)
    sets[5] = np.unique(np.concatenate([sets[5], [H - 1]])).astype(np.uint32)  # ends just below the midpoint
    sets[6] = np.unique(np.concatenate([sets[6], [H]])).astype(np.uint32)      # starts at the midpoint
    sets.append(sets[0][::3].copy())                                            # a subset of a row
    return si.Collection.from_sets(sets)


@pytest.mark.parametrize("shift", [0, 1])
def test_edges_bitmap_halves(ctx, shift):
    """Bitmap rows in 64 KiB bitmaps, two workgroups per CU
    (edge_bitmap_halves = 1): universe halves with the partners' lists split
    at 2^19 (shift 0, universe 2^20), and whole-universe 64 KiB bitmaps
    (shift 1: the same lists halved into a 2^19 universe).  Counts equal the
    oracle's and the 128 KiB path's, with small work items (rows cut into
    several items, partners cut by wave boundaries), twice on one plan."""
    coll = _halves_collection()
    if shift:
        coll = si.Collection.from_sets([np.unique(s >> 1).astype(np.uint32) for s in (coll.set(i) for i in range(coll.k))])
    k = coll.k
    rng = np.random.default_rng(5)
    u = rng.integers(0, k, 4000).astype(np.uint32)
    v = rng.integers(0, k, 4000).astype(np.uint32)
    ds = ctx.upload(coll)
    exp, etot = oracle.edges(coll, u, v)
    for chunk in (1 << 20, 700):
        t = tuning(ds, bitmap_min=512, chunk_elems_bitmap=chunk, edge_bitmap_halves=1)
        plan = ds.edge_plan(u, v, t)
        counts = ctx.buffer(4 * u.size)
        total = ctx.buffer(8)
        for _ in range(2):
            plan.count_async(counts, total)
            assert np.array_equal(counts.download(np.uint32), exp), chunk
            assert int(total.download(np.uint64)[0]) == etot
        got0, tot0 = ds.edge_count(u, v, tuning(ds, bitmap_min=512, chunk_elems_bitmap=chunk,
                                                edge_bitmap_halves=0))
        assert np.array_equal(got0, exp) and tot0 == etot


def test_edges_arbitrary_pairs(ctx):
    coll = datasets.zipf_collection(200, 1.1, 1, 70000, 2**32, 6)  # big rows, no bitmap
    rng = np.random.default_rng(0)
    u = rng.integers(0, 200, 3000).astype(np.uint32)
    v = rng.integers(0, 200, 3000).astype(np.uint32)
    ds = ctx.upload(coll)
    got, total = ds.edge_count(u, v)
    exp, etot = oracle.edges(coll, u, v)
    assert np.array_equal(got, exp) and total == etot


@pytest.mark.parametrize("case", ["rmat", "arbitrary"])
def test_edges_plain_baseline(ctx, case):
    """The plain per-edge kernel (one wave per edge, binary search of the
    shorter list in the longer) on the same plan: counts and total equal the
    oracle's and the binned dispatcher's, twice in a row (the total's
    workgroup ticket resets itself), including self loops, repeated edges and
    empty lists in the arbitrary pairs."""
    if case == "rmat":
        coll = datasets.rmat_collection(12, 16, seed=4)
        u, v = datasets.edge_list(coll)
    else:
        sets = [np.unique(np.random.default_rng(i).integers(0, 2**32, n, dtype=np.uint64)).astype(np.uint32)
                for i, n in enumerate([0, 1, 5, 64, 65, 4000, 70000, 0, 300])]
        sets[3] = sets[6][::1000].copy()  # a subset of a long list
        coll = si.Collection.from_sets(sets)
        rng = np.random.default_rng(3)
        u = rng.integers(0, coll.k, 5000).astype(np.uint32)
        v = rng.integers(0, coll.k, 5000).astype(np.uint32)
    ds = ctx.upload(coll)
    plan = ds.edge_plan(u, v)
    exp, etot = oracle.edges(coll, u, v)
    counts = ctx.buffer(4 * u.size)
    total = ctx.buffer(8)
    for _ in range(2):
        counts.zero()
        plan.count_plain_async(counts, total)
        assert np.array_equal(counts.download(np.uint32), exp)
        assert int(total.download(np.uint64)[0]) == etot
    plan.count_async(counts, total)
    assert np.array_equal(counts.download(np.uint32), exp)


def test_edges_autotune(ctx):
    """si_edge_autotune returns a valid record within its budget; counts with
    it equal the oracle's."""
    coll = datasets.rmat_collection(11, 16, seed=8)
    u, v = datasets.edge_list(coll)
    ds = ctx.upload(coll)
    t = ds.edge_autotune(u, v, budget_ms=300.0)
    assert t.version == 2 and t.chunk_elems > 0 and t.probe_wg_per_cu > 0
    got, total = ds.edge_count(u, v, t)
    exp, etot = oracle.edges(coll, u, v)
    assert np.array_equal(got, exp) and total == etot
