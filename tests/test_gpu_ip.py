"""GPU parity of the merge-path (ip) kernels against the CPU oracle.

Counts and materialised lists must be bit-exact.  Sizes are chosen to hit
tile (4096 merged elements), wave (64) and segment boundaries, empty and
singleton sets, values straddling 2^31 and the top of the uint32 range.
"""
import os

import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets
from helpers import run_layout_counts

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_pair(ctx, a, b, universe=None):
    coll = si.Collection.from_sets([a, b], universe=universe)
    ds = ctx.upload(coll)
    exp = oracle.materialize(coll.set(0), coll.set(1))
    assert ds.pair_count(0, 1) == exp.size
    assert ds.pair_count(1, 0) == exp.size
    got = ds.pair_materialize(0, 1)
    assert np.array_equal(got, exp)
    got = ds.pair_materialize(1, 0)
    assert np.array_equal(got, exp)
    ds.free()


def test_closed_form_multiples(ctx):
    for N, p, q in [(100_000, 2, 3), (1_000_003, 4, 6), (50_000, 7, 7), (300_000, 1, 5), (2_000_000, 97, 89)]:
        a = np.arange(0, N, p, dtype=np.uint32)
        b = np.arange(0, N, q, dtype=np.uint32)
        check_pair(ctx, a, b)


def test_edge_sets(ctx):
    top = np.array([0, 1, 2**31 - 1, 2**31, 2**31 + 1, 2**32 - 2, 2**32 - 1], dtype=np.uint32)
    empty = np.zeros(0, dtype=np.uint32)
    check_pair(ctx, top, top)
    check_pair(ctx, top, empty)
    check_pair(ctx, empty, empty)
    check_pair(ctx, top, top[[0, 3, 6]])
    check_pair(ctx, np.array([5], np.uint32), np.array([5], np.uint32))
    check_pair(ctx, np.array([5], np.uint32), np.array([6], np.uint32))
    # disjoint interleaved and disjoint ranges
    check_pair(ctx, np.arange(0, 20000, 2, dtype=np.uint32), np.arange(1, 20000, 2, dtype=np.uint32))
    check_pair(ctx, np.arange(0, 9000, dtype=np.uint32), np.arange(9000, 30000, dtype=np.uint32))
    # all of A before all of B except one shared element at the seam
    check_pair(ctx, np.arange(0, 9001, dtype=np.uint32), np.arange(9000, 30000, dtype=np.uint32))


@pytest.mark.parametrize("na,nb", [(63, 64), (64, 65), (4095, 1), (4096, 4096), (4097, 4095), (8191, 8193),
                                   (1, 100_000), (100_000, 3), (123_457, 98_765)])
def test_boundary_sizes(ctx, na, nb):
    rng = np.random.default_rng(na * 7 + nb)
    U = max(na, nb) * 3
    a = datasets.uniform_sorted(na, U, 11 + na)
    b = datasets.uniform_sorted(nb, U, 13 + nb)
    # plant matches so that even sparse draws intersect
    if na and nb:
        extra = rng.choice(a, min(na, nb, 50), replace=False)
        b = np.union1d(b, extra).astype(np.uint32)
    check_pair(ctx, a, b)


@pytest.mark.parametrize("seed", range(4))
def test_random_ratios(ctx, seed):
    rng = np.random.default_rng(seed)
    for _ in range(6):
        na = int(rng.integers(0, 300_000))
        nb = int(rng.integers(0, 300_000))
        U = int(rng.choice([na + nb + 1, 4 * (na + nb) + 7, 2**32]))
        a = datasets.uniform_sorted(min(na, U), U, int(rng.integers(1 << 30)))
        b = datasets.uniform_sorted(min(nb, U), U, int(rng.integers(1 << 30)))
        check_pair(ctx, a, b)


def test_sisi_2p20_in_2p24(ctx):
    """The CPU-runnable SISI workload: |A|=|B|=2^20 uniform from [0,2^24)."""
    coll = datasets.make_workload("sisi_2^20_in_2^24")
    ds = ctx.upload(coll)
    exp = oracle.materialize(coll.set(0), coll.set(1))
    assert 60_000 < exp.size < 71_000  # E = 65,536
    assert ds.pair_count(0, 1) == exp.size
    assert np.array_equal(ds.pair_materialize(0, 1), exp)


def test_sisi_large_oracle(ctx):
    """2^25 x 2^25 from [0,2^27): many segments per launch, oracle-checked."""
    coll = datasets.sisi_pair(1 << 25, 1 << 25, 1 << 27, 99)
    ds = ctx.upload(coll)
    exp = oracle.materialize(coll.set(0), coll.set(1))
    assert ds.pair_count(0, 1) == exp.size
    assert np.array_equal(ds.pair_materialize(0, 1), exp)


@pytest.mark.parametrize("max_run", [40, 6000])
def test_sisi_split_prediction_misses(ctx, max_run):
    """Alternating runs of A-only and B-only values of random lengths, with a
    value in both sets after every third run: A's share of the merge path
    jumps from tile to tile, so splits predicted from the previous tile miss
    (runs up to 6000, beyond the +-256 bounded search) and the full search and
    its back-off take over.  Count and materialised list against the oracle,
    also with the set roles swapped."""
    rng = np.random.default_rng(max_run)
    a, b, v = [], [], 0
    while v < 12_000_000:
        n = int(rng.integers(1, max_run))
        (a if len(a) <= len(b) else b).append(np.arange(v, v + n, dtype=np.uint64))
        v += n
        if rng.integers(3) == 0:
            a.append(np.array([v], np.uint64))
            b.append(np.array([v], np.uint64))
            v += 1
    a = np.concatenate(a).astype(np.uint32)
    b = np.concatenate(b).astype(np.uint32)
    coll = si.Collection.from_sets([a, b])
    ds = ctx.upload(coll)
    exp = oracle.materialize(a, b)
    assert exp.size > 1000
    for x, y in ((0, 1), (1, 0)):
        assert ds.pair_count(x, y) == exp.size
        assert np.array_equal(ds.pair_materialize(x, y), exp)
    ds.free()


def test_async_buffers_and_repeat(ctx):
    coll = datasets.sisi_pair(300_000, 200_000, 1 << 21, 5)
    ds = ctx.upload(coll)
    exp = oracle.materialize(coll.set(0), coll.set(1))
    out = ctx.buffer(8 * 4)
    for i in range(4):  # repeated launches reuse the self-clearing ticket counter
        ds.pair_count_async(0, 1, out, offset=8 * i)
    ctx.sync()
    assert out.download(np.uint64).tolist() == [exp.size] * 4
    lst = ctx.buffer(200_000 * 4)
    cnt = ctx.buffer(8)
    ds.pair_materialize_async(0, 1, lst, cnt)
    n = int(cnt.download(np.uint64)[0])
    assert n == exp.size
    assert np.array_equal(lst.download(np.uint32, count=n), exp)


def test_errors(ctx):
    coll = si.Collection.from_sets([[1, 2, 3], [2, 3, 4]])
    ds = ctx.upload(coll)
    with pytest.raises(si.SIError):
        ds.pair_count(0, 5)
    with pytest.raises(si.SIError) as e:
        ctx.upload(si.Collection.from_sets([[1, 3, 2]]))
    assert e.value.status == -5  # SI_ERR_UNSORTED
    with pytest.raises(si.SIError):
        ctx.upload(si.Collection.from_sets([[1, 1, 2]]))


def test_materialise_unvalidated(ctx):
    """A dataset uploaded without validation is checked by its first
    materialisation: a run of duplicates (which would match more often than
    min(|a|, |b|)) is refused with SI_ERR_UNSORTED; valid sets materialise."""
    dup = si.Collection.from_sets([np.repeat(np.arange(0, 5000, dtype=np.uint32), 3),
                                   np.repeat(np.arange(0, 5000, dtype=np.uint32), 2)])
    ds = ctx.upload(dup, validate=False)
    with pytest.raises(si.SIError) as e:
        ds.pair_materialize(0, 1)
    assert e.value.status == -5
    good = datasets.sisi_pair(100_000, 80_000, 1 << 20, 3)
    ds2 = ctx.upload(good, validate=False)
    assert np.array_equal(ds2.pair_materialize(0, 1), oracle.materialize(good.set(0), good.set(1)))


def _skewed_pairs():
    """(large, small) pairs 32x to 1024x apart: the materialisation searches
    the small set's elements in the large one.  Small-set elements below the
    large set's first, above its last, equal to both ends, a subset (every
    probe matches), a disjoint set (none does), and probe sets spanning
    several chunks per workgroup."""
    rng = np.random.default_rng(17)
    big = np.unique(rng.integers(1000, 2**32 - 1000, 1 << 22, dtype=np.uint64)).astype(np.uint32)
    yield big, np.unique(np.concatenate([rng.choice(big, 3000), rng.integers(0, 2**32, 3000, dtype=np.uint64),
                                         [0, 999, big[0], big[-1], 2**32 - 1]]).astype(np.uint32))
    yield big, big[::33].copy()                                                  # every probe matches
    yield big, np.setdiff1d(big[::40].astype(np.uint64) + 1, big).astype(np.uint32)  # no probe matches
    yield big, np.sort(rng.choice(big, 99_997, replace=False))                     # a ragged last chunk
    huge = np.unique(rng.integers(0, 2**32, 1 << 25, dtype=np.uint64)).astype(np.uint32)
    yield huge, np.unique(np.concatenate([rng.choice(huge, 1 << 19), rng.integers(0, 2**32, 1 << 19,
                                                                                  dtype=np.uint64)])).astype(np.uint32)


def test_materialise_skewed(ctx, monkeypatch):
    """Pairs whose sizes are at least 32x apart materialise by search
    (k_obs_mat): the list equals the oracle's in both argument orders, twice
    in a row (the look-back states and tickets reset themselves), and equals
    the merge path's (forced with SI_SISI_SHAPE)."""
    monkeypatch.setenv("SI_SISI_SHAPE", "1")
    merge_ctx = si.Context(0)
    try:
        for big, small in _skewed_pairs():
            assert big.size // small.size >= 32
            coll = si.Collection.from_sets([big, small])
            exp = oracle.materialize(big, small)
            ds = ctx.upload(coll)
            for a, b in ((0, 1), (1, 0), (0, 1)):
                got = ds.pair_materialize(a, b)
                assert got.size == exp.size and np.array_equal(got, exp), (big.size, small.size, a)
            ds.free()
            mds = merge_ctx.upload(coll)
            assert np.array_equal(mds.pair_materialize(1, 0), exp)
            mds.free()
    finally:
        merge_ctx.close()


# ------------------------------------------------------------------ MISI --
@pytest.mark.parametrize("name", ["closed_form", "random_k12", "random_k23"])
@pytest.mark.parametrize("partition", [3, 5, 10000])
def test_golden_runs(ctx, name, partition):
    coll = datasets.read_bin(os.path.join(GOLDEN, name + ".bin"))
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, si.TECH_IP, partition)
    counts, _ = oracle.all_pairs(coll)
    k = coll.k
    for (i, j), c in got.items():
        assert c == counts[oracle.triangular_index(k, i, j)], (i, j)
    assert len(got) == k * (k - 1) // 2


def test_runs_random_collection(ctx):
    coll = datasets.zipf_collection(300, 1.2, 4, 5000, 1 << 16, 3)
    ds = ctx.upload(coll)
    counts, _ = oracle.all_pairs(coll)
    got = run_layout_counts(ds, coll, si.TECH_IP, 128)
    k = coll.k
    for (i, j), c in got.items():
        assert c == counts[oracle.triangular_index(k, i, j)]


def _adversarial_pairs():
    """Pairs whose merge path runs far ahead on one side (tiles cut short),
    ends one set mid-tile, or matches on every step."""
    r = np.random.default_rng(5)
    runs_a, runs_b, v = [], [], 0
    for i in range(60):
        la, lb = int(r.integers(1, 9000)), int(r.integers(1, 9000))
        runs_a.append(np.arange(v, v + la))
        v += la
        runs_b.append(np.arange(v - (i % 3), v - (i % 3) + lb))  # seams share 0-2 values
        v += lb
    yield np.unique(np.concatenate(runs_a)).astype(np.uint32), np.unique(np.concatenate(runs_b)).astype(np.uint32)
    big = datasets.uniform_sorted(1_000_000, 1 << 32, 3)
    yield big, big[::997].copy()
    yield big, big
    yield big[:500_000].copy(), big[250_000:].copy()
    yield big[::2].copy(), big[1::2].copy()
    yield np.arange(0, 3_000_000, 3, dtype=np.uint32), np.arange(0, 3_000_000, 5, dtype=np.uint32)
    # runs of shared values (up to one match per two merge steps) alternate
    # with runs of interleaved disjoint values at random lengths: in one tile
    # some waves hold few matches and others more than a wave's share of the
    # tile's match buffer
    ra, rb, v = [], [], 0
    while v < 3_000_000:
        n = int(r.integers(50, 3000))
        if r.integers(2):
            ra.append(np.arange(v, v + n))
            rb.append(np.arange(v, v + n))
        else:
            ra.append(np.arange(v, v + 2 * n, 2))
            rb.append(np.arange(v + 1, v + 2 * n, 2))
        v += 2 * n
    yield np.concatenate(ra).astype(np.uint32), np.concatenate(rb).astype(np.uint32)


SISI_SHAPES = 7  # tile shapes of the SISI merge kernel (si_params.reserved[0]; 0 = automatic)


@pytest.mark.parametrize("shape", range(SISI_SHAPES))
def test_tile_shapes(shape, monkeypatch):
    """Every SISI tile shape (count: si_params.reserved[0]; materialise:
    SI_SISI_SHAPE at context open) is exact on adversarial pairs."""
    monkeypatch.setenv("SI_SISI_SHAPE", str(shape))
    c = si.Context(0)
    try:
        p = si.params()
        p.reserved[0] = shape
        for a, b in _adversarial_pairs():
            coll = si.Collection.from_sets([a, b])
            ds = c.upload(coll)
            exp = oracle.materialize(coll.set(0), coll.set(1))
            assert ds.pair_count(0, 1, p=p) == exp.size, (shape, a.size, b.size)
            assert ds.pair_count(1, 0, p=p) == exp.size, (shape, a.size, b.size)
            for blocks in (1, 7, 300):
                q = si.params(blocks=blocks)
                q.reserved[0] = shape
                assert ds.pair_count(0, 1, p=q) == exp.size, (shape, blocks)
            assert np.array_equal(ds.pair_materialize(0, 1), exp), (shape, a.size, b.size)
            assert np.array_equal(ds.pair_materialize(1, 0), exp), (shape, a.size, b.size)
            ds.free()
    finally:
        c.close()


@pytest.mark.parametrize("prio", [0, 1])
def test_priority_rotation(prio, monkeypatch):
    """The SISI merge, obs, hi and bitmap_dynamic kernels rotate their waves'
    issue priority per tile, probe round or bucket (SI_SISI_PRIO, read at
    context open; on by default): counts and lists are exact either way."""
    monkeypatch.setenv("SI_SISI_PRIO", str(prio))
    c = si.Context(0)
    try:
        pairs = list(_adversarial_pairs())
        pairs.append((datasets.uniform_sorted(3_000_000, 1 << 26, 5), datasets.uniform_sorted(2_500_000, 1 << 26, 6)))
        for a, b in pairs:
            coll = si.Collection.from_sets([a, b])
            ds = c.upload(coll)
            exp = oracle.materialize(coll.set(0), coll.set(1))
            for t in (si.TECH_IP, si.TECH_OBS, si.TECH_HI, si.TECH_BITMAP_DYNAMIC):
                assert ds.pair_count(0, 1, technique=t) == exp.size, (prio, t, a.size, b.size)
                assert ds.pair_count(1, 0, technique=t) == exp.size, (prio, t, a.size, b.size)
            assert np.array_equal(ds.pair_materialize(0, 1), exp), (prio, a.size, b.size)
            ds.free()
    finally:
        c.close()


@pytest.mark.parametrize("ramp", [0, 4, 100])
def test_segment_ramp(ramp, monkeypatch):
    """Materialising segments whose lengths rise linearly with their ticket
    (SI_SISI_RAMP / 256 of the mean either side; 4 by default, 100 a steep
    ramp): the lists are exact for every tile shape's staging capacity."""
    monkeypatch.setenv("SI_SISI_RAMP", str(ramp))
    c = si.Context(0)
    try:
        pairs = list(_adversarial_pairs())
        pairs.append((datasets.uniform_sorted(3_000_000, 1 << 26, 7), datasets.uniform_sorted(2_500_000, 1 << 26, 8)))
        pairs.append((np.arange(0, 4_000_000, 2, dtype=np.uint32), np.arange(0, 4_000_000, 2, dtype=np.uint32)))
        for a, b in pairs:
            coll = si.Collection.from_sets([a, b])
            ds = c.upload(coll)
            exp = oracle.materialize(coll.set(0), coll.set(1))
            assert np.array_equal(ds.pair_materialize(0, 1), exp), (ramp, a.size, b.size)
            assert np.array_equal(ds.pair_materialize(1, 0), exp), (ramp, a.size, b.size)
            ds.free()
    finally:
        c.close()


def test_sisi_full_size():
    """The headline pair at full size: |A| = |B| = 2^28 from [0, 2^32).
    Exact against the oracle's std::set_intersection; the materialised list is
    strictly increasing and every technique's count agrees."""
    coll = datasets.sisi_pair(1 << 28, 1 << 28, 1 << 32, 20261203)
    exp = oracle.materialize(coll.set(0), coll.set(1))
    c = si.Context(0)
    try:
        ds = c.upload(coll)
        got = ds.pair_materialize(0, 1)
        assert got.size == exp.size and np.array_equal(got, exp)
        assert np.all(np.diff(got.astype(np.int64)) > 0)
        for t in (si.TECH_IP, si.TECH_OBS, si.TECH_HI, si.TECH_BITMAP_DYNAMIC):
            assert ds.pair_count(0, 1, technique=t) == exp.size, t
            assert ds.pair_count(1, 0, technique=t) == exp.size, t
    finally:
        c.close()
