"""GPU parity of the obs, hi and bitmap kernels against the CPU oracle.

Every count must equal the oracle's bit for bit: SISI pairs (edge cases,
closed forms, skewed size ratios), MISI runs through the reference's
multi-run layout for block- and wave-level variants, and bitmap_naive's
all-pairs counts.
"""
import os

import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets
from helpers import check_all_pairs, edge_case_pairs, run_layout_counts

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SISI_TECHS = [si.TECH_IP, si.TECH_OBS, si.TECH_HI, si.TECH_BITMAP_DYNAMIC]
RUN_TECHS = [si.TECH_IP, si.TECH_OBS, si.TECH_HI, si.TECH_BITMAP_DYNAMIC]


PATH = si.params(path=True)


def check_pair(ctx, a, b, techs=SISI_TECHS):
    """Every technique in both role orders; obs and hi also with --path
    (merge-path segments, k_path.hip) at the default geometry and with a
    single workgroup (one range walked segment by segment)."""
    coll = si.Collection.from_sets([a, b])
    ds = ctx.upload(coll)
    exp = oracle.count(coll.set(0), coll.set(1))
    for t in techs:
        assert ds.pair_count(0, 1, t) == exp, t
        assert ds.pair_count(1, 0, t) == exp, t
    for t in (si.TECH_OBS, si.TECH_HI):
        for p in (PATH, si.params(path=True, blocks=1), si.params(path=True, blocks=7)):
            assert ds.pair_count(0, 1, t, p) == exp, (t, p.blocks, "path")
            assert ds.pair_count(1, 0, t, p) == exp, (t, p.blocks, "path")
    ds.free()


def test_sisi_edge_cases(ctx):
    for a, b in edge_case_pairs():
        check_pair(ctx, a, b)


@pytest.mark.parametrize("na,nb,U", [(1 << 16, 1 << 16, 1 << 18), (1 << 20, 1 << 10, 1 << 24),
                                     (3, 1 << 20, 1 << 22), (8191, 8193, 20000), (100_000, 99_999, 2**32)])
def test_sisi_sizes(ctx, na, nb, U):
    a = datasets.uniform_sorted(na, U, na + 1)
    b = datasets.uniform_sorted(nb, U, nb + 2)
    if U == 2**32:
        b = np.union1d(b, a[::97]).astype(np.uint32)
    check_pair(ctx, a, b)


def test_sisi_dense_and_clustered(ctx):
    """Dense value ranges (buckets filled to the brim), clusters separated by
    empty stretches of the universe, and values at both ends of u32."""
    ev = np.arange(0, 1 << 21, 2, dtype=np.uint32)
    th = np.arange(0, 1 << 21, 3, dtype=np.uint32)
    check_pair(ctx, ev, th)
    top = np.uint64(2**32 - 1)
    a = np.unique(np.concatenate([np.arange(0, 5000), np.arange(2**31, 2**31 + 70000, 7),
                                  np.arange(int(top) - 90000, int(top) + 1)])).astype(np.uint32)
    b = np.unique(np.concatenate([np.arange(0, 5000, 2), np.arange(2**31, 2**31 + 70000, 5),
                                  np.arange(int(top) - 40000, int(top) + 1, 3)])).astype(np.uint32)
    check_pair(ctx, a, b)


def test_sisi_bitmap_few_blocks(ctx):
    """bitmap_dynamic with --blocks 1 on a dense pair: the bucket bounds do not
    fit one workgroup's LDS, so the launch takes more workgroups."""
    a = np.arange(0, 1 << 25, 2, dtype=np.uint32)  # 2^24 elements: 4096 buckets of 2^13 values
    b = np.arange(0, 1 << 25, 3, dtype=np.uint32)
    coll = si.Collection.from_sets([a, b])
    ds = ctx.upload(coll)
    exp = oracle.count(a, b)
    assert ds.pair_count(0, 1, si.TECH_BITMAP_DYNAMIC, si.params(blocks=1)) == exp
    ds.free()


def test_sisi_skewed_2p24x2p12(ctx):
    """A scaled-down copy of the |A|=2^27, |B|=2^16 size ratio."""
    coll = datasets.sisi_pair(1 << 24, 1 << 13, 2**32, 17)
    a, b = coll.set(0), coll.set(1)
    b = np.union1d(b, a[::4096]).astype(np.uint32)
    check_pair(ctx, a, b)


# --warp and --path select kernels only for obs and hi
@pytest.mark.parametrize("tech,warp,path", [(t, False, False) for t in RUN_TECHS] +
                         [(t, w, p) for t in (si.TECH_OBS, si.TECH_HI) for w, p in ((True, False), (False, True))])
@pytest.mark.parametrize("name,partition", [("closed_form", 10000), ("closed_form", 3), ("random_k12", 5),
                                            ("random_k23", 10000), ("random_k23", 4)])
def test_golden_runs(ctx, tech, warp, path, name, partition):
    coll = datasets.read_bin(os.path.join(GOLDEN, name + ".bin"))
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, tech, partition, si.params(warp=warp, path=path))
    check_all_pairs(got, coll)


@pytest.mark.parametrize("tech", [si.TECH_OBS, si.TECH_HI])
def test_path_runs_long_pairs(ctx, tech):
    """Runs with --path on pairs longer than one segment (hash: 2048 merged
    elements per segment, obs: 512), across tiles (partition 40 of 90 sets),
    with dense value ranges so that matches fall on segment seams."""
    sizes = [70000, 9000, 4097, 2049, 2048, 513, 100, 3, 1, 0] * 9
    coll = datasets.sized_collection(sizes, 1 << 17, 12)
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, tech, 40, si.params(path=True))
    check_all_pairs(got, coll)


@pytest.mark.parametrize("tech", RUN_TECHS)
def test_runs_zipf_collection(ctx, tech):
    coll = datasets.zipf_collection(400, 1.2, 4, 20000, 1 << 20, 9)
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, tech, 150)
    check_all_pairs(got, coll)


@pytest.mark.parametrize("tech", [si.TECH_OBS, si.TECH_HI])
def test_runs_warp_zipf(ctx, tech):
    coll = datasets.zipf_collection(200, 1.1, 1, 70000, 1 << 20, 5)  # rows beyond the LDS table size
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, tech, 10000, si.params(warp=True))
    check_all_pairs(got, coll)


def test_bitmap_run_global_universe(ctx):
    """Universe beyond the LDS bitmap (2^22 bits) takes the global-memory path."""
    coll = datasets.zipf_collection(120, 1.2, 8, 3000, 1 << 22, 4)
    ds = ctx.upload(coll)
    got = run_layout_counts(ds, coll, si.TECH_BITMAP_DYNAMIC, 10000)
    check_all_pairs(got, coll)


@pytest.mark.parametrize("name", ["closed_form_small", "random_k12", "random_k23"])
def test_bitmap_naive_all_pairs(ctx, name):
    if name == "closed_form_small":
        coll = si.Collection.from_sets([np.arange(0, 1000, p, dtype=np.uint32) for p in (2, 3, 5, 7, 11)] + [[]])
    else:
        coll = datasets.read_bin(os.path.join(GOLDEN, name + ".bin"))
    ds = ctx.upload(coll)
    got = ds.all_pairs_bitmap()
    counts, _ = oracle.all_pairs(coll)
    assert np.array_equal(got, counts)


def test_bitmap_naive_multi_tile(ctx):
    coll = datasets.density_collection(300, 1 << 14, 0.01, 0.2, 3)  # 5 x 5 tiles of 64 pairs
    ds = ctx.upload(coll)
    got = ds.all_pairs_bitmap()
    counts, _ = oracle.all_pairs(coll)
    assert np.array_equal(got, counts)
