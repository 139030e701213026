"""The allocation-free SISI count (si_pair_count_noalloc) against
si_pair_count_async, through the C ABI: every technique (obs and hi also
with --path), two calls in a row with the second on a smaller pair, each
count equal to the oracle's; and the SI_ERR_RANGE refusal of a launch whose
scratch would not fit the static array.  Uses a context of its own, so it
passes alone in a fresh process and after the other tests."""
import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets

pytestmark = pytest.mark.gpu

VARIANTS = [(si.TECH_IP, None), (si.TECH_OBS, None), (si.TECH_HI, None), (si.TECH_BITMAP_DYNAMIC, None),
            (si.TECH_OBS, si.params(path=True)), (si.TECH_HI, si.params(path=True)),
            (si.TECH_IP, si.params(blocks=7))]


def test_noalloc_matches_async():
    big = datasets.sisi_pair(1 << 22, 3 << 20, 1 << 26, 31)
    small = datasets.sisi_pair(50_000, 70_000, 1 << 18, 32)
    coll = si.Collection.from_sets([big.set(0), big.set(1), small.set(0), small.set(1)])
    exp = {(0, 1): oracle.count(coll.set(0), coll.set(1)), (2, 3): oracle.count(coll.set(2), coll.set(3))}
    assert min(exp.values()) > 0
    c = si.Context(0)
    try:
        ds = c.upload(coll)
        ref = c.buffer(64)
        got = c.buffer(64)
        for tech, p in VARIANTS:
            # two calls in a row on the shared scratch, the second on a smaller pair
            for k, (a, b) in enumerate([(0, 1), (2, 3)]):
                ds.pair_count_async(a, b, ref, offset=8 * k, technique=tech, p=p)
                ds.pair_count_noalloc(a, b, got, offset=8 * k, technique=tech, p=p)
            r = ref.download(np.uint64, count=2).tolist()
            g = got.download(np.uint64, count=2).tolist()
            assert g == r == [exp[(0, 1)], exp[(2, 3)]], (tech, p and p.blocks, g, r)
        # an empty set: the output is written as zero
        e = c.upload(si.Collection.from_sets([np.zeros(0, np.uint32), np.arange(5, dtype=np.uint32)]))
        got.upload(np.full(8, 7, np.uint64))
        e.pair_count_noalloc(0, 1, got, offset=0)
        assert int(got.download(np.uint64, count=1)[0]) == 0
    finally:
        c.close()


def test_noalloc_scratch_limit():
    """obs on a 2^24-element probe side with --blocks 2^20 would launch 16384
    workgroups: 128 KiB of partials, beyond the 64 KiB array.  The call is
    refused with SI_ERR_RANGE; the same pair with the default geometry counts
    exactly afterwards."""
    coll = datasets.sisi_pair(1 << 24, 1 << 24, 1 << 27, 33)
    exp = oracle.count(coll.set(0), coll.set(1))
    c = si.Context(0)
    try:
        ds = c.upload(coll)
        out = c.buffer(8)
        with pytest.raises(si.SIError) as err:
            ds.pair_count_noalloc(0, 1, out, technique=si.TECH_OBS, p=si.params(blocks=1 << 20))
        assert err.value.status == -4  # SI_ERR_RANGE
        for tech in (si.TECH_OBS, si.TECH_IP):
            ds.pair_count_noalloc(0, 1, out, technique=tech)
            assert int(out.download(np.uint64)[0]) == exp
    finally:
        c.close()
