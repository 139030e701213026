"""Host replay of the chunk ranges the SISI merge kernel's window refills
read (merge_segment_stream in set_intersection_amd/csrc/si_merge.h), checked
against the element array's allocation: total elements + 64 padding
elements (si_dataset_upload).

A side's window holds its elements from the head up to a multiple of 4; each
tile's refill reads the whole 16-byte chunks from the old window end to the
new one, and the window end is clamped at the chunk after the set's last
element (endA4 = ceil((oA + nA) / 4)).  So the last chunk any refill reads
is endA4 - 1, at most 3 elements past the set, inside the padding.

A variant measured earlier issued the next tile's refill a tile ahead as a
fixed W/4 chunks per side past the window end, before the tile's
consumption is known.  For the pair's second set, which ends the element
array, the last segment's window end reaches endB4 and the ahead range
[endB4, endB4 + W/4) runs W - 64 elements (16,128 bytes for W = 4096) past
the allocation: the memory-aperture fault that variant hit.  The replay
below shows both facts on the merge path of a real pair."""
import numpy as np

from set_intersection_amd import datasets

W = 4096          # window elements per side (StreamCfg::W of the 256 x 31 tile)
T = 256 * 31      # merged elements per tile
PAD = 64          # padding elements after the last element of the dataset


def tile_coranks(a, b, d0, d1):
    """Co-ranks of the merge path (A first on ties) at d0, d0 + T, ..., d1."""
    ds = list(range(d0, d1, T)) + [d1]
    out = []
    for d in ds:
        lo, hi = max(0, d - b.size), min(d, a.size)
        while lo < hi:  # smallest i with a[i] > b[d - 1 - i]
            mid = (lo + hi) // 2
            if a[mid] <= b[d - 1 - mid]:
                lo = mid + 1
            else:
                hi = mid
        out.append((d, lo))
    return out


def replay(a, b, nseg, ahead):
    """Largest element index + 1 any refill of the launch reads (A at
    offset 0, B right after it, so B ends the array)."""
    oA, nA, oB, nB = 0, a.size, a.size, b.size
    endA4, endB4 = (oA + nA + 3) // 4, (oB + nB + 3) // 4
    n = nA + nB
    seg = -(-n // nseg)
    top = 0
    for s in range(nseg):
        d0, d1 = s * seg, min((s + 1) * seg, n)
        if d0 >= d1:
            continue
        steps = tile_coranks(a, b, d0, d1)
        for (d, ia), (dn, ian) in zip(steps[:-1], steps[1:]):
            ib, ibn = d - ia, dn - ian
            if ahead:  # fixed W/4 chunks past the window end, per side
                xa4 = min((oA + ia + W) // 4, endA4)
                xb4 = min((oB + ib + W) // 4, endB4)
                top = max(top, 4 * (xa4 + W // 4), 4 * (xb4 + W // 4))
            else:      # [old window end, new window end), clamped at the set's end
                top = max(top, 4 * min((oA + ian + W) // 4, endA4), 4 * min((oB + ibn + W) // 4, endB4))
    return top


def test_refill_reads_stay_in_the_padding():
    a = datasets.uniform_sorted(300_001, 1 << 24, 5)
    b = datasets.uniform_sorted(299_999, 1 << 24, 6)
    total = a.size + b.size
    for nseg in (1, 7, 64):
        assert replay(a, b, nseg, ahead=False) <= total + PAD
        # the tile-ahead form overruns the padding by W - 64 elements at the array's end
        assert replay(a, b, nseg, ahead=True) - (total + PAD) >= W - PAD - 3
