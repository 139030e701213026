"""The asynchronous transfer calls behind the executables' multi-run pipeline
(si_host_alloc, si_buffer_download_async, si_event_wait): runs computed into
two device buffers in turn, each copied to page-locked memory on the copy
stream while the next run computes, must give the oracle's counts.  The
executables' pipelined result files are checked against the golden files in
test_gpu_apps.py."""
import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets
from helpers import check_all_pairs

pytestmark = pytest.mark.gpu


def test_download_async_roundtrip(ctx):
    data = (np.arange(1000, dtype=np.uint32) * 2654435761).astype(np.uint32)
    buf = ctx.buffer(data.nbytes)
    buf.upload(data)
    host = ctx.host_array(1000)
    e0, e1 = buf.download_async(host)
    ctx.wait(e1)
    assert np.array_equal(host, data)
    assert ctx.elapsed_ms(e0, e1) >= 0.0
    # pageable destination and an offset
    pg = np.zeros(500, np.uint32)
    _, e = buf.download_async(pg, offset=400)
    ctx.wait(e)
    assert np.array_equal(pg, data[100:600])
    # empty copy
    _, e = buf.download_async(np.zeros(0, np.uint32))
    ctx.wait(e)
    ctx.host_free(host)
    with pytest.raises(si.SIError):
        buf.download_async(np.zeros(1001, np.uint32))


@pytest.mark.parametrize("technique", [si.TECH_IP, si.TECH_BITMAP_DYNAMIC])
def test_pipelined_runs_match_oracle(ctx, technique):
    coll = datasets.zipf_collection(600, 1.2, 16, 4096, 1 << 16, 11)
    partition = 128
    k = coll.k
    ds = ctx.upload(coll)
    tiles = [si.tile(t, t * partition, min((t + 1) * partition, k)) for t in range((k + partition - 1) // partition)]
    runs = [(tiles[x], tiles[y]) for x in range(len(tiles)) for y in range(x, len(tiles))]
    assert len(runs) == 15
    dev = [ctx.buffer(4 * partition * partition) for _ in range(2)]
    host = [ctx.host_array(partition * partition) for _ in range(2)]
    done = [None, None]
    got = {}

    def collect(it):
        A, B = runs[it]
        s = it & 1
        ctx.wait(done[s])
        c = host[s]
        self_join = A.id == B.id
        for i in range(A.start, A.end):
            for j in range(i + 1 if self_join else B.start, B.end):
                li, lj = i - A.id * partition, j - B.id * partition
                got[(i, j)] = int(c[oracle.triangular_index(partition, li, lj) if self_join else li * partition + lj])

    for it, (A, B) in enumerate(runs):
        s = it & 1
        ds.run_count_async(technique, A, B, partition, dev[s])
        slots = partition * (partition - 1) // 2 if A.id == B.id else partition * partition
        _, done[s] = dev[s].download_async(host[s][:slots])
        if it:
            collect(it - 1)
    collect(len(runs) - 1)
    check_all_pairs(got, coll)
    for h in host:
        ctx.host_free(h)
