"""GPU parity at the full size of every benchmark workload, bit-exact against
the CPU oracle:

  * SISI 2^27 x 2^16 from [0, 2^32): ip / obs / hi / bitmap_dynamic counts in
    both role orders and the materialised list (ip);
  * MISI 16384 sets, Zipf(1.2) sizes in [16, 65536], elements in [0, 2^20):
    the batched dispatcher (default and autotuned tuning, hipGraph on and off)
    against the oracle's 134,209,536 pair counts, and the plain per-run
    kernels (ip, bitmap_dynamic) through the reference's multi-run layout at
    partition 10000 (3 runs; ip.cu:141-177);
  * bitmap GEMM, 8192 sets over 2^16, density U[1%, 10%]: fp4, int8 and the
    fp8 FP32-accumulate variant against the AND/popcount kernel on the whole
    8192 x 8192 matrix and against the oracle on row blocks spread over it
    (matrix_multiplication.cu:139-184);
  * R-MAT scale 20, edge factor 16: per-edge counts and the total against the
    oracle on every edge, the total a multiple of 3, for the binned dispatcher
    and the plain per-edge kernel.
"""
import numpy as np
import pytest

import oracle
import set_intersection_amd as si
from set_intersection_amd import datasets
from helpers import run_layout_triangular

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- SISI skew --
@pytest.fixture(scope="module")
def skew():
    coll = datasets.make_workload("sisi_2^27x2^16_in_2^32")
    exp = oracle.materialize(coll.set(0), coll.set(1))
    return coll, exp


@pytest.mark.parametrize("tech", [si.TECH_IP, si.TECH_OBS, si.TECH_HI, si.TECH_BITMAP_DYNAMIC])
def test_c2b_counts(ctx, skew, tech):
    coll, exp = skew
    assert 1500 < exp.size < 2600  # E|A∩B| = 2^27 * 2^16 / 2^32 = 2048
    ds = ctx.upload(coll)
    try:
        assert ds.pair_count(0, 1, tech) == exp.size
        assert ds.pair_count(1, 0, tech) == exp.size
    finally:
        ds.free()


def test_c2b_materialise(ctx, skew):
    coll, exp = skew
    ds = ctx.upload(coll)
    try:
        assert np.array_equal(ds.pair_materialize(0, 1), exp)
        assert np.array_equal(ds.pair_materialize(1, 0), exp)
    finally:
        ds.free()


# --------------------------------------------------------------------- MISI --
@pytest.fixture(scope="module")
def misi():
    coll = datasets.make_workload("misi_zipf1.2_16384")
    exp, _ = oracle.all_pairs(coll)
    assert exp.size == 134_209_536
    return coll, exp


@pytest.fixture(scope="module")
def misi_ds(misi):
    c = si.Context(0)
    ds = c.upload(misi[0])
    yield c, ds
    c.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_c3_batched_default(misi, misi_ds, graph):
    _, exp = misi
    c, ds = misi_ds
    t = ds.default_tuning()
    t.use_graph = graph
    out = c.buffer(exp.size * 4)
    for _ in range(2):  # the second launch replays the plan (and the graph)
        out.zero()
        ds.batched_count_async(out, t)
        assert np.array_equal(out.download(np.uint32), exp)


def test_c3_batched_autotuned(misi, misi_ds):
    _, exp = misi
    _, ds = misi_ds
    t = ds.autotune(1500.0)
    assert t.version == si._abi.SI_TUNING_VERSION
    for graph in (0, 1):
        t.use_graph = graph
        assert np.array_equal(ds.batched_count(t), exp), graph


@pytest.mark.parametrize("tech", [si.TECH_IP, si.TECH_BITMAP_DYNAMIC])
def test_c3_plain_runs(misi, misi_ds, tech):
    coll, exp = misi
    _, ds = misi_ds
    got = run_layout_triangular(ds, coll, tech, 10000)
    assert np.array_equal(got, exp)


# ----------------------------------------------------------------------- mm --
@pytest.fixture(scope="module")
def mm():
    coll = datasets.make_workload("mm_8192x2^16_density1-10")
    c = si.Context(0)
    ds = c.upload(coll)
    A = si.tile(0, 0, coll.k)
    ref = ds.mm_counts(A, A, si.MM_POPCOUNT)
    yield coll, c, ds, ref
    c.close()


def test_c4_popcount_rows(mm):
    """The popcount matrix (the reference for the MFMA variants) against the
    oracle on 4 blocks of 16 rows, first, last and two inside, all columns;
    and symmetric, with each set's size on the diagonal."""
    coll, _, _, ref = mm
    k = coll.k
    for r0 in (0, 2048 + 5, 5000, k - 16):
        assert np.array_equal(ref[r0:r0 + 16], oracle.tile_matrix(coll, r0, r0 + 16, 0, k)), r0
    assert np.array_equal(ref, ref.T)
    assert np.array_equal(np.diag(ref), coll.sizes)


@pytest.mark.parametrize("variant", [si.MM_FP4, si.MM_I8, si.MM_F32])
def test_c4_variants(mm, variant):
    coll, _, ds, ref = mm
    A = si.tile(0, 0, coll.k)
    got = ds.mm_counts(A, A, variant)
    mism = int(np.count_nonzero(got != ref))
    assert mism == 0, (variant, mism)


# -------------------------------------------------------------------- edges --
@pytest.fixture(scope="module")
def rmat():
    coll = datasets.make_workload("rmat20_ef16")
    u, v = datasets.edge_list(coll)
    exp, etot = oracle.edges(coll, u, v)
    return coll, u, v, exp, etot


def test_c5_edges(rmat):
    coll, u, v, exp, etot = rmat
    assert etot % 3 == 0
    c = si.Context(0)
    try:
        ds = c.upload(coll)
        plan = ds.edge_plan(u, v)
        counts = c.buffer(4 * u.size)
        total = c.buffer(8)
        for _ in range(2):
            counts.zero()
            plan.count_async(counts, total)
            assert np.array_equal(counts.download(np.uint32), exp)
            assert int(total.download(np.uint64)[0]) == etot
        counts.zero()
        plan.count_plain_async(counts, total)
        assert np.array_equal(counts.download(np.uint32), exp)
        assert int(total.download(np.uint64)[0]) == etot
    finally:
        c.close()
