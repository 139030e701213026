"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it.

Checks use facts fixed independently of the oracle's code: closed forms,
brute-force Python sets, the committed golden fixtures, and agreement of the
four restated methods (std::set_intersection, serial merge, bitset probe,
SIMD block compare) which share no code.
"""
import math
import os
import random

import numpy as np
import pytest

import oracle
from set_intersection_amd import datasets

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
METHODS = ["std", "serial", "bitset", "simd"]


def parse_result(path):
    out = []
    with open(path) as f:
        for line in f:
            lhs, c = line.split(": ")
            i, j = lhs.strip("()").split(",")
            out.append((int(i), int(j), int(c)))
    return out


@pytest.mark.parametrize("method", METHODS)
def test_closed_form_multiples(method):
    N = 100_003
    for p, q in [(2, 3), (4, 6), (7, 7), (1, 5), (12, 18), (97, 89)]:
        a = np.arange(0, N, p, dtype=np.uint32)
        b = np.arange(0, N, q, dtype=np.uint32)
        lcm = p * q // math.gcd(p, q)
        assert oracle.count(a, b, method, universe=N) == math.ceil(N / lcm)


@pytest.mark.parametrize("method", METHODS)
def test_special_cases(method):
    a = np.array([0, 5, 9, 2**31 - 1, 2**31, 2**32 - 1], dtype=np.uint32)
    empty = np.zeros(0, dtype=np.uint32)
    disjoint = np.array([1, 6, 10, 2**31 + 1], dtype=np.uint32)
    assert oracle.count(a, a, method, universe=2**32 - 1) == a.size
    assert oracle.count(a, empty, method, universe=2**32 - 1) == 0
    assert oracle.count(empty, a, method, universe=2**32 - 1) == 0
    assert oracle.count(a, disjoint, method, universe=2**32 - 1) == 0
    sub = a[[1, 3, 5]]
    assert oracle.count(sub, a, method, universe=2**32 - 1) == 3
    assert oracle.count(a, sub, method, universe=2**32 - 1) == 3


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_random(seed):
    rng = random.Random(seed)
    for it in range(60):
        U = rng.choice([10, 100, 5000, 2**20, 2**32])
        na, nb = rng.randrange(0, 300), rng.randrange(0, 300)
        na, nb = min(na, U), min(nb, U)
        a = sorted(rng.sample(range(U), na)) if U < 10**6 else sorted(set(rng.randrange(U) for _ in range(na)))
        b = sorted(rng.sample(range(U), nb)) if U < 10**6 else sorted(set(rng.randrange(U) for _ in range(nb)))
        if U == 2**32 and a and b:  # force some overlap on the sparse universe
            b = sorted(set(b) | set(a[::3]))
        exp = sorted(set(a) & set(b))
        for m in METHODS:
            if m == "bitset" and U == 2**32 and it > 5:
                continue  # a 2^32-bit bitset per call is slow; covered by the first iterations
            assert oracle.count(a, b, m, universe=U - 1) == len(exp), m
        assert oracle.materialize(a, b).tolist() == exp


def test_simd_skewed_path():
    # |b|/|a| >= 50 takes the V1 (one-element-vs-block) path
    rng = np.random.default_rng(1)
    big = np.unique(rng.integers(0, 1 << 24, 200_000, dtype=np.uint64)).astype(np.uint32)
    small = np.sort(rng.choice(big, 1000, replace=False))
    small = np.unique(np.concatenate([small, np.array([3, 7, 11], dtype=np.uint32)]))
    exp = len(set(small.tolist()) & set(big.tolist()))
    assert oracle.count(small, big, "simd") == exp
    assert oracle.count(big, small, "simd") == exp


def test_triangular_index_enumeration():
    # row-major strict upper triangle, util.hpp:68-70
    for n in (2, 3, 7, 50, 201):
        idx = 0
        for i in range(n):
            for j in range(i + 1, n):
                assert oracle.triangular_index(n, i, j) == idx
                idx += 1


@pytest.mark.parametrize("name", ["closed_form", "random_k12", "random_k23"])
@pytest.mark.parametrize("method", METHODS)
def test_golden_all_pairs(name, method):
    coll = datasets.read_bin(os.path.join(GOLDEN, name + ".bin"))
    counts, _ = oracle.all_pairs(coll, method, threads=3)
    expected = parse_result(os.path.join(GOLDEN, name + ".txt"))
    k = coll.k
    got = [(a + 1, b + 1, int(counts[oracle.triangular_index(k, a, b)])) for a in range(k) for b in range(a + 1, k)]
    assert got == expected


def test_edges_and_pairs_agree_with_all_pairs():
    coll = datasets.read_bin(os.path.join(GOLDEN, "random_k23.bin"))
    counts, _ = oracle.all_pairs(coll, "std", threads=2)
    k = coll.k
    pa = np.array([a for a in range(k) for b in range(a + 1, k)], dtype=np.uint32)
    pb = np.array([b for a in range(k) for b in range(a + 1, k)], dtype=np.uint32)
    got, _ = oracle.pairs(coll, pa, pb, "serial", threads=2)
    assert np.array_equal(got, counts)
    ec, tot = oracle.edges(coll, pa, pb, threads=2)
    assert np.array_equal(ec, counts) and tot == int(counts.sum())
    M = oracle.tile_matrix(coll, 0, k, 0, k)
    for a in range(k):
        assert M[a, a] == coll.sizes[a]
        for b in range(a + 1, k):
            assert M[a, b] == M[b, a] == counts[oracle.triangular_index(k, a, b)]


def test_planted_mistake_is_caught():
    """A deliberately wrong count (a dropped element, as an off-by-one merge
    would produce) must disagree with the closed form."""
    a = np.arange(0, 1000, 2, dtype=np.uint32)
    b = np.arange(0, 1000, 3, dtype=np.uint32)
    truth = math.ceil(1000 / 6)
    wrong = oracle.count(a[a != 996], b)  # drops one common element (996 = 166 * 6)
    assert oracle.count(a, b) == truth
    assert wrong != truth
