"""Writes the golden fixtures under tests/golden/.

The reference ships no tests, fixtures or golden outputs, so these vectors are
created here with values fixed independently of the oracle and of the GPU
code: closed forms (|multiples of p ∩ multiples of q below N| =
ceil(N / lcm(p, q))) checked against brute-force Python set intersection.

Files (all small, committed):
  <name>.bin        dataset in the reference .bin layout (include/io.hpp:25-54)
  <name>.txt        expected single-run result file, "(a+1,b+1): count" for all
                    a < b in row-major order (include/io.hpp:91-102)
  <name>_p<P>.txt   expected multi-run result file for --partition P
                    (include/io.hpp:104-136 with findTilePairs, util.hpp:40-52)
The mm executable's writer (io.hpp:104-136 with mm=true) differs only in
where it reads each count, so its expected text is the same <name>_p<P>.txt.
Run:  python tests/golden/make_fixtures.py
"""
import math
import os
import random
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def write_bin(path, sets, universe):
    with open(path, "wb") as f:
        f.write(struct.pack("<IIQ", len(sets), universe, sum(len(s) for s in sets)))
        f.write(struct.pack(f"<{len(sets)}I", *[len(s) for s in sets]))
        for s in sets:
            f.write(struct.pack(f"<{len(s)}I", *s))


def brute(a, b):
    return len(set(a) & set(b))


def single_run_lines(sets):
    k = len(sets)
    return [f"({a + 1},{b + 1}): {brute(sets[a], sets[b])}\n" for a in range(k) for b in range(a + 1, k)]


def multi_run_lines(sets, partition):
    # runs = tile pairs (i <= j) of tiles [t*P, min((t+1)*P, k)); inside a run
    # i walks tile A, j walks tile B (j > i in a self-join run)
    k = len(sets)
    ntiles = math.ceil(k / partition)
    tiles = [(t, t * partition, min((t + 1) * partition, k)) for t in range(ntiles)]
    lines = []
    for x in range(ntiles):
        for y in range(x, ntiles):
            A, B = tiles[x], tiles[y]
            self_join = A[0] == B[0]
            for i in range(A[1], A[2]):
                for j in range(i + 1 if self_join else B[1], B[2]):
                    lines.append(f"({i + 1},{j + 1}): {brute(sets[i], sets[j])}\n")
    return lines


def closed_form_sets():
    N = 1000
    mult = {p: list(range(0, N, p)) for p in (2, 3, 5, 7)}
    # closed forms pin the brute force itself
    for p in mult:
        for q in mult:
            lcm = p * q // math.gcd(p, q)
            assert brute(mult[p], mult[q]) == math.ceil(N / lcm)
    sets = [mult[2], mult[3], mult[5], mult[7], list(mult[3]),  # identical copy of set 2
            [],                                                     # empty set
            list(range(1, N, 2)),                                   # odds: disjoint from evens
            [0, 1, 2**31 - 1, 2**31, 2**31 + 1, 2**32 - 2, 2**32 - 1],  # signedness / top of range
            [2**31, 2**32 - 1],
            list(range(990, 1000)) + [2**31 + 1, 2**32 - 1]]
    return sets, 2**32 - 1


def random_sets(seed, k, max_size, universe):
    rng = random.Random(seed)
    sets = []
    for i in range(k):
        n = rng.choice([0, 1, 2, 3, 63, 64, 65, rng.randrange(0, max_size + 1)])
        n = min(n, universe)
        sets.append(sorted(rng.sample(range(universe), n)))
    return sets


def main():
    fixtures = {
        "closed_form": closed_form_sets(),
        "random_k12": (random_sets(7, 12, 200, 500), 499),
        "random_k23": (random_sets(11, 23, 300, 4096), 4095),
    }
    for name, (sets, universe) in fixtures.items():
        write_bin(os.path.join(HERE, f"{name}.bin"), sets, universe)
        with open(os.path.join(HERE, f"{name}.txt"), "w") as f:
            f.writelines(single_run_lines(sets))
        for P in (3, 5):
            with open(os.path.join(HERE, f"{name}_p{P}.txt"), "w") as f:
                f.writelines(multi_run_lines(sets, P))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
