// Deterministic synthetic dataset generators.
//
// The reference's generator (src/generate_dataset.cpp:100-107, 141-181) draws
// `iterations` values with replacement into a bitset and is seeded from
// std::random_device (generate_dataset.cpp:149-150), so its output cannot be
// reproduced.  These generators are seeded (per-set seed = base seed + set
// index) and produce exact set sizes, which the benchmark workloads need:
//   uniformSorted    k distinct values uniform over [lo, lo+N), sorted
//   zipfSizes        set sizes Zipf(s) on {1..max}, clamped to [minSize, max]
//   densitySet       Bernoulli(p) membership over [0, universe)
//   rmatGraph        R-MAT edges -> undirected, deduplicated, loop-free CSR
// plus the reference's own modes (uniform / normal / bernoulli / zipf draws
// into a bitset, values i+1 as in bitsetToVector, generate_dataset.cpp:191-201).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace si {

// xoshiro256** seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (auto& x : s) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            x = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, n), n <= 2^32, unbiased (Lemire's multiply-and-reject)
    uint64_t below(uint64_t n) {
        if (n == 0) return 0;
        if (n == (1ull << 32)) return next() >> 32;
        uint64_t x = next() >> 32;
        uint64_t m = x * n;
        uint32_t l = uint32_t(m);
        if (l < n) {
            uint32_t t = uint32_t(-uint32_t(n)) % uint32_t(n);
            while (l < t) {
                x = next() >> 32;
                m = x * n;
                l = uint32_t(m);
            }
        }
        return m >> 32;
    }
    double uniform01() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// LSD radix sort of uint32 keys (2 passes of 16 bits).
inline void radixSort(std::vector<uint32_t>& v) {
    if (v.size() < 4096) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<uint32_t> tmp(v.size());
    std::vector<uint64_t> cnt(65536);
    for (int pass = 0; pass < 2; ++pass) {
        const int sh = pass * 16;
        std::fill(cnt.begin(), cnt.end(), 0);
        for (uint32_t x : v) ++cnt[(x >> sh) & 0xFFFF];
        uint64_t sum = 0;
        for (auto& c : cnt) {
            uint64_t t = c;
            c = sum;
            sum += t;
        }
        for (uint32_t x : v) tmp[cnt[(x >> sh) & 0xFFFF]++] = x;
        v.swap(tmp);
    }
}

// n distinct values drawn uniformly (as a set) from [lo, lo + N), sorted.
// Draw-with-replacement rounds are deduplicated and topped up until at least
// n distinct values exist; surplus values are then removed at uniformly random
// positions.  Every step is symmetric in the values, so the result is a
// uniformly random n-subset.
inline std::vector<uint32_t> uniformSorted(uint64_t n, uint64_t N, uint64_t seed, uint64_t lo = 0) {
    if (N == 0 || n > N) throw std::invalid_argument("uniformSorted: n > N");
    if (lo + N > (1ull << 32)) throw std::invalid_argument("uniformSorted: range exceeds 2^32");
    Rng rng(seed);
    if (n * 2 > N) {
        // dense: pick the complement and invert
        std::vector<uint32_t> out_;
        std::vector<uint32_t> excl = uniformSorted(N - n, N, seed ^ 0xA5A5A5A5ull, 0);
        out_.reserve(n);
        size_t e = 0;
        for (uint64_t x = 0; x < N; ++x) {
            if (e < excl.size() && excl[e] == x) {
                ++e;
                continue;
            }
            out_.push_back(uint32_t(lo + x));
        }
        return out_;
    }
    std::vector<uint32_t> vals;
    while (vals.size() < n) {
        const uint64_t need = n - vals.size();
        const uint64_t m = need + need / 8 + 16;
        std::vector<uint32_t> d(m);
        for (auto& x : d) x = uint32_t(lo + rng.below(N));
        radixSort(d);
        d.erase(std::unique(d.begin(), d.end()), d.end());
        if (vals.empty()) {
            vals.swap(d);
        } else {
            std::vector<uint32_t> merged;
            merged.reserve(vals.size() + d.size());
            std::merge(vals.begin(), vals.end(), d.begin(), d.end(), std::back_inserter(merged));
            merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
            vals.swap(merged);
        }
    }
    if (vals.size() > n) {
        // remove r = size - n random positions (Floyd's sampling of positions)
        const uint64_t sz = vals.size(), r = sz - n;
        std::vector<uint8_t> drop(sz, 0);
        for (uint64_t j = sz - r; j < sz; ++j) {
            uint64_t t = rng.below(j + 1);
            if (drop[t]) t = j;
            drop[t] = 1;
        }
        size_t w = 0;
        for (size_t i = 0; i < sz; ++i)
            if (!drop[i]) vals[w++] = vals[i];
        vals.resize(n);
    }
    return vals;
}

// Zipf(s) on {1..maxv} by inverse CDF, clamped to [minv, maxv].
struct ZipfSampler {
    std::vector<double> cdf;
    ZipfSampler(uint32_t maxv, double s) : cdf(maxv) {
        double acc = 0.0;
        for (uint32_t i = 0; i < maxv; ++i) {
            acc += std::pow(double(i + 1), -s);
            cdf[i] = acc;
        }
        for (auto& c : cdf) c /= acc;
    }
    uint32_t operator()(Rng& r) const {
        double u = r.uniform01();
        return uint32_t(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    }
};

// Bernoulli(p) membership over [0, universe).
inline std::vector<uint32_t> densitySet(uint32_t universe, double p, uint64_t seed) {
    Rng rng(seed);
    std::vector<uint32_t> s;
    s.reserve(size_t(universe * p * 1.2) + 16);
    const uint64_t thr = uint64_t(p * 18446744073709551616.0 >= 18446744073709551615.0 ? 18446744073709551615.0
                                                                                         : p * 18446744073709551616.0);
    for (uint32_t x = 0; x < universe; ++x)
        if (rng.next() < thr) s.push_back(x);
    return s;
}

struct Csr {
    uint32_t n = 0;
    std::vector<uint64_t> offsets;  // n + 1
    std::vector<uint32_t> adj;      // sorted within each row
};

// R-MAT graph of 2^scale vertices and edgeFactor * 2^scale generated edges
// (Chakrabarti, Zhan, Faloutsos 2004), made undirected, with duplicate edges
// and self loops removed; adjacency lists sorted.
inline Csr rmatGraph(uint32_t scale, uint32_t edgeFactor, double a, double b, double c, uint64_t seed) {
    const uint32_t n = 1u << scale;
    const uint64_t m = uint64_t(edgeFactor) << scale;
    std::vector<uint64_t> edges;
    edges.reserve(2 * m);
    Rng rng(seed);
    const double ab = a + b, abc = a + b + c;
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t u = 0, v = 0;
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            const double r = rng.uniform01();
            const uint32_t bit = 1u << (scale - 1 - lvl);
            if (r < a) {
            } else if (r < ab) {
                v |= bit;
            } else if (r < abc) {
                u |= bit;
            } else {
                u |= bit;
                v |= bit;
            }
        }
        if (u == v) continue;
        edges.push_back((uint64_t(u) << 32) | v);
        edges.push_back((uint64_t(v) << 32) | u);
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    Csr g;
    g.n = n;
    g.offsets.assign(size_t(n) + 1, 0);
    g.adj.resize(edges.size());
    for (size_t i = 0; i < edges.size(); ++i) {
        g.offsets[(edges[i] >> 32) + 1]++;
        g.adj[i] = uint32_t(edges[i]);
    }
    for (uint32_t i = 0; i < n; ++i) g.offsets[i + 1] += g.offsets[i];
    return g;
}

}  // namespace si
