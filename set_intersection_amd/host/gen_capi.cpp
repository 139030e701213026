// libsigen.so: C interface of the synthetic dataset generators (generate.hpp)
// for the Python tooling (tests, bench.py).  Collections are returned through
// an opaque handle holding sizes[k] and elements[total] in .bin order.
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "generate.hpp"

namespace {
struct Coll {
    uint32_t k = 0;
    uint32_t universe = 0;
    std::vector<uint32_t> sizes;
    std::vector<uint32_t> elements;
};
thread_local std::string g_err;

// Fills elements for given sizes: set i is a uniform sample of sizes[i]
// distinct values from [0, N), seeded with seed + i.
void fillUniform(Coll& c, uint64_t N, uint64_t seed) {
    std::vector<uint64_t> off(c.k + 1, 0);
    for (uint32_t i = 0; i < c.k; ++i) off[i + 1] = off[i] + c.sizes[i];
    c.elements.resize(off[c.k]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < int64_t(c.k); ++i) {
        std::vector<uint32_t> s = si::uniformSorted(c.sizes[i], N, seed + uint64_t(i));
        std::memcpy(c.elements.data() + off[i], s.data(), s.size() * 4);
    }
}
}  // namespace

extern "C" {

const char* sigen_last_error(void) { return g_err.c_str(); }

// Sorted unique uniform sample of n values from [lo, lo + N).
int sigen_uniform_sorted(uint64_t n, uint64_t N, uint64_t seed, uint64_t lo, uint32_t* out) {
    try {
        std::vector<uint32_t> v = si::uniformSorted(n, N, seed, lo);
        std::memcpy(out, v.data(), v.size() * 4);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// k sets, sizes Zipf(s) on {1..max_size} clamped to [min_size, max_size],
// elements uniform without replacement from [0, universe).
void* sigen_zipf_collection(uint32_t k, double s, uint32_t min_size, uint32_t max_size, uint64_t universe,
                            uint64_t seed) {
    try {
        auto* c = new Coll();
        c->k = k;
        c->universe = uint32_t(universe - 1);
        si::ZipfSampler z(max_size, s);
        si::Rng rng(seed);
        c->sizes.resize(k);
        for (auto& x : c->sizes) {
            uint32_t v = z(rng);
            x = v < min_size ? min_size : v;
        }
        fillUniform(*c, universe, seed + 1000003ull);
        return c;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// k sets with explicit sizes, elements uniform without replacement from [0, universe).
void* sigen_sized_collection(uint32_t k, const uint32_t* sizes, uint64_t universe, uint64_t seed) {
    try {
        auto* c = new Coll();
        c->k = k;
        c->universe = uint32_t(universe - 1);
        c->sizes.assign(sizes, sizes + k);
        fillUniform(*c, universe, seed);
        return c;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// k sets over [0, universe), set i has Bernoulli density p_i ~ U[pmin, pmax].
void* sigen_density_collection(uint32_t k, uint32_t universe, double pmin, double pmax, uint64_t seed) {
    try {
        auto* c = new Coll();
        c->k = k;
        c->universe = universe - 1;
        std::vector<std::vector<uint32_t>> sets(k);
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < int64_t(k); ++i) {
            si::Rng r(seed + uint64_t(i));
            double p = pmin + (pmax - pmin) * r.uniform01();
            sets[i] = si::densitySet(universe, p, seed * 31 + uint64_t(i));
        }
        c->sizes.resize(k);
        size_t total = 0;
        for (uint32_t i = 0; i < k; ++i) total += (c->sizes[i] = uint32_t(sets[i].size()));
        c->elements.reserve(total);
        for (auto& s : sets) c->elements.insert(c->elements.end(), s.begin(), s.end());
        return c;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// R-MAT graph as a collection: set v = sorted adjacency list of vertex v.
void* sigen_rmat_collection(uint32_t scale, uint32_t edge_factor, double a, double b, double c_, uint64_t seed) {
    try {
        si::Csr g = si::rmatGraph(scale, edge_factor, a, b, c_, seed);
        auto* c = new Coll();
        c->k = g.n;
        c->universe = g.n - 1;
        c->sizes.resize(g.n);
        for (uint32_t i = 0; i < g.n; ++i) c->sizes[i] = uint32_t(g.offsets[i + 1] - g.offsets[i]);
        c->elements = std::move(g.adj);
        return c;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

uint32_t sigen_coll_k(void* h) { return static_cast<Coll*>(h)->k; }
uint32_t sigen_coll_universe(void* h) { return static_cast<Coll*>(h)->universe; }
uint64_t sigen_coll_total(void* h) { return static_cast<Coll*>(h)->elements.size(); }
void sigen_coll_sizes(void* h, uint32_t* out) {
    auto* c = static_cast<Coll*>(h);
    std::memcpy(out, c->sizes.data(), c->sizes.size() * 4);
}
void sigen_coll_elements(void* h, uint32_t* out) {
    auto* c = static_cast<Coll*>(h);
    std::memcpy(out, c->elements.data(), c->elements.size() * 4);
}
void sigen_coll_free(void* h) { delete static_cast<Coll*>(h); }

}  // extern "C"
