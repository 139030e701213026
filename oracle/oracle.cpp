// oracle.cpp -- CPU restatement of the reference's set-intersection algorithms.
//
// TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
// GPU path and the CPU baseline that bench.py times; it is never linked into
// or called by the product (libsi.so, the executables under
// set_intersection_amd/apps).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it.
//
// What each function restates (paths relative to the reference repository):
//   oracle_std_count / oracle_std_materialize
//       std_set_intersection.cpp:68-73 -- std::set_intersection into a vector of
//       min(|a|,|b|) elements; the count is the resulting size.
//   oracle_serial_count
//       serial_intersect.cpp:93-113 -- two-pointer merge count.
//   oracle_bitset_count
//       boost_bitset.cpp:49,68-80 -- bitset of universe+1 bits built from set a,
//       each element of b probed.  (std::vector<uint64_t> replaces
//       boost::dynamic_bitset, which is absent offline; universe+1 is computed in
//       64 bits, so universe = 2^32-1 works, unlike boost_bitset.cpp:49.)
//   oracle_simd_count
//       simd.cpp:77 calls SIMDCompressionAndIntersection @ b666a60's "simd"
//       intersection (not vendored, unavailable offline).  Restated from the
//       published method (Lemire, Boytsov, Kurz, "SIMD Compression and the
//       Intersection of Sorted Integers", Softw. Pract. Exper. 2016, sec. 4:
//       V1 block comparison for skewed sizes, all-pairs block compare
//       otherwise) with SSE2 intrinsics.
//   oracle_all_pairs
//       the OpenMP driver of std_set_intersection.cpp:58-76 (same loop for
//       serial_intersect.cpp:61-76, simd.cpp:61-83, boost_bitset.cpp:58-84):
//       thread t owns a in [k*t/T, k*(t+1)/T), all b > a, result at
//       triangular_index(k, a, b) (util.hpp:68-70).
//   oracle_edges
//       sum over an edge list of |N(u) ∩ N(v)| with std::set_intersection.
//
// Parity pins (tests/test_oracle.py): closed-form inputs (multiples of p and
// q, identical, disjoint, subset, boundary values 0 / 2^31 / 2^32-1), brute
// force against Python sets on small random inputs, agreement of all four
// methods, triangular-index enumeration order, and the committed golden
// fixtures under tests/golden/.
#include <emmintrin.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <vector>

namespace {

// Output iterator that only counts assignments.
struct CountingIterator {
    using iterator_category = std::output_iterator_tag;
    using value_type = void;
    using difference_type = std::ptrdiff_t;
    using pointer = void;
    using reference = void;
    uint64_t* n;
    CountingIterator& operator*() { return *this; }
    CountingIterator& operator=(uint32_t) {
        ++*n;
        return *this;
    }
    CountingIterator& operator++() { return *this; }
    CountingIterator operator++(int) { return *this; }
};

uint64_t stdCount(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    uint64_t n = 0;
    std::set_intersection(a, a + na, b, b + nb, CountingIterator{&n});
    return n;
}

uint64_t serialCount(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    uint64_t i = 0, j = 0, count = 0;
    while (i < na && j < nb) {
        if (a[i] == b[j]) {
            ++i;
            ++j;
            ++count;
        } else if (a[i] < b[j]) {
            ++i;
        } else {
            ++j;
        }
    }
    return count;
}

// Bitset of `bits` bits (universe + 1).
struct Bitset {
    std::vector<uint64_t> w;
    explicit Bitset(uint64_t bits) : w((bits + 63) / 64, 0) {}
    void set(uint64_t x) { w[x >> 6] |= 1ull << (x & 63); }
    bool test(uint64_t x) const { return (x >> 6) < w.size() && ((w[x >> 6] >> (x & 63)) & 1); }
};

uint64_t bitsetCount(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint64_t universe) {
    uint64_t maxv = universe;
    for (uint64_t i = 0; i < na; ++i) maxv = std::max<uint64_t>(maxv, a[i]);
    Bitset bs(maxv + 1);
    for (uint64_t i = 0; i < na; ++i) bs.set(a[i]);
    uint64_t c = 0;
    for (uint64_t i = 0; i < nb; ++i) c += bs.test(b[i]);
    return c;
}

// ---- SIMD intersection (Lemire, Boytsov & Kurz 2016), SSE2 -----------------
// Unsigned compare via sign flip: SSE2 has only signed 32-bit compares.
inline __m128i flip(__m128i x) { return _mm_xor_si128(x, _mm_set1_epi32(int(0x80000000u))); }

// V1 (sec. 4.1 of the paper): for every element of the small set, skip 4-wide
// blocks of the large set whose last element is smaller, then compare the
// element against the block with one vector equality test.
uint64_t simdV1(const uint32_t* small, uint64_t ns, const uint32_t* large, uint64_t nl) {
    uint64_t count = 0, j = 0;
    for (uint64_t i = 0; i < ns; ++i) {
        uint32_t x = small[i];
        while (j + 4 <= nl && large[j + 3] < x) j += 4;
        if (j + 4 <= nl) {
            __m128i blk = _mm_loadu_si128(reinterpret_cast<const __m128i*>(large + j));
            __m128i eq = _mm_cmpeq_epi32(blk, _mm_set1_epi32(int(x)));
            count += _mm_movemask_ps(_mm_castsi128_ps(eq)) != 0;
        } else {
            while (j < nl && large[j] < x) ++j;
            if (j == nl) break;
            count += large[j] == x;
        }
    }
    return count;
}

// Block-vs-block comparison for similar sizes (the "SIMD" all-pairs compare of
// sec. 4: four rotations of a 4-element block against a 4-element block, then
// advance the block with the smaller maximum).
uint64_t simdBlocks(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    uint64_t i = 0, j = 0, count = 0;
    while (i + 4 <= na && j + 4 <= nb) {
        __m128i va = _mm_loadu_si128(reinterpret_cast<const __m128i*>(a + i));
        __m128i vb = _mm_loadu_si128(reinterpret_cast<const __m128i*>(b + j));
        __m128i m = _mm_cmpeq_epi32(va, vb);
        m = _mm_or_si128(m, _mm_cmpeq_epi32(va, _mm_shuffle_epi32(vb, _MM_SHUFFLE(0, 3, 2, 1))));
        m = _mm_or_si128(m, _mm_cmpeq_epi32(va, _mm_shuffle_epi32(vb, _MM_SHUFFLE(1, 0, 3, 2))));
        m = _mm_or_si128(m, _mm_cmpeq_epi32(va, _mm_shuffle_epi32(vb, _MM_SHUFFLE(2, 1, 0, 3))));
        count += __builtin_popcount(unsigned(_mm_movemask_ps(_mm_castsi128_ps(m))));
        uint32_t amax = a[i + 3], bmax = b[j + 3];
        if (amax <= bmax) i += 4;
        if (bmax <= amax) j += 4;
    }
    // scalar tail
    while (i < na && j < nb) {
        if (a[i] == b[j]) {
            ++count;
            ++i;
            ++j;
        } else if (a[i] < b[j]) {
            ++i;
        } else {
            ++j;
        }
    }
    return count;
}

uint64_t simdCount(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    if (na > nb) {
        std::swap(a, b);
        std::swap(na, nb);
    }
    if (na == 0) return 0;
    if (nb / na >= 50) return simdV1(a, na, b, nb);
    return simdBlocks(a, na, b, nb);
}

uint64_t methodCount(int method, const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb,
                     uint64_t universe) {
    switch (method) {
        case 0: return stdCount(a, na, b, nb);
        case 1: return serialCount(a, na, b, nb);
        case 2: return bitsetCount(a, na, b, nb, universe);
        default: return simdCount(a, na, b, nb);
    }
}

inline uint64_t triangularIndex(uint64_t n, uint64_t i, uint64_t j) {
    return (n * (n - 1) / 2) - (n - i) * ((n - i) - 1) / 2 + j - i - 1;
}

}  // namespace

extern "C" {

uint64_t oracle_std_count(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    return stdCount(a, na, b, nb);
}

uint64_t oracle_std_materialize(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint32_t* out) {
    return static_cast<uint64_t>(std::set_intersection(a, a + na, b, b + nb, out) - out);
}

uint64_t oracle_serial_count(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    return serialCount(a, na, b, nb);
}

uint64_t oracle_bitset_count(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb, uint64_t universe) {
    return bitsetCount(a, na, b, nb, universe);
}

uint64_t oracle_simd_count(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    return simdCount(a, na, b, nb);
}

uint64_t oracle_triangular_index(uint64_t n, uint64_t i, uint64_t j) { return triangularIndex(n, i, j); }

// All k(k-1)/2 pair counts in triangular order.  method: 0 std, 1 serial,
// 2 bitset, 3 simd.  offsets has k+1 entries.  Returns the wall time of the
// pair loop in seconds (the HostTimer interval of std_set_intersection.cpp:57-77).
double oracle_all_pairs(int method, uint32_t k, uint32_t universe, const uint64_t* offsets,
                        const uint32_t* elements, uint32_t threads, uint32_t* counts) {
    if (threads == 0) threads = 1;
    auto t0 = std::chrono::steady_clock::now();
    if (method == 2) {
        // boost_bitset.cpp:60-84: one bitset per a, probe every b > a.
#pragma omp parallel num_threads(threads)
        {
            uint32_t t = uint32_t(omp_get_thread_num());
            uint32_t lo = uint32_t(uint64_t(k) * t / threads), hi = uint32_t(uint64_t(k) * (t + 1) / threads);
            for (uint32_t a = lo; a < hi; ++a) {
                const uint32_t* sa = elements + offsets[a];
                uint64_t na = offsets[a + 1] - offsets[a];
                uint64_t maxv = universe;
                for (uint64_t i = 0; i < na; ++i) maxv = std::max<uint64_t>(maxv, sa[i]);
                Bitset bs(maxv + 1);
                for (uint64_t i = 0; i < na; ++i) bs.set(sa[i]);
                for (uint32_t b = a + 1; b < k; ++b) {
                    const uint32_t* sb = elements + offsets[b];
                    uint64_t nb = offsets[b + 1] - offsets[b], c = 0;
                    for (uint64_t i = 0; i < nb; ++i) c += bs.test(sb[i]);
                    counts[triangularIndex(k, a, b)] = uint32_t(c);
                }
            }
        }
    } else {
#pragma omp parallel num_threads(threads)
        {
            uint32_t t = uint32_t(omp_get_thread_num());
            uint32_t lo = uint32_t(uint64_t(k) * t / threads), hi = uint32_t(uint64_t(k) * (t + 1) / threads);
            for (uint32_t a = lo; a < hi; ++a)
                for (uint32_t b = a + 1; b < k; ++b)
                    counts[triangularIndex(k, a, b)] = uint32_t(
                        methodCount(method, elements + offsets[a], offsets[a + 1] - offsets[a],
                                    elements + offsets[b], offsets[b + 1] - offsets[b], universe));
        }
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Counts for pairs listed explicitly (pa[i], pb[i]); used for sampled checks
// of large collections.  Returns seconds.
double oracle_pairs(int method, uint32_t universe, const uint64_t* offsets, const uint32_t* elements,
                    const uint32_t* pa, const uint32_t* pb, uint64_t npairs, uint32_t threads, uint32_t* counts) {
    if (threads == 0) threads = 1;
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(npairs); ++i) {
        uint32_t a = pa[i], b = pb[i];
        counts[i] = uint32_t(methodCount(method, elements + offsets[a], offsets[a + 1] - offsets[a],
                                         elements + offsets[b], offsets[b + 1] - offsets[b], universe));
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Edge-wise counts: counts[e] = |N(u[e]) ∩ N(v[e])|; returns the uint64 total.
uint64_t oracle_edges(const uint64_t* offsets, const uint32_t* elements, const uint32_t* u, const uint32_t* v,
                      uint64_t m, uint32_t threads, uint32_t* counts) {
    if (threads == 0) threads = 1;
    uint64_t total = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024) reduction(+ : total)
    for (int64_t e = 0; e < int64_t(m); ++e) {
        uint32_t a = u[e], b = v[e];
        uint64_t c = stdCount(elements + offsets[a], offsets[a + 1] - offsets[a], elements + offsets[b],
                              offsets[b + 1] - offsets[b]);
        counts[e] = uint32_t(c);
        total += c;
    }
    return total;
}

// Dense count matrix of tile pair (rows [a0,a1), cols [b0,b1)):
// C[(i-a0)*ld + (j-b0)] = |S_i ∩ S_j|.  Used to check the bitmap-GEMM path.
void oracle_tile_matrix(const uint64_t* offsets, const uint32_t* elements, uint32_t a0, uint32_t a1, uint32_t b0,
                        uint32_t b1, uint32_t ld, uint32_t threads, uint32_t* C) {
    if (threads == 0) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
    for (int64_t i = a0; i < int64_t(a1); ++i)
        for (uint32_t j = b0; j < b1; ++j)
            C[uint64_t(i - a0) * ld + (j - b0)] = uint32_t(stdCount(elements + offsets[i], offsets[i + 1] - offsets[i],
                                                                   elements + offsets[j], offsets[j + 1] - offsets[j]));
}

}  // extern "C"
