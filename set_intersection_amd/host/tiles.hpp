// Pair-space tiling and pair indexing used by every executable.
//
// Follows include/util.hpp of the reference:
//   tile / splitToTiles / findTilePairs   util.hpp:11-52
//   num_pairs (combination(n, 2))         util.hpp:54-66, every call site uses k = 2
//   triangular_index / quadratic_index    util.hpp:68-74
//   formatBytes                           util.hpp:76-96
// The names and the run order are kept because the multi-run result writer
// (format.hpp, writeResult) depends on them.  All index arithmetic is 64-bit.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

namespace si {

struct tile {
    uint32_t id;
    uint32_t start;
    uint32_t end;
    uint32_t length;
    tile(uint32_t id_, uint32_t start_, uint32_t end_)
        : id(id_), start(start_), end(end_), length(end_ - start_) {}
};

using tile_pair = std::pair<tile, tile>;

// Tiles [i*partition, min((i+1)*partition, cardinality)) for i < ceil(k/partition).
inline std::vector<tile> splitToTiles(uint32_t cardinality, uint32_t partition) {
    std::vector<tile> tiles;
    if (partition == 0) return tiles;
    uint64_t n = (static_cast<uint64_t>(cardinality) + partition - 1) / partition;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t start = i * partition;
        uint64_t end = start + partition;
        if (end > cardinality) end = cardinality;
        tiles.emplace_back(static_cast<uint32_t>(i), static_cast<uint32_t>(start),
                           static_cast<uint32_t>(end));
    }
    return tiles;
}

// All tile pairs (i <= j) in row-major order: the "runs" of the drivers.
inline std::vector<tile_pair> findTilePairs(uint32_t cardinality, uint32_t partition) {
    std::vector<tile_pair> pairs;
    std::vector<tile> tiles = splitToTiles(cardinality, partition);
    for (size_t i = 0; i < tiles.size(); ++i)
        for (size_t j = i; j < tiles.size(); ++j) pairs.emplace_back(tiles[i], tiles[j]);
    return pairs;
}

// Number of unordered pairs of n sets, n(n-1)/2 in 64 bits: the only binomial
// the drivers need (combination(n, 2) at every call site of the reference).
inline uint64_t num_pairs(uint32_t n) { return n < 2 ? 0 : uint64_t(n) * (n - 1) / 2; }

// Row-major index of pair (i, j), i < j, in the strict upper triangle of n x n.
inline unsigned long long triangular_index(unsigned long long n, unsigned long long i,
                                           unsigned long long j) {
    return (n * (n - 1) / 2) - (n - i) * ((n - i) - 1) / 2 + j - i - 1;
}

inline unsigned long long quadratic_index(unsigned long long n, unsigned long long i,
                                          unsigned long long j) {
    return i * n + j;
}

// "1.500 MB"-style size string with the same thresholds as the reference.
inline std::string formatBytes(size_t bytes) {
    const size_t gb = 1073741824, mb = 1048576, kb = 1024;
    char buf[64];
    if (bytes >= gb)
        std::snprintf(buf, sizeof buf, "%.3f GB", double(bytes) / double(gb));
    else if (bytes >= mb)
        std::snprintf(buf, sizeof buf, "%.3f MB", double(bytes) / double(mb));
    else if (bytes >= kb)
        std::snprintf(buf, sizeof buf, "%.3f KB", double(bytes) / double(kb));
    else
        std::snprintf(buf, sizeof buf, "%zu bytes", bytes);
    return buf;
}

}  // namespace si
