// Host-only check of the result writers (set_intersection_amd/host/format.hpp):
// reads a .bin, counts every pair with a two-pointer merge written here, and
// writes the result file in the single-run layout (partition 0) or the
// multi-run layout for the given partition, exactly as the executables do.
// Usage: check_writers <in.bin> <out.txt> <partition> [mm]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "format.hpp"
#include "tiles.hpp"

static uint32_t mergeCount(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
    uint64_t i = 0, j = 0;
    uint32_t c = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { ++c; ++i; ++j; }
    }
    return c;
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    si::Dataset d = si::readDataset(argv[1]);
    const std::string out = argv[2];
    const uint32_t partition = uint32_t(std::atoi(argv[3]));
    const bool mm = argc > 4 && std::string(argv[4]) == "mm";
    auto cnt = [&](uint32_t a, uint32_t b) {
        return mergeCount(d.set(a), d.sizes[a], d.set(b), d.sizes[b]);
    };
    if (partition == 0) {
        std::vector<uint32_t> c(si::num_pairs(d.cardinality));
        for (uint32_t a = 0; a < d.cardinality; ++a)
            for (uint32_t b = a + 1; b < d.cardinality; ++b) c[si::triangular_index(d.cardinality, a, b)] = cnt(a, b);
        si::writeResult(d.cardinality, c, out);
        return 0;
    }
    auto runs = si::findTilePairs(d.cardinality, partition);
    const uint64_t P2 = uint64_t(partition) * partition;
    std::vector<uint32_t> c(runs.size() * P2, 0);
    uint64_t iter = 0;
    for (auto& r : runs) {
        const bool self = r.first.id == r.second.id;
        for (uint32_t i = r.first.start; i < r.first.end; ++i)
            for (uint32_t j = self ? i + 1 : r.second.start; j < r.second.end; ++j) {
                const uint64_t li = i - uint64_t(r.first.id) * partition, lj = j - uint64_t(r.second.id) * partition;
                const uint64_t idx = (self && !mm) ? si::triangular_index(partition, li, lj)
                                                   : si::quadratic_index(partition, li, lj);
                c[iter * P2 + idx] = cnt(i, j);
            }
        ++iter;
    }
    if (mm)
        si::writeResult<uint32_t, true>(runs, partition, c, out);
    else
        si::writeResult<uint32_t, false>(runs, partition, c, out);
    return 0;
}
