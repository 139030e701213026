// degree_cdf -- set and element degree statistics of a .bin collection;
// counterpart of the reference's src/degree_cdf.cpp:16-92.
// Usage: degree_cdf <input.bin> <output-prefix>
// Writes, one "first second" pair per line, sorted by the first value:
//   <prefix>_degree_set.txt      (|S_i|, i+1)
//   <prefix>_degree_element.txt  (number of sets containing e, e), e in [0, universe]
//   <prefix>_set_cdf.txt         (|S_i|, running sum of sum_{e in S_i} deg(e))
//   <prefix>_element_cdf.txt     (deg(e), running sum of deg(e)^2)
// The element CDF is the one computed at degree_cdf.cpp:66-71 (the reference
// writes the set CDF into that file by mistake, degree_cdf.cpp:88); the
// inverted index covers elements up to and including `universe`
// (degree_cdf.cpp:23 sizes it one short); ties keep input order.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "format.hpp"
#include "tune_store.hpp"

template <typename A, typename B>
static bool writePairs(const std::vector<std::pair<A, B>>& v, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (auto& p : v) std::fprintf(f, "%llu %llu\n", (unsigned long long)p.first, (unsigned long long)p.second);
    return std::fclose(f) == 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("Usage: %s <input.bin> <output-prefix>\n", argv[0]);
        return 1;
    }
    si::Dataset d;
    try {
        d = si::readDataset(argv[1]);
    } catch (const si::FormatError& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    uint32_t maxv = d.universe;
    for (uint32_t x : d.elements) maxv = std::max(maxv, x);
    std::vector<uint64_t> deg(uint64_t(maxv) + 1, 0);
    for (uint32_t x : d.elements) deg[x]++;

    std::vector<std::pair<uint64_t, uint64_t>> degreeSet, setCdf, degreeElement, elementCdf;
    for (uint32_t i = 0; i < d.cardinality; ++i) {
        degreeSet.emplace_back(d.sizes[i], uint64_t(i) + 1);
        uint64_t inv = 0;
        for (uint64_t k = d.offsets[i]; k < d.offsets[i + 1]; ++k) inv += deg[d.elements[k]];
        setCdf.emplace_back(d.sizes[i], inv);
    }
    for (uint64_t e = 0; e < deg.size(); ++e) {
        degreeElement.emplace_back(deg[e], e);
        elementCdf.emplace_back(deg[e], deg[e]);
    }
    auto byFirst = [](const auto& a, const auto& b) { return a.first < b.first; };
    std::stable_sort(degreeSet.begin(), degreeSet.end(), byFirst);
    std::stable_sort(degreeElement.begin(), degreeElement.end(), byFirst);
    std::stable_sort(setCdf.begin(), setCdf.end(), byFirst);
    std::stable_sort(elementCdf.begin(), elementCdf.end(), byFirst);
    for (size_t i = 0; i < elementCdf.size(); ++i)
        elementCdf[i].second = (i ? elementCdf[i - 1].second : 0) + elementCdf[i].second * elementCdf[i].second;
    for (size_t i = 1; i < setCdf.size(); ++i) setCdf[i].second += setCdf[i - 1].second;

    const std::string pre = argv[2];
    bool ok = writePairs(degreeSet, pre + "_degree_set.txt") && writePairs(degreeElement, pre + "_degree_element.txt") &&
              writePairs(setCdf, pre + "_set_cdf.txt") && writePairs(elementCdf, pre + "_element_cdf.txt");
    if (!ok) {
        std::printf("write error\n");
        return 1;
    }
    // the summary the batched dispatcher's tuning store keys its records on
    const si::ShapeFeatures f = si::features(d);
    std::printf("set size q50 %u q90 %u q99 %u max %u\n", f.q50, f.q90, f.q99, f.maxSize);
    std::printf("distinct elements %llu  element degree mean %.6g  second moment %.6g\n",
                (unsigned long long)f.distinct, f.degMean, f.deg2);
    std::printf("Finished\n");
    return 0;
}
