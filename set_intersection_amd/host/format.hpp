// The reference's binary dataset format and result-file writers.
//
// .bin layout (include/io.hpp:25-54, little-endian, packed, no magic):
//     u32 k | u32 universe | u64 totalElements | u32 sizes[k] | u32 elements[total]
// Sets are stored back to back in `sizes` order; each set is sorted ascending.
//
// Differences from the reference reader, all on malformed input only:
//   - a missing or truncated file, or sum(sizes) != totalElements, is an error
//     (the reference reads past EOF silently, io.hpp:29-49);
//   - offsets has k+1 64-bit entries (the reference never fills offsets).
// Result writers produce byte-identical text to io.hpp:91-136.
#pragma once

#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "tiles.hpp"

namespace si {

struct Dataset {
    uint32_t cardinality = 0;  // k
    uint32_t universe = 0;
    uint64_t totalElements = 0;
    std::vector<uint32_t> sizes;
    std::vector<uint32_t> elements;
    std::vector<uint64_t> offsets;  // k+1 entries, offsets[k] == totalElements

    const uint32_t* set(uint32_t i) const { return elements.data() + offsets[i]; }
};

class FormatError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

inline void computeOffsets(Dataset& d) {
    d.offsets.assign(static_cast<size_t>(d.cardinality) + 1, 0);
    for (uint32_t i = 0; i < d.cardinality; ++i) d.offsets[i + 1] = d.offsets[i] + d.sizes[i];
}

// readDataset: io.hpp:25-54.  Throws FormatError on a missing/short file.
inline Dataset readDataset(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw FormatError("cannot open dataset file '" + path + "'");
    Dataset d;
    auto readExact = [&](void* dst, size_t bytes, const char* what) {
        if (bytes && std::fread(dst, 1, bytes, f) != bytes) {
            std::fclose(f);
            throw FormatError(std::string("truncated dataset file '") + path + "' while reading " + what);
        }
    };
    readExact(&d.cardinality, 4, "k");
    readExact(&d.universe, 4, "universe");
    readExact(&d.totalElements, 8, "totalElements");
    d.sizes.resize(d.cardinality);
    readExact(d.sizes.data(), sizeof(uint32_t) * static_cast<size_t>(d.cardinality), "sizes");
    uint64_t sum = 0;
    for (uint32_t s : d.sizes) sum += s;
    if (sum != d.totalElements) {
        std::fclose(f);
        throw FormatError("dataset '" + path + "': sum of set sizes (" + std::to_string(sum) +
                          ") != totalElements (" + std::to_string(d.totalElements) + ")");
    }
    d.elements.resize(d.totalElements);
    readExact(d.elements.data(), sizeof(uint32_t) * d.totalElements, "elements");
    std::fclose(f);
    computeOffsets(d);
    return d;
}

// datasetToCollection: io.hpp:56-68.
inline std::vector<std::vector<uint32_t>> datasetToCollection(const Dataset& d) {
    std::vector<std::vector<uint32_t>> sets(d.cardinality);
    for (uint32_t i = 0; i < d.cardinality; ++i)
        sets[i].assign(d.elements.begin() + d.offsets[i], d.elements.begin() + d.offsets[i + 1]);
    return sets;
}

// Builds a Dataset from a collection (the inverse of datasetToCollection).
inline Dataset collectionToDataset(const std::vector<std::vector<uint32_t>>& sets, uint32_t universe) {
    Dataset d;
    d.cardinality = static_cast<uint32_t>(sets.size());
    d.universe = universe;
    d.sizes.resize(sets.size());
    for (size_t i = 0; i < sets.size(); ++i) {
        d.sizes[i] = static_cast<uint32_t>(sets[i].size());
        d.totalElements += sets[i].size();
    }
    d.elements.reserve(d.totalElements);
    for (auto& s : sets) d.elements.insert(d.elements.end(), s.begin(), s.end());
    computeOffsets(d);
    return d;
}

// writeDataset: io.hpp:70-89 (same bytes, written in bulk).
inline void writeDataset(uint32_t k, uint32_t universe, uint64_t totalElements,
                         const std::vector<std::vector<uint32_t>>& dataset, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw FormatError("cannot create '" + path + "'");
    bool ok = std::fwrite(&k, 4, 1, f) == 1 && std::fwrite(&universe, 4, 1, f) == 1 &&
              std::fwrite(&totalElements, 8, 1, f) == 1;
    for (auto& s : dataset) {
        uint32_t n = static_cast<uint32_t>(s.size());
        ok = ok && std::fwrite(&n, 4, 1, f) == 1;
    }
    for (auto& s : dataset)
        ok = ok && (s.empty() || std::fwrite(s.data(), 4, s.size(), f) == s.size());
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw FormatError("write error on '" + path + "'");
}

inline void writeDataset(const Dataset& d, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw FormatError("cannot create '" + path + "'");
    bool ok = std::fwrite(&d.cardinality, 4, 1, f) == 1 && std::fwrite(&d.universe, 4, 1, f) == 1 &&
              std::fwrite(&d.totalElements, 8, 1, f) == 1;
    ok = ok && (d.sizes.empty() || std::fwrite(d.sizes.data(), 4, d.sizes.size(), f) == d.sizes.size());
    ok = ok && (d.elements.empty() ||
                std::fwrite(d.elements.data(), 4, d.elements.size(), f) == d.elements.size());
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw FormatError("write error on '" + path + "'");
}

// Buffered text sink for the "(i,j): c" lines.
class LineWriter {
   public:
    explicit LineWriter(const std::string& path) : f_(std::fopen(path.c_str(), "w")) {
        if (!f_) throw FormatError("cannot create '" + path + "'");
        buf_.resize(1 << 20);
    }
    // A destructor must not throw: on an early return or unwind a failed final
    // write is dropped here (the normal path calls close(), which reports it).
    ~LineWriter() {
        try {
            close();
        } catch (...) {
            if (f_) std::fclose(f_);
            f_ = nullptr;
        }
    }
    template <typename T>
    void line(uint64_t i, uint64_t j, T value) {
        if (pos_ + 96 > buf_.size()) flush();
        char* p = buf_.data() + pos_;
        *p++ = '(';
        p = std::to_chars(p, p + 24, i).ptr;
        *p++ = ',';
        p = std::to_chars(p, p + 24, j).ptr;
        *p++ = ')';
        *p++ = ':';
        *p++ = ' ';
        p = putValue(p, value);
        *p++ = '\n';
        pos_ = static_cast<size_t>(p - buf_.data());
    }
    void close() {
        if (!f_) return;
        flush();
        std::fclose(f_);
        f_ = nullptr;
    }

   private:
    template <typename T>
    static char* putValue(char* p, T v) {
        return std::to_chars(p, p + 32, v).ptr;
    }
    // Counts stored as float (the reference's mm path) are integral; they are
    // printed as integers, which is the shortest round-trip text of an integral
    // float below 2^24 (what {fmt} >= 7 emits for "{}").
    static char* putValue(char* p, float v) {
        if (v >= 0.0f && v < 16777216.0f && v == static_cast<float>(static_cast<uint32_t>(v)))
            return std::to_chars(p, p + 32, static_cast<uint32_t>(v)).ptr;
        return p + std::snprintf(p, 32, "%g", static_cast<double>(v));
    }
    void flush() {
        if (pos_ && std::fwrite(buf_.data(), 1, pos_, f_) != pos_) throw FormatError("write error");
        pos_ = 0;
    }
    FILE* f_;
    std::vector<char> buf_;
    size_t pos_ = 0;
};

// writeResult(k, counts): io.hpp:91-102 -- single run, triangular layout.
template <typename T>
inline void writeResult(uint32_t k, const std::vector<T>& counts, const std::string& output) {
    LineWriter w(output);
    for (uint64_t a = 0; a < k; ++a)
        for (uint64_t b = a + 1; b < k; ++b) w.line(a + 1, b + 1, counts[triangular_index(k, a, b)]);
    w.close();
}

// The lines of one run of the multi-run layout (the body of the run loop of
// io.hpp:104-136); `block` is the run's partition^2 block of counts.
template <typename T, bool mm = false>
inline void writeRun(LineWriter& w, const tile_pair& run, uint32_t partition, const T* block) {
    const tile& A = run.first;
    const tile& B = run.second;
    bool selfJoin = (A.id == B.id);
    for (uint64_t i = A.start; i < A.end; ++i) {
        for (uint64_t j = (selfJoin ? i + 1 : B.start); j < B.end; ++j) {
            uint64_t pairOffset;
            uint64_t li = i - uint64_t(A.id) * partition, lj = j - uint64_t(B.id) * partition;
            if (selfJoin && !mm)
                pairOffset = triangular_index(partition, li, lj);
            else
                pairOffset = quadratic_index(partition, li, lj);
            w.line(i + 1, j + 1, block[pairOffset]);
        }
    }
}

// writeResult<T, mm>(runs, partition, counts): io.hpp:104-136 -- multi-run
// layout, one partition^2 block of counts per run in findTilePairs order.
template <typename T, bool mm = false>
inline void writeResult(const std::vector<tile_pair>& runs, uint32_t partition, const std::vector<T>& counts,
                        const std::string& output) {
    LineWriter w(output);
    uint64_t iter = 0;
    for (auto& run : runs) {
        writeRun<T, mm>(w, run, partition, counts.data() + iter * uint64_t(partition) * partition);
        ++iter;
    }
    w.close();
}

}  // namespace si
