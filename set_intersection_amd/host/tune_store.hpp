// Dataset shape features and the tuning store of the batched dispatcher.
//
// The features are the statistics the reference's degree_cdf tool computes
// (degree_cdf.cpp:16-92), summarised:
//   set degrees      the set-size CDF at its median, 90th and 99th percentile
//                    and its maximum (degree_cdf's "_degree_set" / "_set_cdf");
//   element degrees  the number of distinct elements, the mean element degree
//                    and the second moment sum_e deg(e)^2 / total (the last
//                    value of degree_cdf's element CDF, divided by the
//                    element count): the average number of sets sharing an
//                    element, i.e. how much the pairs of the collection
//                    intersect.
// They decide how the dispatcher's work splits over its bins (bitmap rows
// by set size, probe vs merge by size ratio, match density), so a record
// tuned on one collection is a good starting point for another with nearby
// features.
//
// The store is one JSON object: {"<shape key>": {"features": {...},
// "tuning": {...}}, ...}.  Lookups take the exact key if present, else the
// record with the nearest features (log-scale L1 distance) below a threshold.
// Records that do not parse, carry a zero knob or miss a field are skipped,
// so a stale or damaged file never yields an invalid tuning.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "format.hpp"
#include "set_intersection.h"

namespace si {

struct ShapeFeatures {
    uint32_t k = 0, ubits = 0;
    uint64_t total = 0;
    uint32_t q50 = 0, q90 = 0, q99 = 0, maxSize = 0;  // set-size CDF points
    uint64_t distinct = 0;                             // distinct elements
    double degMean = 0.0;                              // total / distinct
    double deg2 = 0.0;                                 // sum_e deg(e)^2 / total
};

// Set-size CDF points of a size vector.
inline void sizeQuantiles(std::vector<uint32_t> s, ShapeFeatures& f) {
    std::sort(s.begin(), s.end());
    auto q = [&](double x) { return s.empty() ? 0u : s[std::min<size_t>(s.size() - 1, size_t(x * s.size()))]; };
    f.q50 = q(0.5);
    f.q90 = q(0.9);
    f.q99 = q(0.99);
    f.maxSize = s.empty() ? 0 : s.back();
}

// Element-degree statistics by sorting a copy of the elements (no array over
// the universe, which may be 2^32 values): deg(e) is the length of e's run.
inline void elementDegrees(const std::vector<uint32_t>& elements, ShapeFeatures& f) {
    std::vector<uint32_t> e(elements);
    std::sort(e.begin(), e.end());
    uint64_t distinct = 0;
    double sq = 0.0;
    for (size_t i = 0; i < e.size();) {
        size_t j = i + 1;
        while (j < e.size() && e[j] == e[i]) ++j;
        const double d = double(j - i);
        sq += d * d;
        ++distinct;
        i = j;
    }
    f.distinct = distinct;
    f.degMean = distinct ? double(e.size()) / double(distinct) : 0.0;
    f.deg2 = e.empty() ? 0.0 : sq / double(e.size());
}

inline ShapeFeatures features(const Dataset& d) {
    ShapeFeatures f;
    f.k = d.cardinality;
    while (f.ubits < 32 && (uint64_t(1) << f.ubits) <= d.universe) ++f.ubits;
    f.total = d.totalElements;
    sizeQuantiles(d.sizes, f);
    elementDegrees(d.elements, f);
    return f;
}

// Exact key of a shape (the integer features).
inline std::string shapeKey(const ShapeFeatures& f) {
    std::ostringstream o;
    o << "k=" << f.k << ";ubits=" << f.ubits << ";total=" << f.total << ";q50=" << f.q50 << ";q90=" << f.q90
      << ";q99=" << f.q99 << ";max=" << f.maxSize;
    return o.str();
}

// Log-scale L1 distance over the features: a factor of 2 in one feature
// counts 1.
inline double featureDistance(const ShapeFeatures& a, const ShapeFeatures& b) {
    auto l = [](double x, double y) { return std::fabs(std::log2(x + 1.0) - std::log2(y + 1.0)); };
    return l(a.k, b.k) + std::fabs(double(a.ubits) - double(b.ubits)) + l(double(a.total), double(b.total)) +
           l(a.q50, b.q50) + l(a.q90, b.q90) + l(a.q99, b.q99) + l(a.maxSize, b.maxSize) +
           l(a.degMean, b.degMean) + l(a.deg2, b.deg2);
}

struct StoreEntry {
    ShapeFeatures f;
    si_tuning t{};
};

namespace store_detail {
inline const char* const kTuning[] = {"bitmap_min", "ratio_search", "chunk_elems", "chunk_elems_bitmap",
                                      "probe_wg_per_cu", "use_graph"};
inline uint32_t* tfield(si_tuning& t, int i) {
    uint32_t* f[] = {&t.bitmap_min, &t.ratio_search, &t.chunk_elems, &t.chunk_elems_bitmap, &t.probe_wg_per_cu,
                     &t.use_graph};
    return f[i];
}
// The number after "name": in body, or false if missing or malformed.
inline bool number(const std::string& body, const char* name, double& v) {
    const std::string n = std::string("\"") + name + "\"";
    const size_t p = body.find(n);
    if (p == std::string::npos) return false;
    const size_t c = body.find(':', p + n.size());
    if (c == std::string::npos) return false;
    const char* s = body.c_str() + c + 1;
    char* end = nullptr;
    v = std::strtod(s, &end);
    return end != s && std::isfinite(v) && v >= 0.0;
}
}  // namespace store_detail

// Records of the store at `path` (empty if the file is missing); malformed or
// invalid records are skipped.
inline std::vector<StoreEntry> loadStore(const std::string& path) {
    using namespace store_detail;
    std::vector<StoreEntry> out;
    std::ifstream in(path);
    if (!in) return out;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    std::vector<size_t> keys;  // each record starts at its "k=... key
    for (size_t p = s.find("\"k="); p != std::string::npos; p = s.find("\"k=", p + 1)) keys.push_back(p);
    for (size_t r = 0; r < keys.size(); ++r) {
        const std::string rec = s.substr(keys[r], (r + 1 < keys.size() ? keys[r + 1] : s.size()) - keys[r]);
        const size_t fb = rec.find("\"features\""), tb = rec.find("\"tuning\"");
        if (fb == std::string::npos || tb == std::string::npos) continue;
        const size_t fo = rec.find('{', fb), to = rec.find('{', tb);
        const size_t fc = fo == std::string::npos ? fo : rec.find('}', fo);
        const size_t tc = to == std::string::npos ? to : rec.find('}', to);
        if (fc == std::string::npos || tc == std::string::npos) continue;
        const std::string fbody = rec.substr(fo, fc - fo), tbody = rec.substr(to, tc - to);
        StoreEntry e;
        double v[12];
        const char* fn[] = {"k", "ubits", "total", "q50", "q90", "q99", "max", "distinct", "deg_mean", "deg2"};
        bool ok = true;
        for (int i = 0; i < 10 && ok; ++i) ok = number(fbody, fn[i], v[i]);
        if (!ok) continue;
        e.f.k = uint32_t(v[0]);
        e.f.ubits = uint32_t(v[1]);
        e.f.total = uint64_t(v[2]);
        e.f.q50 = uint32_t(v[3]);
        e.f.q90 = uint32_t(v[4]);
        e.f.q99 = uint32_t(v[5]);
        e.f.maxSize = uint32_t(v[6]);
        e.f.distinct = uint64_t(v[7]);
        e.f.degMean = v[8];
        e.f.deg2 = v[9];
        e.t.version = SI_TUNING_VERSION;
        for (int i = 0; i < 6 && ok; ++i) {
            double x;
            ok = number(tbody, kTuning[i], x) && x < 4294967296.0;
            if (ok) *tfield(e.t, i) = uint32_t(x);
        }
        // every knob but use_graph must be positive (si_batched_count refuses zeros)
        for (int i = 0; i < 5 && ok; ++i) ok = *tfield(e.t, i) > 0;
        if (ok) out.push_back(e);
    }
    return out;
}

inline bool saveStore(const std::string& path, const std::vector<StoreEntry>& entries) {
    using namespace store_detail;
    std::ofstream f(path);
    if (!f) return false;
    f << std::setprecision(17) << "{\n";
    for (size_t n = 0; n < entries.size(); ++n) {
        const StoreEntry& e = entries[n];
        si_tuning t = e.t;
        f << "  \"" << shapeKey(e.f) << "\": {\"features\": {\"k\": " << e.f.k << ", \"ubits\": " << e.f.ubits
          << ", \"total\": " << e.f.total << ", \"q50\": " << e.f.q50 << ", \"q90\": " << e.f.q90
          << ", \"q99\": " << e.f.q99 << ", \"max\": " << e.f.maxSize << ", \"distinct\": " << e.f.distinct
          << ", \"deg_mean\": " << e.f.degMean << ", \"deg2\": " << e.f.deg2 << "}, \"tuning\": {";
        for (int i = 0; i < 6; ++i) f << (i ? ", " : "") << "\"" << kTuning[i] << "\": " << *tfield(t, i);
        f << "}}" << (n + 1 < entries.size() ? "," : "") << "\n";
    }
    f << "}\n";
    return bool(f);
}

// Replaces the record with the same key, or appends.
inline void putEntry(std::vector<StoreEntry>& entries, const StoreEntry& e) {
    const std::string key = shapeKey(e.f);
    for (auto& x : entries)
        if (shapeKey(x.f) == key) {
            x = e;
            return;
        }
    entries.push_back(e);
}

// The record for f: the exact key, else the nearest within maxDistance
// (nullptr if none); *distance receives the distance of the one returned.
inline const StoreEntry* nearestEntry(const std::vector<StoreEntry>& entries, const ShapeFeatures& f,
                                      double maxDistance, double* distance = nullptr) {
    const StoreEntry* best = nullptr;
    double bd = 1e300;
    for (const auto& e : entries) {
        const double dd = shapeKey(e.f) == shapeKey(f) ? 0.0 : featureDistance(e.f, f);
        if (dd < bd) {
            bd = dd;
            best = &e;
        }
    }
    if (!best || bd > maxDistance) return nullptr;
    if (distance) *distance = bd;
    return best;
}

}  // namespace si
