// Host-only check of the batched dispatcher's tuning store
// (set_intersection_amd/host/tune_store.hpp).
// Usage: check_tune_store <scratch-dir> [<in.bin>]
//   without a .bin: store round trip, lookups and damaged files;
//   with a .bin: prints "distinct deg_mean deg2 total" of its features.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "tune_store.hpp"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "FAILED line %d: %s\n", __LINE__, #c); \
            return 1;                                              \
        }                                                          \
    } while (0)

static si::StoreEntry entry(uint32_t k, uint32_t q50, double deg2, uint32_t bm) {
    si::StoreEntry e;
    e.f.k = k;
    e.f.ubits = 20;
    e.f.total = uint64_t(k) * 100;
    e.f.q50 = q50;
    e.f.q90 = q50 * 4;
    e.f.q99 = q50 * 16;
    e.f.maxSize = q50 * 64;
    e.f.distinct = 1000;
    e.f.degMean = 3.5;
    e.f.deg2 = deg2;
    e.t.version = SI_TUNING_VERSION;
    e.t.bitmap_min = bm;
    e.t.ratio_search = 32;
    e.t.chunk_elems = 65536;
    e.t.chunk_elems_bitmap = 1u << 20;
    e.t.probe_wg_per_cu = 4;
    e.t.use_graph = 1;
    return e;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    if (argc > 2) {
        const si::ShapeFeatures f = si::features(si::readDataset(argv[2]));
        std::printf("%llu %.17g %.17g %llu\n", (unsigned long long)f.distinct, f.degMean, f.deg2,
                    (unsigned long long)f.total);
        return 0;
    }
    // closed form: {0..9}, {0..9}, {5..14}: degrees 2 (0-4), 3 (5-9), 1 (10-14)
    si::Dataset d;
    d.cardinality = 3;
    d.universe = 14;
    for (int s = 0; s < 3; ++s) {
        const uint32_t lo = s == 2 ? 5 : 0;
        for (uint32_t x = lo; x < lo + 10; ++x) d.elements.push_back(x);
        d.sizes.push_back(10);
    }
    d.totalElements = 30;
    const si::ShapeFeatures f = si::features(d);
    CHECK(f.distinct == 15 && std::fabs(f.degMean - 2.0) < 1e-12);
    CHECK(std::fabs(f.deg2 - 70.0 / 30.0) < 1e-12);  // (5*4 + 5*9 + 5*1) / 30
    CHECK(f.q50 == 10 && f.maxSize == 10 && f.ubits == 4 && f.k == 3);

    // round trip: every field survives save + load
    std::vector<si::StoreEntry> st = {entry(1000, 20, 4.0, 512), entry(16384, 16, 9.5, 1024)};
    const std::string path = dir + "/store.json";
    CHECK(si::saveStore(path, st));
    auto back = si::loadStore(path);
    CHECK(back.size() == 2);
    for (int i = 0; i < 2; ++i) {
        CHECK(si::shapeKey(back[i].f) == si::shapeKey(st[i].f));
        CHECK(std::fabs(back[i].f.deg2 - st[i].f.deg2) < 1e-6 && back[i].f.distinct == st[i].f.distinct);
        CHECK(back[i].t.bitmap_min == st[i].t.bitmap_min && back[i].t.chunk_elems == st[i].t.chunk_elems &&
              back[i].t.use_graph == 1 && back[i].t.version == SI_TUNING_VERSION);
    }
    // replacing a record keeps one per key
    si::putEntry(back, entry(1000, 20, 4.0, 2048));
    CHECK(back.size() == 2 && back[0].t.bitmap_min == 2048);
    si::putEntry(back, entry(5000, 20, 4.0, 256));
    CHECK(back.size() == 3);

    // lookups: exact key, nearest within the distance, none beyond it
    double dist = -1;
    const si::StoreEntry* e = si::nearestEntry(st, st[1].f, 4.0, &dist);
    CHECK(e == &st[1] && dist == 0.0);
    si::ShapeFeatures g = st[0].f;
    g.k = 1100;  // about 0.14 in log2 from 1000
    g.total = 110000;
    e = si::nearestEntry(st, g, 4.0, &dist);
    CHECK(e == &st[0] && dist > 0.0 && dist < 1.0);
    g.q50 = 2000;  // the size distribution far away from both
    g.q90 = 8000;
    g.q99 = 32000;
    g.maxSize = 128000;
    CHECK(si::nearestEntry(st, g, 4.0) == nullptr);
    CHECK(si::nearestEntry({}, g, 4.0) == nullptr);

    // damaged files: a missing file, garbage, a truncated record, a zero knob
    // and a missing field are skipped; the good record still loads
    CHECK(si::loadStore(dir + "/missing.json").empty());
    {
        std::ofstream o(dir + "/garbage.json");
        o << "not json {{{ \"k= }";
    }
    CHECK(si::loadStore(dir + "/garbage.json").empty());
    std::string text;
    {
        std::ifstream i(path);
        std::getline(i, text, '\0');
    }
    const std::string good = text.substr(text.find("\"k=1000"), text.find("\n", text.find("\"k=1000")) - text.find("\"k=1000"));
    std::string zero = good, nofield = good;
    zero.replace(zero.find("\"probe_wg_per_cu\": 4"), 20, "\"probe_wg_per_cu\": 0");
    nofield.replace(nofield.find("\"deg2\""), 6, "\"degX\"");
    const std::string bad = "{\n  " + good.substr(0, good.size() / 2) + "\n  " +
                            zero.substr(0, zero.find("\"k=")) + zero.replace(0, 8, "\"k=2000") + "\n  " +
                            nofield.replace(0, 8, "\"k=3000") + "\n  " + good + "\n}\n";
    {
        std::ofstream o(dir + "/damaged.json");
        o << bad;
    }
    auto dam = si::loadStore(dir + "/damaged.json");
    CHECK(dam.size() == 1 && dam[0].t.probe_wg_per_cu == 4 && dam[0].f.k == 1000);
    std::printf("ok\n");
    return 0;
}
