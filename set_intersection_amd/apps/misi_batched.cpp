// misi_batched -- all pair counts of a collection with the size-binned,
// persistent work-queue dispatcher (si_batched_count).  No single reference
// counterpart: it does the job of the ip/obs/hi/bitmap_dynamic executables on
// a multi-set dataset (same --input/--output/--partition contract, same result
// file: single-run order when k <= partition, else the multi-run order of
// io.hpp:104-136), with the technique chosen per size bin.
//   --tune            time candidate settings on this dataset first
//   --tune-file P     JSON store of tuning records keyed by dataset shape:
//                     read before the run, updated after --tune
//   --no-graph        launch the kernels directly instead of a hipGraph replay
//   --repeat N        run the count N times (reports the best time)
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "set_intersection.h"
#include "tiles.hpp"
#include "timers.hpp"

namespace {

int die(si_context* ctx) {
    std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
    if (ctx) si_close(ctx);
    return 1;
}

// Shape key: cardinality, universe bits, total and size quantiles.
std::string shapeKey(const si::Dataset& d) {
    std::vector<uint32_t> s(d.sizes);
    std::sort(s.begin(), s.end());
    auto q = [&](double f) { return s.empty() ? 0u : s[std::min<size_t>(s.size() - 1, size_t(f * s.size()))]; };
    uint32_t ubits = 0;
    while (ubits < 32 && (uint64_t(1) << ubits) <= d.universe) ++ubits;
    std::ostringstream o;
    o << "k=" << d.cardinality << ";ubits=" << ubits << ";total=" << d.totalElements << ";q50=" << q(0.5)
      << ";q90=" << q(0.9) << ";q99=" << q(0.99) << ";max=" << (s.empty() ? 0 : s.back());
    return o.str();
}

const char* kFields[] = {"bitmap_min", "ratio_search", "chunk_elems", "chunk_elems_bitmap", "probe_wg_per_cu",
                         "use_graph"};

uint32_t* field(si_tuning& t, int i) {
    uint32_t* f[] = {&t.bitmap_min, &t.ratio_search, &t.chunk_elems, &t.chunk_elems_bitmap, &t.probe_wg_per_cu,
                     &t.use_graph};
    return f[i];
}

// The store is one JSON object {"<shape>": {"field": value, ...}, ...}.
std::map<std::string, si_tuning> loadStore(const std::string& path) {
    std::map<std::string, si_tuning> m;
    std::ifstream f(path);
    if (!f) return m;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    size_t p = 0;
    while ((p = s.find("\"k=", p)) != std::string::npos) {
        const size_t e = s.find('"', p + 1);
        const std::string key = s.substr(p + 1, e - p - 1);
        const size_t ob = s.find('{', e), cb = s.find('}', ob);
        if (ob == std::string::npos || cb == std::string::npos) break;
        const std::string body = s.substr(ob, cb - ob);
        si_tuning t{};
        t.version = SI_TUNING_VERSION;
        for (int i = 0; i < 6; ++i) {
            const std::string name = std::string("\"") + kFields[i] + "\"";
            const size_t fp = body.find(name);
            if (fp != std::string::npos) *field(t, i) = uint32_t(std::stoul(body.substr(body.find(':', fp) + 1)));
        }
        m[key] = t;
        p = cb;
    }
    return m;
}

void saveStore(const std::string& path, const std::map<std::string, si_tuning>& m) {
    std::ofstream f(path);
    f << "{\n";
    size_t n = 0;
    for (auto& kv : m) {
        f << "  \"" << kv.first << "\": {";
        si_tuning t = kv.second;
        for (int i = 0; i < 6; ++i) f << (i ? ", " : "") << "\"" << kFields[i] << "\": " << *field(t, i);
        f << "}" << (++n < m.size() ? "," : "") << "\n";
    }
    f << "}\n";
}

}  // namespace

int main(int argc, char** argv) {
    si::printTitle("Batched (size-binned) GPU set intersection");
    si_context* ctx = nullptr;
    if (si_open(0, &ctx) != SI_OK) return die(nullptr);
    std::string input, output, tuneFile;
    uint32_t partition = 10000, repeat = 1;
    bool tune = false, noGraph = false;
    si::Options options(argv[0], "Help");
    options.add("input", "Input dataset path", &input)
        .add("output", "Output result path", &output)
        .add("partition", "Result-file order: single run if k <= partition, else the multi-run order", &partition)
        .add("tune", "Autotune the dispatcher on this dataset first", &tune)
        .add("tune-file", "JSON store of tuning records keyed by dataset shape", &tuneFile)
        .add("no-graph", "Launch kernels directly instead of replaying a hipGraph", &noGraph)
        .add("repeat", "Run the count N times", &repeat)
        .addFlag("help", "Print help");
    try {
        options.parse(argc, argv);
    } catch (const si::CliError& e) {
        std::printf("%s\n", e.what());
        si_close(ctx);
        return 1;
    }
    if (options.count("help")) {
        std::printf("%s\n", options.help().c_str());
        si_close(ctx);
        return 0;
    }
    if (!options.count("input")) {
        std::printf("%s\n", "No input dataset given! Exiting...");
        si_close(ctx);
        return 1;
    }
    si::Dataset d;
    try {
        d = si::readDataset(input);
    } catch (const si::FormatError& e) {
        std::printf("%s\n", e.what());
        si_close(ctx);
        return 1;
    }
    si::printTable("Dataset characteristics", {{"Cardinality", si::toText(d.cardinality)},
                                               {"Universe", si::toText(d.universe)},
                                               {"Total elements", si::toText(d.totalElements)}});
    si::DeviceTimer deviceTimer(ctx);
    si_dataset* ds = nullptr;
    if (si_dataset_upload(ctx, d.cardinality, d.universe, d.totalElements, d.sizes.data(), d.elements.data(), 1, &ds) !=
        SI_OK)
        return die(ctx);
    deviceTimer.addUpload(ds);

    const std::string key = shapeKey(d);
    si_tuning t;
    si_batched_default_tuning(ds, &t);
    std::map<std::string, si_tuning> store;
    bool fromStore = false;
    if (!tuneFile.empty()) {
        store = loadStore(tuneFile);
        auto it = store.find(key);
        if (it != store.end()) {
            t = it->second;
            fromStore = true;
        }
    }
    if (tune) {
        si::HostTimer ht;
        si::Interval* iv = ht.add("Autotune");
        if (si_batched_autotune(ds, 20000.0, &t) != SI_OK) return die(ctx);
        si::HostTimer::finish(iv);
        ht.print();
        if (!tuneFile.empty()) {
            store[key] = t;
            saveStore(tuneFile, store);
        }
    }
    if (noGraph) t.use_graph = 0;
    si::printTable("Dispatcher", {{"Shape", key.substr(0, 25)},
                                  {"Tuning", tune ? "autotuned" : fromStore ? "stored" : "default"},
                                  {"bitmap_min", si::toText(t.bitmap_min)},
                                  {"ratio_search", si::toText(t.ratio_search)},
                                  {"chunk_elems", si::toText(t.chunk_elems)},
                                  {"chunk_elems_bitmap", si::toText(t.chunk_elems_bitmap)},
                                  {"probe_wg_per_cu", si::toText(t.probe_wg_per_cu)},
                                  {"hipGraph", t.use_graph ? "yes" : "no"}});

    const uint64_t pairs = si::num_pairs(d.cardinality);
    si::EventPair* alloc = deviceTimer.add("Device memory allocation");
    si_buffer* dCounts = nullptr;
    if (si_buffer_alloc(ctx, std::max<uint64_t>(pairs, 1) * 4, &dCounts) != SI_OK) return die(ctx);
    si::DeviceTimer::finish(alloc);
    float best = 1e30f;
    for (uint32_t r = 0; r < std::max(1u, repeat); ++r) {
        si::EventPair* ev = deviceTimer.add("Batched intersection");
        if (si_batched_count_async(ds, &t, dCounts) != SI_OK) return die(ctx);
        si::DeviceTimer::finish(ev);
        float ms = 0.f;
        si_event_elapsed(ctx, ev->start, ev->end, &ms);
        best = std::min(best, ms);
    }
    std::vector<uint32_t> counts(pairs);
    si::EventPair* tr = deviceTimer.add("Transfer result");
    if (pairs && si_buffer_download(dCounts, 0, counts.data(), pairs * 4) != SI_OK) return die(ctx);
    si::DeviceTimer::finish(tr);
    si_dataset_free(ds);
    si_buffer_free(dCounts);
    si_sync(ctx);
    deviceTimer.print();
    std::printf("pairs: %llu  best: %.3f ms  pairs/s: %.4g\n", (unsigned long long)pairs, best,
                best > 0 ? pairs / (best / 1e3) : 0.0);
    if (!output.empty()) {
        std::printf("Writing result to file %s\n", output.c_str());
        if (partition == 0 || d.cardinality <= partition) {
            si::writeResult(d.cardinality, counts, output);
        } else {
            // multi-run order of the reference, counts looked up in triangular layout
            std::vector<si::tile_pair> runs = si::findTilePairs(d.cardinality, partition);
            si::LineWriter w(output);
            for (auto& run : runs) {
                const si::tile& A = run.first;
                const si::tile& B = run.second;
                const bool self = A.id == B.id;
                for (uint64_t i = A.start; i < A.end; ++i)
                    for (uint64_t j = self ? i + 1 : B.start; j < B.end; ++j)
                        w.line(i + 1, j + 1, counts[si::triangular_index(d.cardinality, i, j)]);
            }
            w.close();
        }
        std::printf("Finished\n");
    }
    si_close(ctx);
    return 0;
}
