// misi_batched -- all pair counts of a collection with the size-binned,
// persistent work-queue dispatcher (si_batched_count).  No single reference
// counterpart: it does the job of the ip/obs/hi/bitmap_dynamic executables on
// a multi-set dataset (same --input/--output/--partition contract, same result
// file: single-run order when k <= partition, else the multi-run order of
// io.hpp:104-136), with the technique chosen per size bin.
//   --tune            time candidate settings on this dataset first
//   --tune-file P     JSON store of tuning records with the dataset's shape
//                     features (set-size and element-degree statistics,
//                     host/tune_store.hpp): the record of the same shape, or
//                     else of the nearest one, is used (and seeds --tune);
//                     --tune writes this dataset's record back
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
#include "tune_store.hpp"

namespace {

int die(si_context* ctx) {
    std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
    if (ctx) si_close(ctx);
    return 1;
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

    si_tuning t;
    si_batched_default_tuning(ds, &t);
    std::string source = "default";
    si::ShapeFeatures shape;
    std::vector<si::StoreEntry> store;
    const si::StoreEntry* seed = nullptr;
    if (tune || !tuneFile.empty()) {
        shape = si::features(d);
        si::printTable("Shape features", {{"Set size q50/q90/q99", si::toText(shape.q50) + "/" + si::toText(shape.q90) +
                                                                       "/" + si::toText(shape.q99)},
                                          {"Set size max", si::toText(shape.maxSize)},
                                          {"Distinct elements", si::toText(shape.distinct)},
                                          {"Element degree mean", si::toText(shape.degMean)},
                                          {"Element degree 2nd moment", si::toText(shape.deg2)}});
    }
    if (!tuneFile.empty()) {
        store = si::loadStore(tuneFile);
        double dist = 0.0;
        seed = si::nearestEntry(store, shape, 4.0, &dist);
        if (seed) {
            t = seed->t;
            source = dist == 0.0 ? "stored" : "nearest stored (distance " + si::toText(dist) + ")";
        }
    }
    if (tune) {
        si::HostTimer ht;
        si::Interval* iv = ht.add("Autotune");
        const int st = seed ? si_batched_autotune_seeded(ds, &seed->t, 20000.0, &t) : si_batched_autotune(ds, 20000.0, &t);
        if (st != SI_OK) return die(ctx);
        si::HostTimer::finish(iv);
        ht.print();
        source = seed ? "autotuned from the stored record" : "autotuned";
        if (!tuneFile.empty()) {
            si::putEntry(store, si::StoreEntry{shape, t});
            if (!si::saveStore(tuneFile, store)) std::printf("cannot write %s\n", tuneFile.c_str());
        }
    }
    if (noGraph) t.use_graph = 0;
    si::printTable("Dispatcher", {{"Tuning", source},
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
