// edges -- edge-wise intersection counts of a graph stored as a .bin
// collection (set v = sorted adjacency list of vertex v).  For every
// undirected edge (u, v), u < v, in CSR order: |N(u) ∩ N(v)|, written as
// "(u+1,v+1): count" lines (the reference's result-line format, io.hpp:97),
// plus the uint64 total (= 3 x triangles).  No reference counterpart.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "set_intersection.h"
#include "timers.hpp"

static int die(si_context* ctx) {
    std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
    if (ctx) si_close(ctx);
    return 1;
}

int main(int argc, char** argv) {
    si::printTitle("Edge-wise GPU set intersection");
    si_context* ctx = nullptr;
    if (si_open(0, &ctx) != SI_OK) return die(nullptr);
    std::string input, output;
    uint32_t repeat = 1, tuneMs = 0;
    si::Options options(argv[0], "Help");
    options.add("input", "Input graph (.bin, one adjacency list per vertex)", &input)
        .add("output", "Output result path", &output)
        .add("repeat", "Run the count N times", &repeat)
        .add("tune", "Autotune the dispatcher for MS milliseconds before planning (0: default tuning)", &tuneMs)
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
    std::vector<uint32_t> eu, ev;
    for (uint32_t a = 0; a < d.cardinality; ++a)
        for (uint64_t i = d.offsets[a]; i < d.offsets[a + 1]; ++i)
            if (d.elements[i] > a) {
                if (d.elements[i] >= d.cardinality) {
                    std::printf("vertex id %u out of range\n", d.elements[i]);
                    si_close(ctx);
                    return 1;
                }
                eu.push_back(a);
                ev.push_back(d.elements[i]);
            }
    si::printTable("Graph characteristics", {{"Vertices", si::toText(d.cardinality)},
                                             {"Adjacency entries", si::toText(d.totalElements)},
                                             {"Edges (u < v)", si::toText(eu.size())}});
    si::DeviceTimer deviceTimer(ctx);
    si_dataset* ds = nullptr;
    if (si_dataset_upload(ctx, d.cardinality, d.universe, d.totalElements, d.sizes.data(), d.elements.data(), 1, &ds) !=
        SI_OK)
        return die(ctx);
    deviceTimer.addUpload(ds);
    si::HostTimer ht;
    si_tuning tuning{};
    const bool tuned = tuneMs > 0;
    if (tuned) {
        si::Interval* tv = ht.add("Autotune");
        if (si_edge_autotune(ds, eu.data(), ev.data(), eu.size(), double(tuneMs), &tuning) != SI_OK) return die(ctx);
        si::HostTimer::finish(tv);
        si::printTable("Tuning", {{"bitmap_min", si::toText(tuning.bitmap_min)},
                                  {"chunk_elems", si::toText(tuning.chunk_elems)},
                                  {"chunk_elems_bitmap", si::toText(tuning.chunk_elems_bitmap)},
                                  {"probe_wg_per_cu", si::toText(tuning.probe_wg_per_cu)},
                                  {"ratio_search", si::toText(tuning.ratio_search)},
                                  {"edge_bitmap_halves", si::toText(tuning.edge_bitmap_halves)}});
    }
    si::Interval* iv = ht.add("Edge plan");
    si_edge_plan* plan = nullptr;
    if (si_edge_plan_create(ds, eu.data(), ev.data(), eu.size(), tuned ? &tuning : nullptr, &plan) != SI_OK)
        return die(ctx);
    si::HostTimer::finish(iv);
    si_buffer *dc = nullptr, *dt = nullptr;
    if (si_buffer_alloc(ctx, std::max<size_t>(eu.size(), 1) * 4, &dc) != SI_OK || si_buffer_alloc(ctx, 8, &dt) != SI_OK)
        return die(ctx);
    float best = 1e30f;
    for (uint32_t r = 0; r < std::max(1u, repeat); ++r) {
        si::EventPair* e = deviceTimer.add("Edge intersection");
        if (si_edge_count_async(plan, dc, dt) != SI_OK) return die(ctx);
        si::DeviceTimer::finish(e);
        float ms = 0.f;
        si_event_elapsed(ctx, e->start, e->end, &ms);
        best = std::min(best, ms);
    }
    std::vector<uint32_t> counts(eu.size());
    uint64_t total = 0;
    si::EventPair* tr = deviceTimer.add("Transfer result");
    if (!eu.empty() && si_buffer_download(dc, 0, counts.data(), eu.size() * 4) != SI_OK) return die(ctx);
    if (si_buffer_download(dt, 0, &total, 8) != SI_OK) return die(ctx);
    si::DeviceTimer::finish(tr);
    si_edge_plan_free(plan);
    si_buffer_free(dc);
    si_buffer_free(dt);
    si_dataset_free(ds);
    si_sync(ctx);
    ht.print();
    deviceTimer.print();
    std::printf("edges: %zu  total: %llu  triangles: %llu  best: %.3f ms  edges/s: %.4g\n", eu.size(),
                (unsigned long long)total, (unsigned long long)(total / 3), best,
                best > 0 ? eu.size() / (best / 1e3) : 0.0);
    if (!output.empty()) {
        std::printf("Writing result to file %s\n", output.c_str());
        si::LineWriter w(output);
        for (size_t e = 0; e < eu.size(); ++e) w.line(uint64_t(eu[e]) + 1, uint64_t(ev[e]) + 1, counts[e]);
        w.close();
        std::printf("Finished\n");
    }
    si_close(ctx);
    return 0;
}
