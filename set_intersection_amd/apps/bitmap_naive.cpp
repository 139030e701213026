// bitmap_naive -- bitmaps of all sets, AND + popcount for every pair; drop-in
// for the reference's src/bitmap_naive.cu executable (flags bitmap_naive.cu:36-41,
// tables :57-80, single-run result file :140-144).
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "set_intersection.h"
#include "timers.hpp"

int main(int argc, char** argv) {
    si::printTitle("Bitmap-naive GPU set intersection");
    si_context* ctx = nullptr;
    if (si_open(0, &ctx) != SI_OK) {
        std::fprintf(stderr, "GPUassert: si_open: %s\n", si_last_error());
        return 1;
    }
    si_device_info info;
    si_device_info_get(ctx, &info);
    std::string input, output;
    uint32_t blocks = uint32_t(info.multiprocessor_count) * 16, blockSize = uint32_t(info.max_threads_per_block) / 2;
    si::Options options(argv[0], "Help");
    options.add("input", "Input dataset path", &input)
        .add("output", "Output result path", &output)
        .add("blocks", "Number of blocks (default: " + std::to_string(blocks) + ")", &blocks)
        .add("threads", "Threads per block (default: " + std::to_string(blockSize) + ")", &blockSize)
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
    const uint64_t words = (uint64_t(d.universe) + 64) / 64;  // universe + 1 bits
    si::printTable("Bitmap Info", {{"Word size", "64"},
                                   {"Bitmap words", si::toText(words)},
                                   {"Required Memory (MB)", si::toText(words * 8 * d.cardinality / 1000000)}});
    const uint64_t pairs = si::num_pairs(d.cardinality);
    std::vector<uint32_t> counts(pairs);
    si::DeviceTimer deviceTimer(ctx);
    si::EventPair* alloc = deviceTimer.add("Device memory allocation");
    si_buffer* dCounts = nullptr;
    if (si_buffer_alloc(ctx, std::max<uint64_t>(pairs, 1) * 4, &dCounts) != SI_OK) {
        std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
        return 1;
    }
    si::DeviceTimer::finish(alloc);
    si_dataset* ds = nullptr;
    if (si_dataset_upload(ctx, d.cardinality, d.universe, d.totalElements, d.sizes.data(), d.elements.data(), 1, &ds) !=
        SI_OK) {
        std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
        return 1;
    }
    deviceTimer.addUpload(ds);
    si_params p{};
    p.blocks = options.count("blocks") ? blocks : 0;
    p.threads = options.count("threads") ? blockSize : 0;
    si::EventPair* inter = deviceTimer.add("Bitmap intersection");
    if (si_all_pairs_bitmap_async(ds, &p, dCounts) != SI_OK) {
        std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
        return 1;
    }
    si::DeviceTimer::finish(inter);
    si::EventPair* tr = deviceTimer.add("Transfer result");
    if (pairs && si_buffer_download(dCounts, 0, counts.data(), pairs * 4) != SI_OK) {
        std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
        return 1;
    }
    si::DeviceTimer::finish(tr);
    si::EventPair* fr = deviceTimer.add("Free device memory");
    si_dataset_free(ds);
    si_buffer_free(dCounts);
    si::DeviceTimer::finish(fr);
    si_sync(ctx);
    deviceTimer.print();
    if (!output.empty()) {
        std::printf("Writing result to file %s\n", output.c_str());
        si::writeResult(d.cardinality, counts, output);
        std::printf("Finished\n");
    }
    si_close(ctx);
    return 0;
}
