// Shared main() of the GPU technique executables (ip, obs, hi, bitmap_dynamic).
//
// Follows the reference drivers step by step (ip.cu:14-205, obs.cu:38-273,
// hi.cu:47-293, bitmap_dynamic.cu:21-238): title box, CLI, readDataset,
// "Dataset characteristics" and "Launch info" tables, runs from findTilePairs,
// per-run kernel + result transfer timed under the reference's labels, the
// "Device Timings (in ms)" table, and writeResult in the single-run or
// multi-run layout.  Several runs go through the transfer/write pipeline of
// pipeline.hpp.  All device work goes through the C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "pipeline.hpp"
#include "set_intersection.h"
#include "tiles.hpp"
#include "timers.hpp"

namespace si {

struct TechniqueSpec {
    const char* title;         // title box text
    int technique;             // si_technique
    const char* kernelLabel;   // DeviceTimer label of the intersection kernel
    bool hasWarpPath;          // --warp / --path flags (obs, hi)
    bool hasBuckets;           // --buckets / --bucket-size (hi)
    bool warpDefault;          // bitmap_dynamic declares --warp default true
};

inline int fail(const char* what, si_context* ctx = nullptr) {
    std::fprintf(stderr, "GPUassert: %s: %s\n", what, si_last_error());
    if (ctx) si_close(ctx);
    return 1;
}

inline int techniqueMain(int argc, char** argv, const TechniqueSpec& spec) {
    printTitle(spec.title);
    si_context* ctx = nullptr;
    if (si_open(0, &ctx) != SI_OK) return fail("si_open");
    si_device_info info;
    si_device_info_get(ctx, &info);

    std::string input, output;
    uint32_t blocks = uint32_t(info.multiprocessor_count) * 16;
    uint32_t blockSize = uint32_t(info.max_threads_per_block) / 2;
    uint32_t partition = 10000, buckets = 512, bucketSize = 1000;
    bool warpBased = spec.warpDefault, path = false;

    Options options(argv[0], "Help");
    options.add("input", "Input dataset path", &input)
        .add("output", "Output result path", &output)
        .add("blocks", "Number of blocks (default: " + std::to_string(blocks) + ")", &blocks)
        .add("threads", "Threads per block (default: " + std::to_string(blockSize) + ")", &blockSize);
    if (spec.hasBuckets)
        options.add("buckets", "Number of buckets (default: 512)", &buckets)
            .add("bucket-size", "Size of each bucket (default: 1000)", &bucketSize);
    if (spec.hasWarpPath || spec.warpDefault)
        options.add("warp", "Launch warp based kernel (default: false, runs block based)", &warpBased);
    options.add("partition", "Number of sets to be processed per GPU invocation", &partition);
    if (spec.hasWarpPath)
        options.add("path", "Adapt intersect path 1st level partitioning to distribute workload across thread blocks",
                    &path);
    options.addFlag("help", "Print help");
    try {
        options.parse(argc, argv);
    } catch (const CliError& e) {
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
    if (partition == 0) {
        std::printf("%s\n", "--partition must be positive");
        si_close(ctx);
        return 1;
    }

    Dataset d;
    try {
        d = readDataset(input);
    } catch (const FormatError& e) {
        std::printf("%s\n", e.what());
        si_close(ctx);
        return 1;
    }
    printTable("Dataset characteristics", {{"Cardinality", toText(d.cardinality)},
                                           {"Universe", toText(d.universe)},
                                           {"Total elements", toText(d.totalElements)}});

    std::vector<tile_pair> runs = findTilePairs(d.cardinality, partition);
    const uint64_t combinations =
        runs.size() == 1 ? num_pairs(d.cardinality) : uint64_t(partition) * partition;
    const uint64_t outputMemory = runs.size() * combinations * sizeof(uint32_t);
    const uint64_t deviceMemory = sizeof(uint32_t) * (uint64_t(d.cardinality) * 3 + d.totalElements + combinations);

    std::vector<std::pair<std::string, std::string>> launch = {
        {"Blocks", toText(blocks)},
        {"Block Size", toText(blockSize)},
        {"Warps per Block", toText(blockSize / 64)},
        {"Level", (warpBased ? "Warp" : "Block")}};
    if (spec.hasBuckets) {
        launch.push_back({"Buckets", toText(buckets)});
        launch.push_back({"Bucket Size", toText(bucketSize)});
    }
    launch.push_back({"Partition", toText(partition)});
    launch.push_back({"GPU invocations", toText(runs.size())});
    launch.push_back({"Required memory (Output)", formatBytes(outputMemory)});
    launch.push_back({"Required memory (GPU)", formatBytes(deviceMemory)});
    printTable("Launch info", launch);

    si_params p{};
    // explicit --blocks/--threads override the library's own launch geometry
    p.blocks = options.count("blocks") ? blocks : 0;
    p.threads = options.count("threads") ? blockSize : 0;
    p.warp = warpBased;
    p.path = path;
    p.buckets = buckets;
    p.bucket_size = bucketSize;

    DeviceTimer deviceTimer(ctx);
    const bool pipelined = runs.size() > 1;

    EventPair* devMemAlloc = deviceTimer.add("Device memory allocation");
    si_buffer* dCounts[2] = {nullptr, nullptr};
    for (int b = 0; b < (pipelined ? 2 : 1); ++b)
        if (si_buffer_alloc(ctx, std::max<uint64_t>(combinations, 1) * sizeof(uint32_t), &dCounts[b]) != SI_OK)
            return fail("si_buffer_alloc", ctx);
    DeviceTimer::finish(devMemAlloc);

    si_dataset* ds = nullptr;
    if (si_dataset_upload(ctx, d.cardinality, d.universe, d.totalElements, d.sizes.data(), d.elements.data(), 1,
                          &ds) != SI_OK)
        return fail("si_dataset_upload", ctx);
    deviceTimer.addUpload(ds);

    // one run: counts of all k(k-1)/2 pairs, written in the single-run layout
    std::vector<uint32_t> counts;
    auto launchRun = [&](const tile_pair& run, si_buffer* dst) -> int64_t {
        const tile& A = run.first;
        const tile& B = run.second;
        const bool selfJoin = A.id == B.id;
        const uint32_t numOfSets = selfJoin && runs.size() == 1 ? d.cardinality : partition;
        si_tile tA{A.id, A.start, A.end}, tB{B.id, B.start, B.end};
        EventPair* compute = deviceTimer.add(spec.kernelLabel);
        if (si_run_count_async(ds, spec.technique, &p, tA, tB, numOfSets, dst) != SI_OK) return -1;
        DeviceTimer::finish(compute);
        const uint64_t slots = selfJoin ? uint64_t(numOfSets) * (numOfSets - 1) / 2 : uint64_t(numOfSets) * numOfSets;
        return int64_t(std::min<uint64_t>(slots, combinations));
    };
    if (pipelined) {
        try {
            if (runPipelined<uint32_t>(ctx, runs, partition, dCounts, deviceTimer, nullptr, output,
                                       [&](size_t, const tile_pair& run, si_buffer* dst) {
                                           return launchRun(run, dst);
                                       }) != SI_OK)
                return fail("multi-run pipeline", ctx);
        } catch (const FormatError& e) {
            std::printf("%s\n", e.what());
            si_close(ctx);
            return 1;
        }
    } else {
        counts.resize(combinations);
        const int64_t used = launchRun(runs[0], dCounts[0]);
        if (used < 0) return fail("si_run_count_async", ctx);
        EventPair* countTransfer = deviceTimer.add("Transfer result");
        if (used && si_buffer_download(dCounts[0], 0, counts.data(), uint64_t(used) * sizeof(uint32_t)) != SI_OK)
            return fail("si_buffer_download", ctx);
        DeviceTimer::finish(countTransfer);
    }

    EventPair* freeDevMem = deviceTimer.add("Free device memory");
    si_dataset_free(ds);
    si_buffer_free(dCounts[0]);
    si_buffer_free(dCounts[1]);
    DeviceTimer::finish(freeDevMem);
    si_sync(ctx);
    deviceTimer.print();

    if (!output.empty()) {
        // the pipeline has written the multi-run file as the runs ended
        std::printf("Writing result to file %s\n", output.c_str());
        if (!pipelined) {
            try {
                writeResult(d.cardinality, counts, output);
            } catch (const FormatError& e) {
                std::printf("%s\n", e.what());
                si_close(ctx);
                return 1;
            }
        }
        std::printf("Finished\n");
    }
    si_close(ctx);
    return 0;
}

}  // namespace si
