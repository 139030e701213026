// mm -- bitmap-GEMM set intersection; drop-in for the reference's
// src/matrix_multiplication.cu executable (target name `mm`, CMakeLists.txt:152).
// Same flags (--input, --output, --partition), tables, tile loop (:139-184) and
// result file (writeResult<T, true>, io.hpp:104-136), written run by run as
// the runs end (host/pipeline.hpp).  The incidence matrices
// are built on the device as bit rows and multiplied on the matrix cores:
// --variant auto (default) and i8 take the int8 MFMA path with int32
// accumulation (exact for every count below 2^31); fp4 (e2m1 MFMA) and f32
// (fp8 MFMA) accumulate in FP32 and are exact while counts stay below 2^24,
// as the reference's SGEMM (they refuse larger universes); popcount is the
// bitmap AND + popcount kernel.  fp4 is the fastest of them and is kept as a
// checked comparison beside the int8 path.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "pipeline.hpp"
#include "set_intersection.h"
#include "tiles.hpp"
#include "timers.hpp"

static int die(si_context* ctx) {
    std::fprintf(stderr, "GPUassert: %s\n", si_last_error());
    if (ctx) si_close(ctx);
    return 1;
}

int main(int argc, char** argv) {
    si::printTitle("GPU Matrix multiplication set intersection");
    si_context* ctx = nullptr;
    if (si_open(0, &ctx) != SI_OK) return die(nullptr);
    si_device_info info;
    si_device_info_get(ctx, &info);
    std::string input, output, variant = "auto";
    uint32_t partition = 10000;
    si::Options options(argv[0], "Help");
    options.add("input", "Input dataset path", &input)
        .add("output", "Output result path", &output)
        .add("partition", "Number of sets to be processed per GPU invocation", &partition)
        .add("variant", "auto (default) | fp4 | i8 | f32 | popcount", &variant)
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
    int32_t v = variant == "f32" ? SI_MM_F32 : variant == "popcount" ? SI_MM_POPCOUNT : variant == "i8" ? SI_MM_I8
              : variant == "fp4" ? SI_MM_FP4 : variant == "auto" ? SI_MM_I8 : -1;
    if (v < 0 || partition == 0) {
        std::printf("bad --variant or --partition\n");
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
    partition = std::min(d.cardinality, partition);
    if (partition == 0) partition = 1;
    std::vector<si::tile> tiles = si::splitToTiles(d.cardinality, partition);
    std::vector<si::tile_pair> runs = si::findTilePairs(d.cardinality, partition);
    const uint64_t combinations = uint64_t(partition) * partition;
    const uint64_t K = uint64_t(d.universe) + 1;
    const uint64_t deviceMemory = (K * partition * 2) + combinations * 4;
    si::printTable("Launch info", {{"Partition", si::toText(partition)},
                                   {"GPU invocations", si::toText(runs.size())},
                                   {"Required memory (Output)", si::formatBytes(runs.size() * combinations * 4)},
                                   {"Required memory (GPU)", si::formatBytes(deviceMemory)}});
    if (deviceMemory > info.free_global_mem) {
        std::printf("Error not enough GPU memory (%s)!\nExiting...", si::formatBytes(info.free_global_mem).c_str());
        si_close(ctx);
        return 1;
    }
    si::HostTimer hostTimer;
    si::DeviceTimer deviceTimer(ctx);

    si::EventPair* devAlloc = deviceTimer.add("Device memory allocation");
    si_buffer* dOut[2] = {nullptr, nullptr};
    for (auto& b : dOut)
        if (si_buffer_alloc(ctx, combinations * 4, &b) != SI_OK) return die(ctx);
    si::DeviceTimer::finish(devAlloc);
    si_dataset* ds = nullptr;
    if (si_dataset_upload(ctx, d.cardinality, d.universe, d.totalElements, d.sizes.data(), d.elements.data(), 1, &ds) !=
        SI_OK)
        return die(ctx);
    deviceTimer.addUpload(ds);

    // runs in findTilePairs order (tiles i <= j, matrix_multiplication.cu:139-150);
    // counts reach the host and the file through the multi-run pipeline
    int st = SI_OK;
    try {
        st = si::runPipelined<uint32_t, true>(
            ctx, runs, partition, dOut, deviceTimer, &hostTimer, output,
            [&](size_t, const si::tile_pair& run, si_buffer* dst) -> int64_t {
                const si::tile& A = run.first;
                const si::tile& B = run.second;
                si::EventPair* mmEv = deviceTimer.add("Matrix multiplication");
                if (si_mm_counts_async(ds, si_tile{A.id, A.start, A.end}, si_tile{B.id, B.start, B.end}, v,
                                       partition, dst) != SI_OK)
                    return -1;
                si::DeviceTimer::finish(mmEv);
                return int64_t(uint64_t(A.length - 1) * partition + B.length);
            });
    } catch (const si::FormatError& e) {
        std::printf("%s\n", e.what());
        si_close(ctx);
        return 1;
    }
    if (st != SI_OK) return die(ctx);
    si::EventPair* fr = deviceTimer.add("Free device memory");
    si_dataset_free(ds);
    si_buffer_free(dOut[0]);
    si_buffer_free(dOut[1]);
    si::DeviceTimer::finish(fr);
    si_sync(ctx);
    hostTimer.print();
    deviceTimer.print();
    if (!output.empty()) {
        // written by the pipeline as the runs ended
        std::printf("Writing result to file %s\n", output.c_str());
        std::printf("Finished\n");
    }
    si_close(ctx);
    return 0;
}
