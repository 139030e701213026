// CPU technique executables of the reference, built from the oracle
// (test infrastructure and CPU baseline, see oracle.cpp):
//   stdintersection   std_set_intersection.cpp:9-91
//   serial_intersect  serial_intersect.cpp:11-91
//   simd              simd.cpp:14-98 (intersection restated, see oracle.cpp)
//   boost_bitset      boost_bitset.cpp:8-99
// Same flags (--input, --output, --threads default 4), same OpenMP split of
// the first set index, same HostTimer label and result file.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "timers.hpp"

extern "C" double oracle_all_pairs(int method, uint32_t k, uint32_t universe, const uint64_t* offsets,
                                   const uint32_t* elements, uint32_t threads, uint32_t* counts);

#if defined(CPU_METHOD_stdintersection)
static const int kMethod = 0;
static const char* kTitle = "CPU std::set_intersection";
static const char* kLabel = "std::set intersection";
#elif defined(CPU_METHOD_serial_intersect)
static const int kMethod = 1;
static const char* kTitle = "CPU serial intersection";
static const char* kLabel = "std::set intersection";
#elif defined(CPU_METHOD_boost_bitset)
static const int kMethod = 2;
static const char* kTitle = "CPU Boost bitset intersection";
static const char* kLabel = "Boost bitset intersection";
#else
static const int kMethod = 3;
static const char* kTitle = "CPU SIMD intersection";
static const char* kLabel = "std::set intersection";
#endif

int main(int argc, char** argv) {
    std::string input, output;
    uint32_t threads = 4;
    si::Options options(argv[0], "Help");
    options.add("input", "Input dataset path", &input)
        .add("output", "Output result path", &output)
        .add("threads", "Number of threads", &threads)
        .addFlag("help", "Print help");
    try {
        options.parse(argc, argv);
    } catch (const si::CliError& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    std::printf("%s\n", kTitle);
    if (options.count("help")) {
        std::printf("%s\n", options.help().c_str());
        return 0;
    }
    if (!options.count("input")) {
        std::printf("%s\n", "No input dataset given! Exiting...");
        return 1;
    }
    si::Dataset d;
    try {
        d = si::readDataset(input);
    } catch (const si::FormatError& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    si::printTable("Dataset characteristics", {{"Cardinality", si::toText(d.cardinality)},
                                               {"Universe", si::toText(d.universe)},
                                               {"Total elements", si::toText(d.totalElements)}});
    si::HostTimer hostTimer;
    std::vector<uint32_t> counts(si::num_pairs(d.cardinality));
    si::Interval* iv = hostTimer.add(kLabel);
    oracle_all_pairs(kMethod, d.cardinality, d.universe, d.offsets.data(), d.elements.data(),
                     threads ? threads : 1, counts.data());
    si::HostTimer::finish(iv);
    hostTimer.print();
    if (!output.empty()) {
        std::printf("Writing result to file %s\n", output.c_str());
        si::writeResult(d.cardinality, counts, output);
        std::printf("Finished\n");
    }
    return 0;
}
