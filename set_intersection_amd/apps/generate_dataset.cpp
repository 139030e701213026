// generate_dataset -- synthetic collections in the .bin format; counterpart of
// the reference's src/generate_dataset.cpp:21-139 (same flags and output
// names: <output>_asc_<tmp>.bin / <output>_desc_<tmp>.bin with
// tmp = <distribution>_<k>_<universe>[_<iterations>][_<param>]).
//
// Reference modes: each set is `iterations` draws with replacement from the
// distribution, deduplicated (generate_dataset.cpp:141-181):
//   uniform    values uniform in [1, universe]
//   normal     round(N(universe/2, stdev)), kept inside [0, universe]
//   bernoulli  every value of [1, universe] with probability --prob
//   zipf       Zipf(--skew) over [1, universe] (inverse CDF for universe <=
//              2^24, otherwise the continuous inverse-power approximation)
// Additions: --seed makes the output reproducible (the reference seeds from
// std::random_device, generate_dataset.cpp:149-150; seed 0 keeps that), and
// --workload writes one of the benchmark workloads with exact set sizes:
//   sisi_2p20   2 sets of 2^20 uniform from [0,2^24)
//   sisi_2p28   2 sets of 2^28 uniform from [0,2^32)
//   sisi_skew   |A| = 2^27, |B| = 2^16 uniform from [0,2^32)
//   misi_zipf   16384 sets, sizes Zipf(1.2) clamped to [16,65536], universe 2^20
//   mm_density  8192 sets over 2^16, density U[1%,10%] per set
//   rmat20      R-MAT scale 20, edge factor 16 (a=.57, b=c=.19), adjacency lists
#include <omp.h>

#include <memory>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cli.hpp"
#include "format.hpp"
#include "generate.hpp"

namespace {

std::vector<uint32_t> drawSet(const std::string& dist, uint64_t universe, uint32_t iterations, double param,
                              const si::ZipfSampler* zipfTable, uint64_t seed) {
    si::Rng rng(seed);
    std::vector<uint32_t> v;
    if (dist == "bernoulli") {
        const uint64_t thr = uint64_t(std::min(param, 1.0) * 18446744073709551615.0);
        for (uint64_t x = 1; x <= universe; ++x)
            if (rng.next() <= thr) v.push_back(uint32_t(x));
        return v;
    }
    v.reserve(iterations);
    for (uint32_t i = 0; i < iterations; ++i) {
        uint64_t x;
        if (dist == "uniform") {
            x = 1 + rng.below(universe);
        } else if (dist == "normal") {
            // Box-Muller
            const double u1 = std::max(rng.uniform01(), 1e-300), u2 = rng.uniform01();
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            const double r = std::round(double(universe) / 2 + param * z);
            if (r < 0 || r > double(universe)) continue;
            x = uint64_t(r);
        } else {  // zipf
            if (zipfTable) {
                x = (*zipfTable)(rng);
            } else {
                const double s = param, u = rng.uniform01();
                if (std::fabs(s - 1.0) < 1e-9)
                    x = uint64_t(std::exp(u * std::log(double(universe))));
                else
                    x = uint64_t(std::pow(u * (std::pow(double(universe), 1 - s) - 1) + 1, 1 / (1 - s)));
                x = std::min<uint64_t>(std::max<uint64_t>(x, 1), universe);
            }
        }
        v.push_back(uint32_t(x));
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

void writeBoth(std::vector<std::vector<uint32_t>>& sets, uint32_t universe, const std::string& prefix,
               const std::string& tmp) {
    uint64_t total = 0;
    for (auto& s : sets) total += s.size();
    std::printf("Sorting sets in ascending order\n");
    std::stable_sort(sets.begin(), sets.end(), [](const auto& a, const auto& b) { return a.size() < b.size(); });
    std::string out = prefix + "_asc_" + tmp;
    std::printf("Writing dataset to %s\n", out.c_str());
    si::writeDataset(uint32_t(sets.size()), universe, total, sets, out);
    std::printf("Sorting sets in descending order\n");
    std::stable_sort(sets.begin(), sets.end(), [](const auto& a, const auto& b) { return a.size() > b.size(); });
    out = prefix + "_desc_" + tmp;
    std::printf("Writing dataset to %s\n", out.c_str());
    si::writeDataset(uint32_t(sets.size()), universe, total, sets, out);
}

}  // namespace

int main(int argc, char** argv) {
    std::string output = "out", distribution = "uniform", workload;
    uint32_t iterations = 10000000, k = 100;
    double prob = 0.5, stdev = 2.0, skew = 0.9;
    uint64_t universe = 0, seed = 0;
    si::Options options(argv[0], "Generate dataset");
    options.add("distribution", "Set element distribution {uniform|normal|bernoulli|zipf} (default: uniform)", &distribution)
        .add("output", "Output prefix name for the generated dataset", &output)
        .add("prob", "Bernoulli probability", &prob)
        .add("stdev", "Standard Deviation for normal distribution (default: 2)", &stdev)
        .add("skew", "Theta (i.e. skew factor for zipf distribution)", &skew)
        .add("iterations", "Number of iterations (i.e. approximate set size, default: 10000000)", &iterations)
        .add("universe", "Universe cardinality", &universe)
        .add("k", "Dataset size (i.e. number of sets)", &k)
        .add("seed", "Random seed (0: std::random_device)", &seed)
        .add("workload", "Benchmark workload {sisi_2p20|sisi_2p28|sisi_skew|misi_zipf|mm_density|rmat20}", &workload)
        .addFlag("help", "Print help");
    try {
        options.parse(argc, argv);
    } catch (const si::CliError& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    if (options.count("help")) {
        std::printf("%s\n", options.help().c_str());
        return 0;
    }
    if (seed == 0) seed = (uint64_t(std::random_device{}()) << 32) ^ std::random_device{}();

    if (!workload.empty()) {
        std::vector<std::vector<uint32_t>> sets;
        uint32_t uni = 0;
        if (workload == "sisi_2p20" || workload == "sisi_2p28" || workload == "sisi_skew") {
            const uint64_t N = workload == "sisi_2p20" ? (1ull << 24) : (1ull << 32);
            const uint64_t na = workload == "sisi_2p20" ? (1ull << 20) : workload == "sisi_2p28" ? (1ull << 28) : (1ull << 27);
            const uint64_t nb = workload == "sisi_skew" ? (1ull << 16) : na;
            sets.resize(2);
#pragma omp parallel for num_threads(2)
            for (int i = 0; i < 2; ++i) sets[i] = si::uniformSorted(i ? nb : na, N, seed + uint64_t(i));
            uni = uint32_t(N - 1);
        } else if (workload == "misi_zipf") {
            si::ZipfSampler z(65536, 1.2);
            si::Rng r(seed);
            sets.resize(16384);
            std::vector<uint32_t> sz(16384);
            for (auto& x : sz) x = std::max<uint32_t>(16, z(r));
#pragma omp parallel for schedule(dynamic, 16)
            for (int i = 0; i < 16384; ++i) sets[i] = si::uniformSorted(sz[i], 1u << 20, seed + 1000003ull + uint64_t(i));
            uni = (1u << 20) - 1;
        } else if (workload == "mm_density") {
            sets.resize(8192);
#pragma omp parallel for schedule(dynamic, 16)
            for (int i = 0; i < 8192; ++i) {
                si::Rng r(seed + uint64_t(i));
                sets[i] = si::densitySet(1u << 16, 0.01 + 0.09 * r.uniform01(), seed * 31 + uint64_t(i));
            }
            uni = (1u << 16) - 1;
        } else if (workload == "rmat20") {
            si::Csr g = si::rmatGraph(20, 16, 0.57, 0.19, 0.19, seed);
            sets.resize(g.n);
            for (uint32_t v = 0; v < g.n; ++v) sets[v].assign(g.adj.begin() + g.offsets[v], g.adj.begin() + g.offsets[v + 1]);
            uni = g.n - 1;
            // vertex order matters for a graph: write it as is
            uint64_t total = g.adj.size();
            const std::string out = output + "_" + workload + ".bin";
            std::printf("Writing graph to %s\n", out.c_str());
            si::writeDataset(uint32_t(sets.size()), uni, total, sets, out);
            return 0;
        } else {
            std::printf("unknown workload '%s'\n", workload.c_str());
            return 1;
        }
        writeBoth(sets, uni, output, workload + ".bin");
        return 0;
    }

    if (!options.count("universe") || universe == 0 || universe > 0xFFFFFFFFull) {
        std::printf("%s\n", "No universe given! Exiting...");
        return 1;
    }
    std::string tmp = distribution + "_" + std::to_string(k) + "_" + std::to_string(universe);
    double param = 0;
    if (distribution == "uniform") {
        tmp += "_" + std::to_string(iterations);
    } else if (distribution == "normal") {
        tmp += "_" + std::to_string(iterations) + "_" + std::to_string(stdev);
        param = stdev;
    } else if (distribution == "bernoulli") {
        tmp += "_" + std::to_string(prob);
        param = prob;
    } else if (distribution == "zipf") {
        tmp += "_" + std::to_string(iterations) + "_" + std::to_string(skew);
        param = skew;
    } else {
        std::printf("unknown distribution '%s'\n", distribution.c_str());
        return 1;
    }
    tmp += ".bin";
    std::unique_ptr<si::ZipfSampler> ztab;
    if (distribution == "zipf" && universe <= (1u << 24)) ztab.reset(new si::ZipfSampler(uint32_t(universe), skew));
    std::vector<std::vector<uint32_t>> sets(k);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < int64_t(k); ++i)
        sets[i] = drawSet(distribution, universe, iterations, param, ztab.get(), seed + uint64_t(i));
    writeBoth(sets, uint32_t(std::min<uint64_t>(universe, 0xFFFFFFFFull)), output, tmp);
    std::printf("Finished!\n");
    return 0;
}
