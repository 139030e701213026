// txt2bin -- text dataset (one set per line, space-separated, ascending) to
// the .bin format; counterpart of the reference's src/txt2bin.cpp:12-73.
// Usage: txt2bin <input.txt> <output-prefix>
// Writes <prefix>_asc.bin (sets ordered by size, ascending) and
// <prefix>_desc.bin (descending); universe = largest last element of a line.
// Differences: values are parsed as uint32 (the reference's std::stoi rejects
// values >= 2^31, txt2bin.cpp:46), the _asc file is ordered by size as its
// name says (the reference writes it in input order, txt2bin.cpp:23-25), and
// malformed lines are reported instead of throwing.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "format.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("Usage: %s <input.txt> <output-prefix>\n", argv[0]);
        return 1;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::printf("cannot open %s\n", argv[1]);
        return 1;
    }
    std::vector<std::vector<uint32_t>> sets;
    uint64_t total = 0;
    uint32_t universe = 0;
    std::string line;
    uint64_t lineNo = 0;
    while (std::getline(in, line)) {
        ++lineNo;
        std::istringstream ss(line);
        std::vector<uint32_t> s;
        std::string tok;
        bool any = false;
        while (ss >> tok) {
            any = true;
            char* end = nullptr;
            unsigned long long v = std::strtoull(tok.c_str(), &end, 10);
            if (*end || tok[0] == '-' || v > 0xFFFFFFFFull) {
                std::printf("line %llu: bad value '%s'\n", (unsigned long long)lineNo, tok.c_str());
                return 1;
            }
            s.push_back(uint32_t(v));
        }
        if (!any) continue;  // empty lines are skipped, as in the reference
        universe = std::max(universe, s.back());
        total += s.size();
        sets.push_back(std::move(s));
    }
    const std::string prefix = argv[2];
    std::stable_sort(sets.begin(), sets.end(), [](const auto& a, const auto& b) { return a.size() < b.size(); });
    std::string out = prefix + "_asc.bin";
    std::printf("Writing dataset to %s\n", out.c_str());
    si::writeDataset(uint32_t(sets.size()), universe, total, sets, out);
    std::printf("Sorting sets in descending order\n");
    std::stable_sort(sets.begin(), sets.end(), [](const auto& a, const auto& b) { return a.size() > b.size(); });
    out = prefix + "_desc.bin";
    std::printf("Writing dataset to %s\n", out.c_str());
    si::writeDataset(uint32_t(sets.size()), universe, total, sets, out);
    return 0;
}
