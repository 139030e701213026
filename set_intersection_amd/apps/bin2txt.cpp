// bin2txt -- .bin dataset to text, one set per line; counterpart of the
// reference's src/bin2txt.cpp:10-37 (each value followed by a space, then a
// newline).  Values are printed unsigned (the reference prints them through
// an ostream_iterator<int>, bin2txt.cpp:33, so values >= 2^31 come out
// negative).  Usage: bin2txt <input.bin> <output-prefix>  -> <prefix>.txt
#include <cstdio>
#include <string>

#include "format.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("Usage: %s <input.bin> <output-prefix>\n", argv[0]);
        return 1;
    }
    si::Dataset d;
    try {
        d = si::readDataset(argv[1]);
    } catch (const si::FormatError& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    const std::string out = std::string(argv[2]) + ".txt";
    std::printf("Writing dataset to %s\n", out.c_str());
    FILE* f = std::fopen(out.c_str(), "w");
    if (!f) {
        std::printf("cannot create %s\n", out.c_str());
        return 1;
    }
    for (uint32_t s = 0; s < d.cardinality; ++s) {
        for (uint64_t i = d.offsets[s]; i < d.offsets[s + 1]; ++i) std::fprintf(f, "%u ", d.elements[i]);
        std::fputc('\n', f);
    }
    std::fclose(f);
    return 0;
}
