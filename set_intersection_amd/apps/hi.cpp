// hi -- Hash-based GPU set intersection; drop-in for the reference's
// src/hi.cu executable (--buckets, --bucket-size, --warp, --path).
#include "driver.hpp"

int main(int argc, char** argv) {
    return si::techniqueMain(argc, argv, {"Hash-based GPU set intersection", SI_TECH_HI, "Hash intersection", true,
                                          true, false});
}
