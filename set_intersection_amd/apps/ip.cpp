// ip -- Intersect-Path (merge path) GPU set intersection; drop-in for the
// reference's src/ip.cu executable (same flags, tables and result file).
#include "driver.hpp"

int main(int argc, char** argv) {
    return si::techniqueMain(argc, argv, {"Intersect-Path GPU set intersection", SI_TECH_IP, "Intersect path",
                                          false, false, false});
}
