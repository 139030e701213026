// obs -- Optimized Binary Search GPU set intersection; drop-in for the
// reference's src/obs.cu executable (--warp, --path, --partition).
#include "driver.hpp"

int main(int argc, char** argv) {
    return si::techniqueMain(argc, argv, {"OBS GPU set intersection", SI_TECH_OBS, "OBS intersection", true, false,
                                          false});
}
