// bitmap_dynamic -- per-set bitmap probing GPU set intersection; drop-in for
// the reference's src/bitmap_dynamic.cu executable (--warp default true).
#include "driver.hpp"

int main(int argc, char** argv) {
    return si::techniqueMain(argc, argv, {"Bitmap-dynamic GPU set intersection", SI_TECH_BITMAP_DYNAMIC,
                                          "Bitmap intersection", false, false, true});
}
