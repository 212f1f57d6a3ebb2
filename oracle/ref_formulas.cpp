// Thin C entry over the reference's generated closed-branch formulas (checker only, built into oracle/_ref/).
// The header itself stays in the reference tree; this file only calls it.
#include "closed_branch.cuh"

extern "C" void ref_closed_branch(const double* p8, double v1, double ph1, double v2, double ph2, double* o) {
    // p8 = (y, g1, g2, b1, b2, ksi, r1, a1); outputs in the scatter-table order (28 values)
    powsybl::openloadflow::updateClosedBranch(p8[0], p8[1], p8[2], p8[3], p8[4], v1, p8[6], v2, p8[5], p8[7], ph1, ph2, 0,
        o + 0, o + 1, o + 2, o + 3, o + 4, o + 5, o + 6, o + 7, o + 8, o + 9, o + 10, o + 11, o + 12, o + 13,
        o + 14, o + 15, o + 16, o + 17, o + 18, o + 19, o + 20, o + 21, o + 22, o + 23, o + 24, o + 25, o + 26, o + 27);
}
