// Host build of the device sincos (olf-gpu_amd/csrc/trig.hpp) for the CPU test: sincos over an array.
#include "../../olf-gpu_amd/csrc/trig.hpp"

extern "C" void trig_sincos(int n, const double *x, double *s, double *c) {
    for (int i = 0; i < n; i++) olfhip::sincosReduced(x[i], s + i, c + i);
}
