// Host build of the kernels' sincos_cw (geon_device.hpp) for the accuracy
// test in tests/test_sincos_cw.py.
#include "../../geon_amd/csrc/kernels/geon_device.hpp"

extern "C" void geon_test_sincos_cw(const double* x, long n, double* s, double* c) {
    for (long i = 0; i < n; ++i) geon::sincos_cw(x[i], &s[i], &c[i]);
}
