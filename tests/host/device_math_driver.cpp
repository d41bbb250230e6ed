// Host build of the kernels' own transcendental routines (geon_device.hpp)
// for the accuracy tests in tests/test_device_math.py.
#include "../../geon_amd/csrc/kernels/geon_device.hpp"

extern "C" void geon_test_sincos_cw(const double* x, long n, double* s, double* c) {
    for (long i = 0; i < n; ++i) geon::sincos_cw(x[i], &s[i], &c[i]);
}

extern "C" void geon_test_tanh_cw(const double* x, long n, double* out) {
    for (long i = 0; i < n; ++i) out[i] = geon::tanh_cw(x[i]);
}
