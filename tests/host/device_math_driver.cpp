// Host build of the kernels' own transcendental routines (geon_device.hpp)
// for the accuracy tests in tests/test_device_math.py.
#include "../../geon_amd/csrc/kernels/geon_device.hpp"

extern "C" void geon_test_sincos_cw(const double* x, long n, double* s, double* c) {
    for (long i = 0; i < n; ++i) geon::sincos_cw(x[i], &s[i], &c[i]);
}

extern "C" void geon_test_tanh_cw(const double* x, long n, double* out) {
    for (long i = 0; i < n; ++i) out[i] = geon::tanh_cw(x[i]);
}

// the FP32 path's sincos (geon_fp32.hpp); the GPU builtins it uses
// elsewhere get host stand-ins
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __expf expf
#include "../../geon_amd/csrc/kernels/geon_fp32.hpp"

extern "C" void geon_test_sincosf_cw(const float* x, long n, float* s, float* c) {
    for (long i = 0; i < n; ++i) geon::sincosf_cw(x[i], &s[i], &c[i]);
}
