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

// sincos_acc_dw: the rounded values and their low words
extern "C" void geon_test_sincos_acc(const double* x, long n, double* s, double* sl, double* c,
                                     double* cl) {
    for (long i = 0; i < n; ++i) geon::sincos_acc_dw(x[i], &s[i], &sl[i], &c[i], &cl[i]);
}

// sin and cos of th[i] by rotation from the two-word values at th0[i], as an
// RK4 stage takes them (the rotation angle is th − th0 as computed)
extern "C" void geon_test_sincos_rot_dw(const double* th0, const double* th, long n, double* s,
                                        double* c) {
    for (long i = 0; i < n; ++i) {
        double s0, s0l, c0, c0l;
        geon::sincos_acc_dw(th0[i], &s0, &s0l, &c0, &c0l);
        geon::sincos_rot_dw(th[i], s0, s0l, c0, c0l, th[i] - th0[i], &s[i], &c[i]);
    }
}
