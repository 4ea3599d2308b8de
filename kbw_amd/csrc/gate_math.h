// The reference's per-pair gate arithmetic (dense_v2.rs:166-277), operand for
// operand, shared by the per-gate HBM kernels and the single-workgroup LDS
// kernel of queued gates.  Included by HIP translation units compiled without
// FP contraction.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace kbwh {
namespace gm {

enum { KX = 0, KY = 1, KZ = 2, KH = 3, KP = 4, KRX = 5, KRY = 6, KRZ = 7 };

// num-complex `a *= b`: re = re*br - im*bi; im = im*br + re*bi
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    double2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.y * b.x + a.x * b.y;
    return r;
}

template <int KIND>
__device__ __forceinline__ void pair_math(double2 &k0, double2 &k1, double c, double s) {
    const double2 t0 = k0, t1 = k1;
    if constexpr (KIND == KX) {  // dense_v2.rs:168-170
        k0 = t1;
        k1 = t0;
    } else if constexpr (KIND == KY) {  // dense_v2.rs:178-182: swap, *= -i, *= i
        k0 = cmul(t1, make_double2(-0.0, -1.0));
        k1 = cmul(t0, make_double2(0.0, 1.0));
    } else if constexpr (KIND == KH) {  // dense_v2.rs:200-205
        const double f = 0.70710678118654752440;
        k0.x = (t0.x + t1.x) * f;
        k0.y = (t0.y + t1.y) * f;
        k1.x = (t0.x - t1.x) * f;
        k1.y = (t0.y - t1.y) * f;
    } else if constexpr (KIND == KRX) {  // dense_v2.rs:229-237
        k0.x = c * t0.x + s * t1.y;
        k0.y = c * t0.y - s * t1.x;
        k1.x = s * t0.y + c * t1.x;
        k1.y = -s * t0.x + c * t1.y;
    } else if constexpr (KIND == KRY) {  // dense_v2.rs:250-258
        k0.x = c * t0.x - s * t1.x;
        k0.y = c * t0.y - s * t1.y;
        k1.x = s * t0.x + c * t1.x;
        k1.y = s * t0.y + c * t1.y;
    }
}

// Diagonal gates on the |1> amplitude (Z, P) or on both (RZ):
// dense_v2.rs:191, 216, 271-272.
template <int KIND>
__device__ __forceinline__ void diag_math(double2 &k0, double2 &k1, double2 ph0, double2 ph1) {
    if constexpr (KIND == KZ) {
        k1 = cmul(k1, make_double2(-1.0, -0.0));
    } else if constexpr (KIND == KP) {
        k1 = cmul(k1, ph1);
    } else {
        k0 = cmul(k0, ph0);
        k1 = cmul(k1, ph1);
    }
}

}  // namespace gm
}  // namespace kbwh
