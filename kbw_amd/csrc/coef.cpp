// Gate coefficients, computed on the host with glibc's cos/sin exactly where
// and how the reference computes them, before its per-amplitude loop.
// Built by the host C++ compiler with -ffp-contract=off.
#include <math.h>

namespace kbwh {

enum { KP = 4, KRX = 5, KRY = 6, KRZ = 7 };

void host_coef(int kind, double angle, double out[6]) {
    for (int i = 0; i < 6; ++i) out[i] = 0.0;
    if (kind == KP) {
        // dense_v2.rs:212: Complex::exp(i * lambda) = from_polar(exp(0), lambda)
        out[4] = 1.0 * cos(angle);
        out[5] = 1.0 * sin(angle);
    } else if (kind == KRX || kind == KRY) {
        // dense_v2.rs:224-225, 245-246
        out[0] = cos(angle / 2.0);
        out[1] = sin(angle / 2.0);
    } else if (kind == KRZ) {
        // dense_v2.rs:266-267: exp(i * (-theta / 2)), exp(i * (theta / 2))
        const double x0 = -angle / 2.0, x1 = angle / 2.0;
        out[2] = 1.0 * cos(x0);
        out[3] = 1.0 * sin(x0);
        out[4] = 1.0 * cos(x1);
        out[5] = 1.0 * sin(x1);
    }
}

}  // namespace kbwh
