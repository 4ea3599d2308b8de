// Internal declarations shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kbwh {

constexpr int kMaxInsert = 64;

// Sorted bit positions to insert into a compressed index (ascending), stored
// as the masks (1 << p) - 1.  Inserting a 0 at each position enumerates the
// indices whose listed bits are all 0; OR-ing `set` afterwards selects which
// of those bits are 1 (controls) and which stay 0 (the target of a pair).
struct Insert {
    uint32_t m;
    uint32_t pad;
    uint64_t mlo[kMaxInsert];
};

Insert make_insert(uint64_t positions_mask);

// Sampling levels: level 0 is the amplitude array, level l >= 1 holds
// count[l] = count[l-1] / 64 block sums.
struct Levels {
    const double *lev[16];
    uint64_t count[16];
    int nlev;
    int pad;
};

struct QubitList {
    uint32_t nq;
    uint8_t q[64];
};

// ---- launchers (kernels.hip) ---------------------------------------------
hipError_t launch_init(double2 *a, uint64_t size, hipStream_t st);
hipError_t launch_gate(double2 *a, uint32_t n, int kind, double angle, uint32_t target,
                       uint64_t ctrl_mask, hipStream_t st);
hipError_t launch_p1_partial(const double2 *a, uint32_t n, uint32_t target, double *partial,
                             uint32_t *nblocks, hipStream_t st);
hipError_t launch_sum_final(const double *partial, uint32_t m, double *out, hipStream_t st);
hipError_t launch_collapse(double2 *a, uint32_t n, uint32_t target, int result, double p,
                           hipStream_t st);
hipError_t launch_pauli_partial(const double2 *a, uint32_t n, uint64_t xmask, uint64_t zmask,
                                uint32_t ny, double *partial, uint32_t *nblocks, hipStream_t st);
uint32_t reduce_blocks(uint64_t count);
hipError_t launch_swap_bits(double2 *a, uint32_t n, uint32_t b_lo, uint32_t b_hi, hipStream_t st);
hipError_t launch_hier_levels(const double2 *a, uint32_t n, double *scratch, Levels *lv,
                              hipStream_t st);
hipError_t launch_hier_sample(const double2 *a, const Levels &lv, const double *uniforms,
                              uint64_t shots, const QubitList &ql, uint64_t *out, double *meta,
                              hipStream_t st);
hipError_t launch_dump_count(const double2 *a, uint64_t size, uint32_t *counts, hipStream_t st);
hipError_t launch_dump_write(const double2 *a, uint64_t size, const uint64_t *offsets,
                             const QubitList &ql, uint64_t *keys, double *re, double *im,
                             hipStream_t st);
constexpr uint32_t kDumpSeg = 1024;

// Gate coefficients computed on the host with libm, as the reference does.
struct Coef {
    double c, s;
    double2 ph0, ph1;
};
Coef make_coef(int kind, double angle);
// coef.cpp (built by the host compiler against glibc): {c, s, ph0.re, ph0.im, ph1.re, ph1.im}
void host_coef(int kind, double angle, double out[6]);

}  // namespace kbwh
