// Device side of the tile machinery: the tile layout, the gate record and
// the per-amplitude gate application in registers.  Self-contained (no
// standard-library or host headers), so the same text also compiles at run
// time into pass-specialised kernels (tile_jit.hip).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
// hiprtc declares the fixed-width integers in its own namespace
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#endif

// Exact gate arithmetic rounds every product and sum on its own, as the
// reference does.  A unit that defines KBWH_TILE_CONTRACT (the dependency-
// ordered runs, equal to the reference only to rounding anyway) lets the
// compiler fuse a product into the following add.
#ifdef KBWH_TILE_CONTRACT
#pragma clang fp contract(fast)
#else
#pragma clang fp contract(off)
#endif

namespace kbwh {
namespace tile {
// Internal linkage: a translation unit may be built with its own tile width
// (KBWH_TILE_BITS), and no tile type crosses translation units.
namespace {

// Tile shape, measured on MI355X for the 30-qubit brickwork circuit: 2^11
// amplitudes per tile (8 per thread), one tile per workgroup and 4 waves per
// SIMD beat 2^12-amplitude tiles and an LDS-staged persistent pipeline.
#ifndef KBWH_TILE_BITS
#define KBWH_TILE_BITS 11
#endif
#ifndef KBWH_TILE_WAVES
#define KBWH_TILE_WAVES 4
#endif
#ifndef KBWH_TILE_TBITS
#define KBWH_TILE_TBITS 8
#endif
constexpr int TL = KBWH_TILE_BITS;   // tile qubits
constexpr int TB = KBWH_TILE_TBITS;  // thread bits: 6 lane bits, then TB - 6 wave bits
constexpr int TT = 1 << TB;          // threads per tile
constexpr int TR = 1 << (TL - TB);   // amplitudes per thread (register bits = TL - TB)
static_assert(TB >= 6 && TL - TB >= 1 && TL - TB <= 4, "tile geometry");
// register bits that gate targets and controls may occupy (a unit that keeps
// the top register bit for its own use lowers this)
#ifndef KBWH_TILE_GATE_RBITS
#define KBWH_TILE_GATE_RBITS (KBWH_TILE_BITS - KBWH_TILE_TBITS)
#endif
constexpr int TGR = KBWH_TILE_GATE_RBITS;
static_assert(TGR >= 1 && TGR <= TL - TB, "gate register bits");
constexpr uint32_t kOutside = 255;
#ifndef KBWH_TILE_CHUNK
#define KBWH_TILE_CHUNK 4
#endif
constexpr int kChunk = KBWH_TILE_CHUNK < TR ? KBWH_TILE_CHUNK : TR;  // register slots per LDS exchange round

struct TGate {
    uint32_t kind;
    uint32_t tpos;  // tile position of the target, kOutside if the target is not in Q
    uint32_t cm;    // controls inside the tile, as tile positions
    uint32_t pad;
    uint64_t ocm;   // controls outside the tile, as global bits
    uint64_t otb;   // the target's global bit (diagonal gates with target outside Q)
    // pair gates: out coefficients (uniform part u, part times the half's sign s)
    double au, as, bu, bs, cru, crs, ciu, cis, d;
    // diagonal gates: the phases of the |0> and |1> halves
    double ph0x, ph0y, ph1x, ph1y;
    // diagonal gates: the thread bits and the register-slot bits that must be
    // 1 (Z and P: target and controls; RZ: controls), set by make_tgate
    uint32_t dlo, dhi;
};
static_assert(sizeof(TGate) == 144, "TGate layout");

struct TileSpec {
    uint64_t qbit[TL];  // qbit[b] = 1 << Q[b]
    uint64_t qlo[TL];   // (1 << p) - 1 for the positions p of Q in ascending order
};

enum { KX = 0, KY = 1, KZ = 2, KH = 3, KP = 4, KRX = 5, KRY = 6, KRZ = 7 };

// Amplitudes in the tile kernel are clang vector values (selects on them are
// value selects; HIP's double2 class would turn a ?: into a select between
// addresses and push the register array out to scratch memory).
typedef double dv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ dv2 mk2(double x, double y) {
    dv2 r;
    r.x = x;
    r.y = y;
    return r;
}

__device__ __forceinline__ dv2 cmul(dv2 a, dv2 b) {
    dv2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.y * b.x + a.x * b.y;
    return r;
}

// m = 0 or the sign bit: flips the sign of a double / of both parts when set
__device__ __forceinline__ double flip_sign_d(double x, uint32_t m) {
    return __longlong_as_double(__double_as_longlong(x) ^ (long long)((unsigned long long)m << 32));
}
__device__ __forceinline__ dv2 flip_sign(dv2 x, uint32_t m) {
    return mk2(flip_sign_d(x.x, m), flip_sign_d(x.y, m));
}

__device__ __forceinline__ dv2 shfl_xor2(dv2 v, int mask) {
    return mk2(__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64));
}

// Every pair closure of the reference (dense_v2.rs:166-263), written for one
// output amplitude from its own value m and its partner's value o:
//     out.re = ((A m.re + B o.re) + Cr o.im) * D
//     out.im = ((A m.im + B o.im) + Ci o.re) * D
// with coefficients that depend on the half (sigma = +1 for the |1> amplitude,
// -1 for the |0> one):
//     X:  A=0     B=1        Cr=0       Ci=0      D=1
//     Y:  A=0     B=0        Cr=-sigma  Ci=sigma  D=1
//     H:  A=-sigma B=1       Cr=0       Ci=0      D=1/sqrt2
//     RX: A=c     B=0        Cr=s       Ci=-s     D=1
//     RY: A=c     B=sigma*s  Cr=0       Ci=0      D=1
// Adding a zero product, multiplying by 1 and taking a product by +-1 are
// exact, and the order of two addends does not change a rounded sum, so each
// output equals the reference's closure in value (the sign of a zero aside).
struct PairCoef {
    double A, B, Cr, Ci, D;
};

__device__ __forceinline__ PairCoef pair_coef(const TGate &G, double sigma) {
    PairCoef k;
    k.A = G.au + G.as * sigma;
    k.B = G.bu + G.bs * sigma;
    k.Cr = G.cru + G.crs * sigma;
    k.Ci = G.ciu + G.cis * sigma;
    k.D = G.d;
    return k;
}

__device__ __forceinline__ dv2 pair_out(dv2 m, dv2 o, const PairCoef &k) {
    return mk2(((k.A * m.x + k.B * o.x) + k.Cr * o.y) * k.D, ((k.A * m.y + k.B * o.y) + k.Ci * o.x) * k.D);
}

// Uncontrolled gates take shorter forms, each equal in value to the generic
// form above (and to the reference's closure):
//     X:  out = o                                   (exchange)
//     Y:  |0>: (o.im, -o.re), |1>: (-o.im, o.re)    (exchange and sign flips)
//     H:  |0>: (m + o) k,     |1>: (o - m) k         (k = 1/sqrt2)
//     RX: (c m.re + s o.im, c m.im - s o.re)
//     RY: c m + sigma s o
// RY (PK_TWO) is out = A m + B o and RX (PK_RX) out = A m + Cr q with
// q = (o.im, -o.re), with no product by D = 1; a path each, so no amplitude
// picks its partner's form at run time.  Controlled gates other than X use
// the generic form (PK_GEN).  Every path is chosen by uniform two-way tests on
// the gate record, never by a multi-way branch (see apply_pair_t).
enum { PK_X = 0, PK_TWO = 1, PK_GEN = 2, PK_Y = 3, PK_H = 4, PK_RX = 5, PK_HU = 6 };
// An uncontrolled H in a compiled pass may leave out its factor 1/sqrt2
// (PK_HU: m + o and o - m), which the pass then folds into its pending
// phase: a factor common to every amplitude of the tile commutes with every
// gate.  gate_at takes it as this pseudo-kind.
constexpr uint32_t KH_UNSCALED = 8;

// Control test for the amplitude in register slot r of thread tid, whose tile
// index is tid + TT r: the thread bits are tested once per gate (tok) and the
// slot bits against the uniform part of cm, so no per-slot index is kept.
__device__ __forceinline__ bool thread_ctrl_ok(uint32_t tid, uint32_t cm) {
    const uint32_t lo = cm & (uint32_t)(TT - 1);
    return (tid & lo) == lo;
}
__device__ __forceinline__ bool slot_ctrl_ok(bool tok, int r, uint32_t cm) {
    const uint32_t hi = cm >> TB;
    return tok && (((uint32_t)r & hi) == hi);
}

template <int PK>
__device__ __forceinline__ dv2 pair_val(dv2 m, dv2 o, bool one, const TGate &G) {
    if constexpr (PK == PK_X) {
        return o;
    } else if constexpr (PK == PK_Y) {
        return one ? mk2(-o.y, o.x) : mk2(o.y, -o.x);
    } else if constexpr (PK == PK_H) {
        const double sx = one ? o.x - m.x : m.x + o.x;
        const double sy = one ? o.y - m.y : m.y + o.y;
        return mk2(sx * G.d, sy * G.d);
    } else if constexpr (PK == PK_HU) {
        return mk2(one ? o.x - m.x : m.x + o.x, one ? o.y - m.y : m.y + o.y);
    } else if constexpr (PK == PK_TWO) {
        // RY's record has as = bu = 0 (set_coefficients): A = au + as sigma is
        // au and B = bu + bs sigma is sigma bs, exactly
        const double B = one ? G.bs : -G.bs;
        return mk2(G.au * m.x + B * o.x, G.au * m.y + B * o.y);
    } else if constexpr (PK == PK_RX) {
        // RX's record has as = crs = 0: A = au and Cr = cru for both halves
        return mk2(G.au * m.x + G.cru * o.y, G.au * m.y + G.cru * -o.x);
    } else {
        return pair_out(m, o, pair_coef(G, one ? 1.0 : -1.0));
    }
}

// pair_val where the half varies across lanes: sg is the sign bit when the
// amplitude is in the |1> half.  H and Y take their half from sign flips
// instead of computing both forms and selecting: -m + o is o - m exactly, and
// (-o.y, o.x) / (o.y, -o.x) are sign flips of the partner.
template <int PK>
__device__ __forceinline__ dv2 pair_val_lane(dv2 m, dv2 o, bool one, uint32_t sg, const TGate &G) {
    if constexpr (PK == PK_Y) {
        return mk2(flip_sign_d(o.y, sg), flip_sign_d(o.x, sg ^ 0x80000000u));
    } else if constexpr (PK == PK_H) {
        return mk2((flip_sign_d(m.x, sg) + o.x) * G.d, (flip_sign_d(m.y, sg) + o.y) * G.d);
    } else if constexpr (PK == PK_HU) {
        return mk2(flip_sign_d(m.x, sg) + o.x, flip_sign_d(m.y, sg) + o.y);
    } else if constexpr (PK == PK_TWO) {
        const double B = flip_sign_d(G.bs, sg ^ 0x80000000u);
        return mk2(G.au * m.x + B * o.x, G.au * m.y + B * o.y);
    } else {
        return pair_val<PK>(m, o, one, G);
    }
}

// The in-tile controls of a pair gate: the record's cm, or CMT when the
// kernel is compiled for its pass (CMT >= 0).  With the mask known at
// compile time, the register-slot part of every control test is a constant:
// slots whose register controls fail are left out altogether (no exchange,
// no arithmetic, no select), and without thread controls no select remains.
template <int64_t CMT>
__device__ __forceinline__ uint32_t ctrl_mask(const TGate &G) {
    if constexpr (CMT >= 0)
        return (uint32_t)CMT;
    else
        return G.cm;
}
// Compiled kernels: a gate's slots in groups of 4.  Scheduled across a whole
// tile, the 16 slots' products and exchanges kept their temporaries live at
// once (up to 132 registers where 128 give 4 waves per SIMD).
__device__ __forceinline__ void slot_group(int r) {
    if (r > 0 && (r & 3) == 0) __builtin_amdgcn_sched_barrier(0);
}

// slot r takes part in a gate whose register-slot controls are cm >> TB
__device__ __forceinline__ constexpr bool slot_in(int r, uint32_t cm) {
    return ((uint32_t)r & (cm >> TB)) == (cm >> TB);
}

// Target on register bit RB: both amplitudes of a pair live in one thread.
// CTRL: some control lies inside the tile, so only the pairs whose index
// matches cm change.
template <int RB, int PK, bool CTRL, int64_t CMT = -1>
__device__ __forceinline__ void reg_pair(dv2 (&v)[TR], uint32_t tid, const TGate &G) {
    // An exchange has no instructions of its own: without this marker the
    // optimizer folds the register-bit variants of X into selects between
    // every candidate partner.
    if constexpr (PK == PK_X) asm volatile("");
    const uint32_t cm = ctrl_mask<CMT>(G);
    const bool tok = thread_ctrl_ok(tid, cm);
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        if constexpr (CMT >= 0) slot_group(r);
        if ((r >> RB) & 1) continue;
        if (CTRL && CMT >= 0 && !slot_in(r, cm)) continue;
        const int r1 = r | (1 << RB);
        const dv2 t0 = v[r], t1 = v[r1];
        const dv2 o0 = pair_val<PK>(t0, t1, false, G);
        const dv2 o1 = pair_val<PK>(t1, t0, true, G);
        if constexpr (CTRL) {
            const bool ok = slot_ctrl_ok(tok, r, cm);
            v[r] = ok ? o0 : t0;
            v[r1] = ok ? o1 : t1;
        } else {
            v[r] = o0;
            v[r1] = o1;
        }
    }
}

// Lane exchanges.  Lane layouts pinned on gfx950:
//   xor 1, 2:  DPP quad_perm 0xB1 / 0x4E gives lane i the value of lane i^1 / i^2
//   xor 8:     DPP row_ror:8 (rotate each 16-lane row by 8) gives lane i^8
//   xor 4:     row_half_mirror (lane i^7 within each 8-lane half row), then
//              quad_perm [3,2,1,0] (lane i^3): two moves, no select
//   xor 16/32: ds_bpermute.  v_permlane16/32_swap of a register with itself
//              (then a select on the lane bit) was measured slower: RY on
//              lane bit 4 or 5 took 510 / 655 us per 30-qubit pass against
//              338 / 399 us.
enum { LX_Q1 = 0, LX_Q2 = 1, LX_R8 = 2, LX_R4 = 3, LX_SHFL = 4 };

template <int LX>
__device__ __forceinline__ unsigned lane_xor_u(unsigned x) {
    if constexpr (LX == LX_Q1) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (LX == LX_Q2) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (LX == LX_R8) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);
    } else {
        const int m7 = __builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, false);
        return (unsigned)__builtin_amdgcn_mov_dpp(m7, 0x1B, 0xF, 0xF, false);
    }
}

template <int LX>
__device__ __forceinline__ double lane_xor_d(double x, int lm) {
    if constexpr (LX == LX_SHFL) {
        return __shfl_xor(x, lm, 64);
    } else {
        const unsigned long long b = (unsigned long long)__double_as_longlong(x);
        const unsigned lo = lane_xor_u<LX>((unsigned)b);
        const unsigned hi = lane_xor_u<LX>((unsigned)(b >> 32));
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
}

template <int PK, bool CTRL, int LX, int64_t CMT = -1>
__device__ __forceinline__ void lane_pair(dv2 (&v)[TR], uint32_t tid, uint32_t tp, const TGate &G) {
    const uint32_t cm = ctrl_mask<CMT>(G);
    const bool tok = thread_ctrl_ok(tid, cm);
    const bool one = ((tid >> tp) & 1u) != 0;
    const uint32_t sg = one ? 0x80000000u : 0u;
    const int lm = 1 << tp;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        if constexpr (CMT >= 0) slot_group(r);
        if (CTRL && CMT >= 0 && !slot_in(r, cm)) continue;
        const dv2 o = mk2(lane_xor_d<LX>(v[r].x, lm), lane_xor_d<LX>(v[r].y, lm));
        const dv2 out = pair_val_lane<PK>(v[r], o, one, sg, G);
        if constexpr (CTRL) {
            v[r] = slot_ctrl_ok(tok, r, cm) ? out : v[r];
        } else {
            v[r] = out;
        }
    }
}

template <int PK, bool CTRL>
__device__ __forceinline__ void thread_pair(dv2 (&v)[TR], dv2 *lds, uint32_t tid, uint32_t tp,
                                            const TGate &G) {
    const uint32_t cm = G.cm;
    const bool tok = thread_ctrl_ok(tid, cm);
    const bool one = ((tid >> tp) & 1u) != 0;
    if (tp < 6) {
        // dispatched on the position's bits (no multi-way branch; see apply_pair_t)
        if (tp & 4u) {
            lane_pair<PK, CTRL, LX_SHFL>(v, tid, tp, G);
        } else if (tp & 2u) {
            if (tp & 1u)
                lane_pair<PK, CTRL, LX_R8>(v, tid, tp, G);
            else
                lane_pair<PK, CTRL, LX_R4>(v, tid, tp, G);
        } else if (tp & 1u) {
            lane_pair<PK, CTRL, LX_Q2>(v, tid, tp, G);
        } else {
            lane_pair<PK, CTRL, LX_Q1>(v, tid, tp, G);
        }
        return;
    }
    // wave bits: through LDS, kChunk register slots at a time
    const uint32_t pt = tid ^ (1u << tp);
#pragma unroll
    for (int q = 0; q < TR; q += kChunk) {
        __syncthreads();  // earlier exchange reads are done
#pragma unroll
        for (int r = 0; r < kChunk; ++r) lds[tid + (r << TB)] = v[q + r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kChunk; ++r) {
            const dv2 o = lds[pt + (r << TB)];
            const dv2 out = pair_val_lane<PK>(v[q + r], o, one, one ? 0x80000000u : 0u, G);
            if constexpr (CTRL) {
                v[q + r] = slot_ctrl_ok(tok, q + r, cm) ? out : v[q + r];
            } else {
                v[q + r] = out;
            }
        }
    }
}

// The wave-bit part of thread_pair on its own, for the compile-time dispatch
// (pair_at): target on wave bit tp (6 <= tp < TB).
template <int PK, bool CTRL, int64_t CMT = -1>
__device__ __forceinline__ void wave_pair(dv2 (&v)[TR], dv2 *lds, uint32_t tid, uint32_t tp, const TGate &G) {
    const uint32_t cm = ctrl_mask<CMT>(G);
    const bool tok = thread_ctrl_ok(tid, cm);
    const bool one = ((tid >> tp) & 1u) != 0;
    const uint32_t pt = tid ^ (1u << tp);
    // with the controls known, only the slots whose register controls hold
    // go through LDS (a partner's slot r has the same register bits), and a
    // chunk without such slots is skipped with its barriers (the same for
    // every thread)
    constexpr bool known = CTRL && CMT >= 0;
#pragma unroll
    for (int q = 0; q < TR; q += kChunk) {
        if (known) {
            bool any = false;
#pragma unroll
            for (int r = 0; r < kChunk; ++r) any = any || slot_in(q + r, cm);
            if (!any) continue;
        }
        __syncthreads();  // earlier exchange reads are done
#pragma unroll
        for (int r = 0; r < kChunk; ++r)
            if (!known || slot_in(q + r, cm)) lds[tid + (r << TB)] = v[q + r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kChunk; ++r) {
            if (known && !slot_in(q + r, cm)) continue;
            const dv2 o = lds[pt + (r << TB)];
            const dv2 out = pair_val_lane<PK>(v[q + r], o, one, one ? 0x80000000u : 0u, G);
            if constexpr (CTRL) {
                v[q + r] = slot_ctrl_ok(tok, q + r, cm) ? out : v[q + r];
            } else {
                v[q + r] = out;
            }
        }
    }
}

template <int PK, bool CTRL>
__device__ __forceinline__ void apply_pair_t(dv2 (&v)[TR], dv2 *lds, uint32_t tid, const TGate &G) {
    const uint32_t tp = G.tpos;
    if (tp >= (uint32_t)TB) {
        // Register bit rb = tp - TB, dispatched on its bits.  A `switch (tp)`
        // here was lowered by the gfx950 backend into a compare tree whose
        // default case had two predecessors, and the structurized code never
        // ran that case (the last register bit's gates were dropped).
        const uint32_t rb = tp - TB;
        if constexpr (TL - TB > 2) {
            if (rb & 2u) {
                if constexpr (TL - TB > 3) {
                    if (rb & 1u)
                        reg_pair<3, PK, CTRL>(v, tid, G);
                    else
                        reg_pair<2, PK, CTRL>(v, tid, G);
                } else {
                    reg_pair<2, PK, CTRL>(v, tid, G);
                }
                return;
            }
        }
        if constexpr (TL - TB > 1) {
            if (rb & 1u) {
                reg_pair<1, PK, CTRL>(v, tid, G);
                return;
            }
        }
        reg_pair<0, PK, CTRL>(v, tid, G);
        return;
    }
    thread_pair<PK, CTRL>(v, lds, tid, tp, G);
}

__device__ __forceinline__ void apply_pair(dv2 (&v)[TR], dv2 *lds, uint32_t tid, const TGate &G) {
    if (G.kind == KX) {
        if (G.cm == 0)
            apply_pair_t<PK_X, false>(v, lds, tid, G);
        else
            apply_pair_t<PK_X, true>(v, lds, tid, G);
        return;
    }
    if (G.cm != 0)
        apply_pair_t<PK_GEN, true>(v, lds, tid, G);
    else if (G.kind == KY)
        apply_pair_t<PK_Y, false>(v, lds, tid, G);
    else if (G.kind == KH)
        apply_pair_t<PK_H, false>(v, lds, tid, G);
    else if (G.kind == KRX)
        apply_pair_t<PK_RX, false>(v, lds, tid, G);
    else
        apply_pair_t<PK_TWO, false>(v, lds, tid, G);
}

// Z, P and RZ: each amplitude is multiplied by the phase of its half
// (dense_v2.rs:190-191, 215-216, 271-272); Z and P leave the |0> half alone.
// Z's product by -1 is the negation in value, a flip of the two sign bits.
//
// Z and P change the amplitudes whose target and control bits are all 1.
// Those bits split into thread bits (a sign mask or a phase per thread),
// register bits (a mask MH of register slots, dispatched once per gate to code
// that touches exactly the slots r with r & MH == MH) and bits outside the
// tile (fixed per tile: the gate is skipped where one is 0).  A thread that is
// not selected multiplies by exactly 1 + 0i, which returns the value (a zero
// may change sign).  RZ changes every slot whose controls hold: the phase
// comes per thread (target on a thread bit) or per slot (target on a register
// bit, or outside the tile).  No per-amplitude branch, no per-slot branch.
enum { DK_Z = 0, DK_P = 1, DK_RZ = 2 };

template <int N>
struct IC {
    static constexpr int value = N;
};

// f(IC<m>()) for the runtime mask m of NB register bits, by tests of its bits
// (no multi-way branch; see apply_pair_t)
template <int NB, int ACC, class F>
__device__ __forceinline__ void with_slot_mask(uint32_t m, F &f) {
    if constexpr (NB == 0) {
        f(IC<ACC>());
    } else {
        if (m & (1u << (NB - 1)))
            with_slot_mask<NB - 1, ACC | (1 << (NB - 1))>(m, f);
        else
            with_slot_mask<NB - 1, ACC>(m, f);
    }
}

template <int DK>
__device__ __forceinline__ void apply_diag_t(dv2 (&v)[TR], uint64_t base, uint32_t tid, const TGate &G) {
    const dv2 ph0 = mk2(G.ph0x, G.ph0y), ph1 = mk2(G.ph1x, G.ph1y);
    const uint32_t cm = G.cm, tp = G.tpos;
    const uint32_t rb = tp - (uint32_t)TB;
    const bool reg_t = rb < 8u;         // target on a register bit
    const bool out_t = tp == kOutside;  // target outside the tile: its bit is fixed per tile
    const bool thr_t = !reg_t && !out_t;
    const uint32_t clo = G.dlo, chi = G.dhi;
    if constexpr (DK == DK_Z || DK == DK_P) {
        // dlo / dhi hold target and controls; a target outside the tile is
        // among the record's outside controls, tested before the gate
        const uint32_t mlo = clo, mhi = chi;
        const bool sel = (tid & mlo) == mlo;
        if constexpr (DK == DK_Z) {
            const uint32_t m = sel ? 0x80000000u : 0u;
            auto body = [&](auto mh) {
                constexpr int MH = decltype(mh)::value;
#pragma unroll
                for (int r = 0; r < TR; ++r)
                    if ((r & MH) == MH) v[r] = flip_sign(v[r], m);
            };
            with_slot_mask<TGR, 0>(mhi, body);
        } else {
            const dv2 f = sel ? ph1 : mk2(1.0, 0.0);
            auto body = [&](auto mh) {
                constexpr int MH = decltype(mh)::value;
#pragma unroll
                for (int r = 0; r < TR; ++r)
                    if ((r & MH) == MH) v[r] = cmul(v[r], f);
            };
            with_slot_mask<TGR, 0>(mhi, body);
        }
    } else {
        const bool ok = (tid & clo) == clo;
        if (!reg_t) {
            // one phase per thread (target on a thread bit) or per tile (outside)
            const bool one = thr_t ? ((tid >> (tp & 31u)) & 1u) != 0 : (base & G.otb) != 0;
            const dv2 f = ok ? (one ? ph1 : ph0) : mk2(1.0, 0.0);
            auto body = [&](auto mh) {
                constexpr int MH = decltype(mh)::value;
#pragma unroll
                for (int r = 0; r < TR; ++r)
                    if ((r & MH) == MH) v[r] = cmul(v[r], f);
            };
            with_slot_mask<TGR, 0>(chi, body);
        } else if (cm == 0) {
            // target on register bit RB: every slot, each with a fixed phase
            auto body = [&](auto tb) {
                constexpr int TM = decltype(tb)::value;
                if constexpr (TM != 0 && (TM & (TM - 1)) == 0) {
#pragma unroll
                    for (int r = 0; r < TR; ++r) v[r] = cmul(v[r], (r & TM) ? ph1 : ph0);
                }
            };
            with_slot_mask<TGR, 0>(1u << rb, body);
        } else {
            // controlled, target on a register bit: the phase of slot r is
            // picked by uniform selects; slots and threads left out multiply by 1
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                const dv2 ph = ((uint32_t)r & chi) == chi ? (((r >> rb) & 1u) ? ph1 : ph0) : mk2(1.0, 0.0);
                v[r] = cmul(v[r], ok ? ph : mk2(1.0, 0.0));
            }
        }
    }
}

__device__ __forceinline__ void apply_diag(dv2 (&v)[TR], uint64_t base, uint32_t tid, const TGate &G) {
    if (G.kind == KZ)
        apply_diag_t<DK_Z>(v, base, tid, G);
    else if (G.kind == KP)
        apply_diag_t<DK_P>(v, base, tid, G);
    else
        apply_diag_t<DK_RZ>(v, base, tid, G);
}

// The tile number spread over the bits outside Q (a 0 inserted at each position of Q).
__device__ __forceinline__ uint64_t tile_base(uint64_t k, const TileSpec &ts) {
#pragma unroll
    for (int j = 0; j < TL; ++j) {
        const uint64_t lo = k & ts.qlo[j];
        k = ((k ^ lo) << 1) | lo;
    }
    return k;
}

__device__ __forceinline__ uint64_t slot_offset(uint64_t dt, int r, const TileSpec &ts) {
    uint64_t o = dt;
#pragma unroll
    for (int b = 0; b < TL - TB; ++b)
        if ((r >> b) & 1) o |= ts.qbit[TB + b];
    return o;
}

// ---- compile-time dispatch (pass-specialised kernels) ----
// The same paths as apply_pair / apply_diag, chosen from the record's
// structural fields given as template arguments, so a kernel compiled for a
// pass instantiates only the path each gate takes.  The coefficients still
// come from the record.

// Exchange of lane bit LB (4 or 5) with register bit K.  v_permlane16_swap
// (LB 4) swaps the odd 16-lane rows of its first register with the even rows
// of its second, v_permlane32_swap (LB 5) the upper 32 lanes of the first with
// the lower 32 of the second.  Applied to the register pair (r, r | 1 << K)
// of every slot r with bit K clear, the amplitude at lane bit LB = x and
// register bit K = y moves to lane bit LB = y and register bit K = x: lane bit
// LB and register bit K trade places, every other bit stays.  Applying it
// twice restores the layout.  Four swaps per register pair, no LDS traffic
// (the lane exchange through ds_bpermute costs four LDS permutes per
// amplitude and gate).
template <int LB>
__device__ __forceinline__ void swap_rows(double &p, double &q) {
    const unsigned long long pb = (unsigned long long)__double_as_longlong(p);
    const unsigned long long qb = (unsigned long long)__double_as_longlong(q);
    unsigned plo = (unsigned)pb, phi = (unsigned)(pb >> 32), qlo = (unsigned)qb, qhi = (unsigned)(qb >> 32);
    if constexpr (LB == 4) {
        const auto lo = __builtin_amdgcn_permlane16_swap(plo, qlo, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(phi, qhi, false, false);
        plo = lo[0]; qlo = lo[1]; phi = hi[0]; qhi = hi[1];
    } else {
        const auto lo = __builtin_amdgcn_permlane32_swap(plo, qlo, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(phi, qhi, false, false);
        plo = lo[0]; qlo = lo[1]; phi = hi[0]; qhi = hi[1];
    }
    p = __longlong_as_double((long long)(((unsigned long long)phi << 32) | plo));
    q = __longlong_as_double((long long)(((unsigned long long)qhi << 32) | qlo));
}

// HI: register-slot controls of the gate (pairs whose controls fail are
// neither changed by it nor moved)
template <int LB, int K, uint32_t HI = 0>
__device__ __forceinline__ void swap_lane_reg(dv2 (&v)[TR]) {
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        if ((r >> K) & 1) continue;
        if (((uint32_t)r & HI) != HI) continue;
        const int r1 = r | (1 << K);
        double ax = v[r].x, ay = v[r].y, bx = v[r1].x, by = v[r1].y;
        swap_rows<LB>(ax, bx);
        swap_rows<LB>(ay, by);
        v[r] = mk2(ax, ay);
        v[r1] = mk2(bx, by);
    }
}

// slot index j with a 0 bit inserted at bit K
__device__ __forceinline__ constexpr int insert_zero_bit(int j, int K) {
    return ((j >> K) << (K + 1)) | (j & ((1 << K) - 1));
}

// Exchange of wave bit W with register bit K through LDS: the amplitude at
// wave bit W = x and register bit K = y moves to wave bit y and register bit
// x; every other bit stays.  Of each register pair (r, r | 1 << K) a thread
// keeps the slot whose bit K equals its wave bit and trades the other with
// its partner across W, so half of a wave-bit gate's LDS bytes move.
template <int W, int K>
__device__ __forceinline__ void swap_wave_reg(dv2 (&v)[TR], dv2 *lds, uint32_t tid) {
    const bool hi = ((tid >> W) & 1u) != 0;
    const uint32_t pt = tid ^ (1u << W);
    constexpr int NP = TR / 2;
    static_assert(NP % kChunk == 0, "register pairs per LDS round");
#pragma unroll
    for (int q = 0; q < NP; q += kChunk) {
        __syncthreads();  // earlier exchange reads are done
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            const int r0 = insert_zero_bit(q + c, K), r1 = r0 | (1 << K);
            // values, not array elements: a select between two elements of
            // the register array becomes a select between their addresses
            // and puts the array in scratch memory
            dv2 x0 = v[r0], x1 = v[r1];
            asm volatile("" : "+v"(x0), "+v"(x1));
            lds[tid + (c << TB)] = hi ? x0 : x1;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            const int r0 = insert_zero_bit(q + c, K), r1 = r0 | (1 << K);
            const dv2 in = lds[pt + (c << TB)];
            v[r0] = hi ? in : v[r0];
            v[r1] = hi ? v[r1] : in;
        }
    }
}

// A gate register bit that is not among the in-tile controls CM (-1: none)
constexpr int free_gate_reg_bit(uint32_t cm) {
    for (int k = TGR - 1; k >= 0; --k)
        if (!((cm >> (TB + k)) & 1u)) return k;
    return -1;
}

// Target on lane bit 4 or 5 (the ds_bpermute lanes): the target trades places
// with a free register bit K, the gate runs as a register-bit gate (both
// amplitudes of each pair in one thread, the half known per slot), and the
// layout is restored.  Controls keep their positions (K is not one of them,
// the target never is), so the control tests are unchanged; the outputs are
// the register-bit form's values, equal to the lane form's (the lane form's
// sign flips give the same operands).
template <uint32_t TP, int PK, bool CTRL, uint32_t CM>
__device__ __forceinline__ void pair_at(dv2 (&v)[TR], dv2 *lds, uint32_t tid, const TGate &G) {
    if constexpr (TP >= (uint32_t)TB) {
        reg_pair<(int)(TP - TB), PK, CTRL, CM>(v, tid, G);
    } else if constexpr ((TP == 4 || TP == 5) && free_gate_reg_bit(CM) >= 0) {
        constexpr int K = free_gate_reg_bit(CM);
        swap_lane_reg<(int)TP, K, (CM >> TB)>(v);
        reg_pair<K, PK, CTRL, CM>(v, tid, G);
        swap_lane_reg<(int)TP, K, (CM >> TB)>(v);
    } else if constexpr (TP < 6) {
        constexpr int LX = (TP & 4u) ? LX_SHFL : (TP & 2u) ? ((TP & 1u) ? LX_R8 : LX_R4) : ((TP & 1u) ? LX_Q2 : LX_Q1);
        lane_pair<PK, CTRL, LX, CM>(v, tid, TP, G);
    } else {
        wave_pair<PK, CTRL, CM>(v, lds, tid, TP, G);
    }
}

// apply_diag_t with its runtime tests resolved: DK, target position TP, cm,
// and the record's thread / register-slot masks DLO / DHI
// OUT: the record's outside bits (G.ocm) are tested here rather than by a
// branch around the gate (Z: the test joins the per-thread sign mask, so a
// tile that is not selected flips nothing; see z_outside_at).
template <int DK, uint32_t TP, uint32_t CM, uint32_t DLO, uint32_t DHI, bool OUT = false>
__device__ __forceinline__ void diag_at(dv2 (&v)[TR], uint64_t base, uint32_t tid, const TGate &G) {
    const dv2 ph0 = mk2(G.ph0x, G.ph0y), ph1 = mk2(G.ph1x, G.ph1y);
    constexpr bool reg_t = TP >= (uint32_t)TB && TP != kOutside;
    constexpr bool out_t = TP == kOutside;
    constexpr bool thr_t = !reg_t && !out_t;
    if constexpr (DK == DK_Z || DK == DK_P) {
        const bool sel = (tid & DLO) == DLO && (!OUT || (base & G.ocm) == G.ocm);
        if constexpr (DK == DK_Z) {
            const uint32_t m = sel ? 0x80000000u : 0u;
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                slot_group(r);
                if (((uint32_t)r & DHI) == DHI) v[r] = flip_sign(v[r], m);
            }
        } else {
            const dv2 f = sel ? ph1 : mk2(1.0, 0.0);
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                slot_group(r);
                if (((uint32_t)r & DHI) == DHI) v[r] = cmul(v[r], f);
            }
        }
    } else {
        const bool ok = (tid & DLO) == DLO;
        if constexpr (!reg_t) {
            bool one;
            if constexpr (thr_t)
                one = ((tid >> TP) & 1u) != 0;
            else
                one = (base & G.otb) != 0;
            const dv2 f = ok ? (one ? ph1 : ph0) : mk2(1.0, 0.0);
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                slot_group(r);
                if (((uint32_t)r & DHI) == DHI) v[r] = cmul(v[r], f);
            }
        } else if constexpr (CM == 0) {
            constexpr uint32_t TM = 1u << (TP - TB);
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                slot_group(r);
                v[r] = cmul(v[r], ((uint32_t)r & TM) ? ph1 : ph0);
            }
        } else {
            constexpr uint32_t rb = TP - TB;
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                slot_group(r);
                const dv2 ph = ((uint32_t)r & DHI) == DHI ? ((((uint32_t)r >> rb) & 1u) ? ph1 : ph0) : mk2(1.0, 0.0);
                v[r] = cmul(v[r], ok ? ph : mk2(1.0, 0.0));
            }
        }
    }
}

// A diagonal gate whose target and controls all lie outside the tile
// multiplies the whole tile by one phase (fixed by the tile's outside bits),
// and a tile-wide scalar commutes with every gate of the pass.  Compiled
// passes therefore collect these gates into one per-tile phase ph (products
// of the records' phases, computed without branches) and apply it once,
// before the store, where it is not exactly 1.  Z and P: the phase of the
// |1> half where the target and outside controls (G.ocm, which holds the
// target) are all 1; RZ: the phase of the target's half where the outside
// controls hold.
template <uint32_t KIND>
__device__ __forceinline__ void tile_phase_acc(dv2 &ph, uint64_t base, const TGate &G) {
    dv2 f;
    if constexpr (KIND == KRZ) {
        const bool ok = (base & G.ocm) == G.ocm;
        const bool one = (base & G.otb) != 0;
        f = ok ? (one ? mk2(G.ph1x, G.ph1y) : mk2(G.ph0x, G.ph0y)) : mk2(1.0, 0.0);
    } else {
        f = (base & G.ocm) == G.ocm ? mk2(G.ph1x, G.ph1y) : mk2(1.0, 0.0);
    }
    ph = cmul(ph, f);
}

__device__ __forceinline__ void tile_phase_apply(dv2 (&v)[TR], dv2 ph) {
    if (ph.x != 1.0 || ph.y != 0.0) {
#pragma unroll
        for (int r = 0; r < TR; ++r) {
            slot_group(r);
            v[r] = cmul(v[r], ph);
        }
    }
}

// A diagonal gate whose target and in-tile controls all lie on thread bits
// or outside the tile (no register bit) multiplies every register slot of a
// thread by the same factor.  Compiled passes fold such gates into the
// thread's pending phase tp (one complex product per gate instead of one
// per slot) and apply tp to the slots only where a later gate or exchange
// moves amplitudes between threads whose pending phases differ, and before
// the store.  The factor is diag_at's, outside tests included: Z and P, the
// |1> phase where the thread bits DLO (target and controls) and the outside
// bits G.ocm (outside controls, and the target when it is outside) are all
// 1; RZ, the target's half's phase where the controls hold.
__device__ __forceinline__ void thread_phase_apply(dv2 (&v)[TR], dv2 tp) {
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        slot_group(r);
        v[r] = cmul(v[r], tp);
    }
}

// An uncontrolled RZ (or fused diagonal gate of its form) on register bit
// TP - TB: the |0> phase multiplies every amplitude of the tile, so it joins
// the pending phase tp, and only the |1> slots take the ratio ph1 conj(ph0)
// (the phases have unit modulus): 8 + 2 complex products instead of 16.
template <uint32_t TP>
__device__ __forceinline__ void rz_reg_fold(dv2 (&v)[TR], dv2 &tp, const TGate &G) {
    static_assert(TP >= (uint32_t)TB && TP != kOutside, "register target");
    constexpr uint32_t TM = 1u << (TP - TB);
    const dv2 ph0 = mk2(G.ph0x, G.ph0y);
    const dv2 ratio = cmul(mk2(G.ph1x, G.ph1y), mk2(ph0.x, -ph0.y));
    tp = cmul(tp, ph0);
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        slot_group(r);
        if ((uint32_t)r & TM) v[r] = cmul(v[r], ratio);
    }
}

template <uint32_t KIND, uint32_t TP, uint32_t DLO, bool OUT>
__device__ __forceinline__ void thread_phase_acc(dv2 &tp, uint64_t base, uint32_t tid, const TGate &G) {
    const bool ok = (tid & DLO) == DLO && (!OUT || (base & G.ocm) == G.ocm);
    if constexpr (KIND == KZ) {
        tp = flip_sign(tp, ok ? 0x80000000u : 0u);
    } else if constexpr (KIND == KP) {
        tp = cmul(tp, ok ? mk2(G.ph1x, G.ph1y) : mk2(1.0, 0.0));
    } else {
        static_assert(KIND == KRZ && (TP < (uint32_t)TB || TP == kOutside), "thread-level RZ");
        bool one;
        if constexpr (TP == kOutside)
            one = (base & G.otb) != 0;
        else
            one = ((tid >> TP) & 1u) != 0;
        tp = cmul(tp, ok ? (one ? mk2(G.ph1x, G.ph1y) : mk2(G.ph0x, G.ph0y)) : mk2(1.0, 0.0));
    }
}

// A Z with outside bits (controls, or its target) and the rest in the tile:
// without a branch around it.  A per-tile branch around a gate makes the
// register tile a join of two versions, and a pass with many of them kept
// about 15 more registers live (a long pass then spilled at 4 waves per
// SIMD); a Z costs a sign flip either way.
template <uint32_t TP, uint32_t CM, uint32_t DLO, uint32_t DHI>
__device__ __forceinline__ void z_outside_at(dv2 (&v)[TR], uint64_t base, uint32_t tid, const TGate &G) {
    diag_at<DK_Z, TP, CM, DLO, DHI, true>(v, base, tid, G);
}

// One gate whose kind, target position, in-tile controls and diagonal masks
// are compile-time constants (the outside-control test stays with the caller).
template <uint32_t KIND, uint32_t TP, uint32_t CM, uint32_t DLO, uint32_t DHI>
__device__ __forceinline__ void gate_at(dv2 (&v)[TR], dv2 *lds, uint64_t base, uint32_t tid, const TGate &G) {
    if constexpr (KIND == KZ) {
        diag_at<DK_Z, TP, CM, DLO, DHI>(v, base, tid, G);
    } else if constexpr (KIND == KP) {
        diag_at<DK_P, TP, CM, DLO, DHI>(v, base, tid, G);
    } else if constexpr (KIND == KRZ) {
        diag_at<DK_RZ, TP, CM, DLO, DHI>(v, base, tid, G);
    } else if constexpr (KIND == KX) {
        pair_at<TP, PK_X, CM != 0, CM>(v, lds, tid, G);
    } else if constexpr (CM != 0) {
        pair_at<TP, PK_GEN, true, CM>(v, lds, tid, G);
    } else if constexpr (KIND == KY) {
        pair_at<TP, PK_Y, false, CM>(v, lds, tid, G);
    } else if constexpr (KIND == KH) {
        pair_at<TP, PK_H, false, CM>(v, lds, tid, G);
    } else if constexpr (KIND == KH_UNSCALED) {
        pair_at<TP, PK_HU, false, CM>(v, lds, tid, G);
    } else if constexpr (KIND == KRX) {
        pair_at<TP, PK_RX, false, CM>(v, lds, tid, G);
    } else {
        pair_at<TP, PK_TWO, false, CM>(v, lds, tid, G);
    }
}

// One gate of a pass on the tile of this workgroup (base: the tile's bits
// outside the tile qubits).
__device__ __forceinline__ void apply_gate(dv2 (&v)[TR], dv2 *lds, uint64_t base, uint32_t tid, const TGate &G) {
    if ((base & G.ocm) == G.ocm) {  // a control outside the tile is 0 here
        if (G.kind == KZ || G.kind == KP || G.kind == KRZ)
            apply_diag(v, base, tid, G);
        else
            apply_pair(v, lds, tid, G);
    }
}

// The tile a workgroup takes.  Workgroups are dispatched in block order,
// round-robin over the 8 XCDs, so the low bits of the block index vary
// fastest over the workgroups in flight.  With ntiles a power of two, the
// block index is rotated right by `order` bits within the tile index's
// log2(ntiles) bits: those low bits become its top bits, and the workgroups
// in flight spread over 2^order regions of the state (order 3: each XCD walks
// its own eighth in tile order).  Which order streams best depends on the
// pass's tile qubits, so it is measured per pass (the order tuner of
// tile_jit.hip).  Other tile counts: order 3 with ntiles a multiple of 8
// gives each XCD its own eighth, any other order is the block index.  Every
// order is a permutation of the tiles, and each tile's arithmetic does not
// depend on which workgroup runs it, so the results are the same bits.
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t ntiles, uint32_t order) {
    if ((ntiles & (ntiles - 1u)) == 0u) {
        const uint32_t L = 31u - (uint32_t)__builtin_clz(ntiles | 1u);
        if (order == 0u || order >= L) return b;
        return ((b >> order) | (b << (L - order))) & (ntiles - 1u);
    }
    if (order == 3u && (ntiles & 7u) == 0u) return (b & 7u) * (ntiles >> 3) + (b >> 3);
    return b;
}

// The thread's amplitude offsets in its tile: dt is the index of its slot 0.
__device__ __forceinline__ uint64_t thread_offset(uint64_t base, uint32_t tid, const TileSpec &ts) {
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    return base | tmask;
}

__device__ __forceinline__ void tile_load(dv2 (&v)[TR], const dv2 *__restrict__ a, uint64_t dt, const TileSpec &ts) {
#pragma unroll
    for (int r = 0; r < TR; ++r) v[r] = __builtin_nontemporal_load(a + slot_offset(dt, r, ts));
}

// tile_store through a permuted layout: position p holds the qubit loaded at
// position AT[p] (a compiled pass that moved qubits between positions).
template <int... AT>
__device__ __forceinline__ void tile_store_at(const dv2 (&v)[TR], dv2 *__restrict__ a, uint64_t base, uint32_t tid,
                                              const TileSpec &ts) {
    static_assert(sizeof...(AT) == TL, "one load position per tile position");
    constexpr int at[TL] = {AT...};
    uint64_t dt = base;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) dt |= ts.qbit[at[b]];
    asm volatile("" : "+v"(dt));
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        uint64_t o = dt;
#pragma unroll
        for (int k = 0; k < TL - TB; ++k)
            if ((r >> k) & 1) o |= ts.qbit[at[TB + k]];
        __builtin_nontemporal_store(v[r], a + o);
    }
}

// |0...0> of every set in registers: amplitude 1 where the index within its
// set (index & set_mask) is 0
__device__ __forceinline__ void tile_zero_state(dv2 (&v)[TR], uint64_t dt, const TileSpec &ts, uint64_t set_mask) {
#pragma unroll
    for (int r = 0; r < TR; ++r) v[r] = mk2((slot_offset(dt, r, ts) & set_mask) == 0 ? 1.0 : 0.0, 0.0);
}

__device__ __forceinline__ void tile_store(const dv2 (&v)[TR], dv2 *__restrict__ a, uint64_t dt, const TileSpec &ts) {
    // recompute the store addresses instead of keeping eight of them live
    // through the gate code
    asm volatile("" : "+v"(dt));
#pragma unroll
    for (int r = 0; r < TR; ++r) __builtin_nontemporal_store(v[r], a + slot_offset(dt, r, ts));
}

// ---- Pauli-string sweeps compiled per string group (exp_value) ----
// A string P = i^ny X^x Z^z gives <psi|P|psi> = sum_i conj(a_{i^x}) a_i
// i^ny (-1)^popcount(i & z).  Per tile, w_i = conj(a_{i^x}) a_i is formed
// once for every string with the same x (the partner comes from a register,
// a lane move or the tile in LDS, chosen at compile time), and each string
// adds the real part of i^ny w_i with the sign of popcount(i & z): the
// register-slot part of that parity and i^ny are compile-time constants, the
// part from the tile's and thread's bits one test per string and tile.

// a string record of a sweep (tiled.hip's string table): z as global bits,
// ny, x as tile positions and the string's coefficient
struct SweepString {
    uint64_t z;
    uint32_t ny, xt;
    double coeff;
};
static_assert(sizeof(SweepString) == 24, "sweep string record");

// lane xor by a compile-time mask: one DPP move for 1, 2 and 8, two for 4,
// a cross-lane shuffle otherwise
template <uint32_t XL>
__device__ __forceinline__ double lane_xor_any(double x) {
    if constexpr (XL == 1u)
        return lane_xor_d<LX_Q1>(x, 1);
    else if constexpr (XL == 2u)
        return lane_xor_d<LX_Q2>(x, 2);
    else if constexpr (XL == 8u)
        return lane_xor_d<LX_R8>(x, 8);
    else if constexpr (XL == 4u)
        return lane_xor_d<LX_R4>(x, 4);
    else
        return __shfl_xor(x, (int)XL, 64);
}

// w[r] = conj(a_{i^x}) a_i for the thread's slots (XT: x as tile positions;
// WAVE: the tile is in LDS, lds[tid | r << TB])
template <uint32_t XT>
__device__ __forceinline__ void sweep_w(dv2 (&w)[TR], const dv2 (&v)[TR], const dv2 *lds, uint32_t tid) {
    constexpr uint32_t XR = XT >> TB, XTH = XT & (uint32_t)(TT - 1), XL = XT & 63u;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const dv2 m = v[r];
        if constexpr (XT == 0) {
            w[r] = mk2(m.x * m.x + m.y * m.y, 0.0);
        } else {
            dv2 o;
            if constexpr (XTH != XL)
                o = lds[(tid ^ XTH) | ((uint32_t)(r ^ (int)XR) << TB)];
            else if constexpr (XL != 0)
                o = mk2(lane_xor_any<XL>(v[r ^ (int)XR].x), lane_xor_any<XL>(v[r ^ (int)XR].y));
            else
                o = v[r ^ (int)XR];
            w[r] = mk2(o.x * m.x + o.y * m.y, o.x * m.y - o.y * m.x);
        }
    }
}

// sum over the thread's slots of (-1)^popcount(r & ZR) Re(i^Q w[r])
// (Re(i^q w) = w.x, -w.y, -w.x, w.y for q = 0..3)
template <uint32_t ZR, uint32_t Q>
__device__ __forceinline__ double sweep_sum(const dv2 (&w)[TR]) {
    double sp = 0.0, sn = 0.0;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const double t = (Q & 1u) ? w[r].y : w[r].x;
        const bool neg = ((__builtin_popcount((uint32_t)r & ZR) & 1) != 0) != (((Q + 1u) & 2u) != 0);
        if (neg)
            sn += t;
        else
            sp += t;
    }
    return sp - sn;
}

// the workgroup's sum of v (every thread returns it)
__device__ __forceinline__ double sweep_block_sum(double v, double *ws) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < TT / 64; ++w) t += ws[w];
    __syncthreads();
    return t;
}

}  // namespace
}  // namespace tile
}  // namespace kbwh
