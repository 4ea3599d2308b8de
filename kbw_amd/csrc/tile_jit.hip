// Pass-specialised tile kernels (KBWH_RUN_COMPILED), compiled at run time.
//
// The generic pass kernel (k_tile_fma) reads each gate record and picks the
// gate's code path by uniform tests on its kind, tile position and masks.
// Those tests cost about 22 scalar instructions per gate, and the register
// allocator copies the register tile where the paths join (about 16 vector
// moves per gate), so a pass of 40+ gates is bound by instruction issue.  A
// specialised kernel is the same load / gates / store, with each gate's
// structural fields (kind, tile position, controls inside and outside the
// tile, diagonal masks) given as template arguments of gate_at (tile_dev.h):
// only the path each gate takes is instantiated (which also keeps the compile
// short: folding the runtime dispatch away instead took 7x longer) and the
// gates run as one straight line.  The coefficients are still read from the
// gate records, so one kernel serves every pass of the same structure
// whatever its angles.
//
// Kernels are compiled with hiprtc (loaded on first use; a library without it
// runs the generic kernel) from the device tile header (tile_dev.h, embedded
// in the library at build time), cached for the process by the pass's
// structure, and loaded once per device.  The passes a run needs are compiled
// in parallel host threads.
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hiprtc.h>

#define KBWH_TILE_CONTRACT 1
#include "state.h"
#include "tile_core.h"

namespace kbwh {

extern const char kTileDevSource[];  // tile_dev.h, embedded by the build
extern const char kAdjDevSource[];   // adj_dev.h, embedded by the build

namespace {

using namespace tile;

// the hiprtc entry points, resolved from libhiprtc at first use
struct Rtc {
    bool ok = false;
    std::string why;
    decltype(&hiprtcCreateProgram) create = nullptr;
    decltype(&hiprtcCompileProgram) compile = nullptr;
    decltype(&hiprtcDestroyProgram) destroy = nullptr;
    decltype(&hiprtcGetCodeSize) code_size = nullptr;
    decltype(&hiprtcGetCode) code = nullptr;
    decltype(&hiprtcGetProgramLogSize) log_size = nullptr;
    decltype(&hiprtcGetProgramLog) log = nullptr;
};

const Rtc &rtc() {
    static Rtc r;
    static std::once_flag once;
    std::call_once(once, [] {
        void *h = dlopen("libhiprtc.so.7", RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("libhiprtc.so", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            r.why = "libhiprtc not found";
            return;
        }
        r.create = (decltype(r.create))dlsym(h, "hiprtcCreateProgram");
        r.compile = (decltype(r.compile))dlsym(h, "hiprtcCompileProgram");
        r.destroy = (decltype(r.destroy))dlsym(h, "hiprtcDestroyProgram");
        r.code_size = (decltype(r.code_size))dlsym(h, "hiprtcGetCodeSize");
        r.code = (decltype(r.code))dlsym(h, "hiprtcGetCode");
        r.log_size = (decltype(r.log_size))dlsym(h, "hiprtcGetProgramLogSize");
        r.log = (decltype(r.log))dlsym(h, "hiprtcGetProgramLog");
        r.ok = r.create && r.compile && r.destroy && r.code_size && r.code && r.log_size && r.log;
        if (!r.ok) r.why = "libhiprtc lacks an entry point";
    });
    return r;
}

// The structure of a pass: the fields of each record that select code, and
// the pass flags.
std::string pass_key(const TGate *g, uint32_t count, uint32_t flags) {
    std::string k;
    k.reserve(count * 56 + 4);
    k.append((const char *)&flags, sizeof flags);
    for (uint32_t i = 0; i < count; ++i) {
        const uint64_t f[7] = {g[i].kind, g[i].tpos, g[i].cm, g[i].ocm, g[i].otb, g[i].dlo, g[i].dhi};
        k.append((const char *)f, sizeof f);
    }
    return k;
}

uint64_t fnv1a(const std::string &s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) h = (h ^ c) * 1099511628211ull;
    return h;
}

// The kernel of one pass.  Arguments: the state, the pass's tile qubits, its
// gate records, the tile count, set_shift (< 32: the state holds parameter
// sets of 2^(set_shift + TL) amplitudes one after another, and tile t takes
// the records gp + (t >> set_shift) * count) and, for an initialising pass
// (JIT_INIT), the amplitude mask of one set: the tile starts as |0...0> of
// its set in registers instead of being loaded.
// A wide pass (JIT_WIDE) runs the wide tile (kWideTileQubits qubits over
// 2^kWideTileTBits threads, 16 amplitudes per thread), built for `waves`
// waves per SIMD (the register budget: 4 gives 128 VGPRs, 3 gives 168).
std::string pass_source(const TGate *g, uint32_t count, uint32_t flags, const std::string &name, int waves = 3) {
    std::string s;
    s.reserve(1024 + count * 200);
    if (flags & kJitWide)
        s += "#define KBWH_TILE_BITS " + std::to_string(kWideTileQubits) + "\n#define KBWH_TILE_TBITS " +
             std::to_string(kWideTileTBits) + "\n#define KBWH_TILE_WAVES " + std::to_string(waves) +
             "\n";
    s += "#define KBWH_TILE_CONTRACT 1\n#include \"tile_dev.h\"\nusing namespace kbwh::tile;\n";
    s += "extern \"C\" __global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, "
         "KBWH_TILE_WAVES))) void ";
    s += name;
    s += "(dv2 *__restrict__ a, TileSpec ts, const TGate *__restrict__ gp, uint32_t ntiles, uint32_t set_shift,\n"
         "    uint64_t set_mask, uint32_t order";
    s += (flags & kJitPermStore) ? ", TileSpec tss) {\n" : ") {\n";
    s += 
         "  __shared__ dv2 lds[TT * kChunk];\n"
         "  const uint32_t tid = threadIdx.x;\n"
         "  const uint32_t tile = tile_of_block(blockIdx.x, ntiles, order);\n"
         "  if (tile >= ntiles) return;\n";
    s += "  if (set_shift < 32) gp += (size_t)(tile >> set_shift) * " + std::to_string(count) + "u;\n";
    s += "  const uint64_t base = tile_base(tile, ts);\n"
         "  const uint64_t dt = thread_offset(base, tid, ts);\n"
         "  dv2 v[TR];\n";
    s += (flags & kJitInit) ? "  tile_zero_state(v, dt, ts, set_mask);\n" : "  tile_load(v, a, dt, ts);\n";
    char line[320];
    // diagonal gates wholly outside the tile: one per-tile phase (tile_phase_acc)
    auto tile_wide = [](const TGate &G) {
        return (G.kind == KZ || G.kind == KP || G.kind == KRZ) && G.tpos == kOutside && G.cm == 0;
    };
    // Layout of a wide pass, followed through the pass (the records hold load
    // positions): a pair gate whose target sits on a wave bit or on lane bit
    // 4 or 5 first trades that position with a register bit (swap_wave_reg
    // through LDS, half of a wave-bit gate's LDS bytes; swap_lane_reg by
    // permlane swaps) and runs as a register-bit gate, and the qubit stays
    // there for its later gates.  The register bit given up is the one whose
    // qubit is next a target furthest ahead.  cur[l]: the position now
    // holding the qubit loaded at position l; at[p]: the load position of
    // the qubit now at p.  The tile is stored through the final layout.
    const bool relayout = (flags & kJitWide) != 0;
    const int kTb = (flags & kJitWide) ? (int)kWideTileTBits : TB;
    const int tl = (flags & kJitWide) ? (int)kWideTileQubits : TL;
    int cur[16], at[16];
    for (int q = 0; q < 16; ++q) cur[q] = at[q] = q;
    bool moved = false;
    auto is_diag = [](const TGate &G) { return G.kind == KZ || G.kind == KP || G.kind == KRZ; };
    auto remap = [&](uint32_t m) {
        uint32_t o = 0;
        for (int q = 0; q < tl; ++q)
            if ((m >> q) & 1u) o |= 1u << cur[q];
        return o;
    };
    // positions of gate G in the current layout: target, in-tile controls, diagonal masks
    auto positions = [&](const TGate &G, uint32_t &tp, uint32_t &cm, uint32_t &dlo, uint32_t &dhi) {
        tp = G.tpos == kOutside ? kOutside : (uint32_t)cur[G.tpos];
        cm = remap(G.cm);
        const uint32_t d = remap(G.dlo | (G.dhi << kTb));
        dlo = d & ((1u << kTb) - 1u);
        dhi = d >> kTb;
    };
    // The pending phase tp of each thread (thread_phase_acc): diagonal gates
    // wholly outside the tile (one factor per tile) and diagonal gates on
    // thread bits and outside bits only are folded into it; tp_dep holds the
    // thread positions (current layout) its factor depends on.  Before a pair
    // gate or a layout swap moves amplitudes across a thread bit in tp_dep,
    // tp is applied to the slots and restarts at 1; what is left is applied
    // before the store.  tp_thread: a thread-level gate was folded since the
    // last application (tp may differ between lanes); tp_any: anything was.
    // uncontrolled H: run unscaled (PK_HU), its 1/sqrt2 joining tp
    auto h_unscaled = [](const TGate &G) { return G.kind == KH && G.cm == 0 && G.ocm == 0; };
    bool any_diag = false;
    for (uint32_t i = 0; i < count; ++i) any_diag = any_diag || is_diag(g[i]) || h_unscaled(g[i]);
    if (any_diag) s += "  dv2 tp = mk2(1.0, 0.0);\n";
    uint32_t tp_dep = 0;
    bool tp_thread = false, tp_any = false;
    auto flush_phase = [&]() {
        s += "  thread_phase_apply(v, tp); tp = mk2(1.0, 0.0);\n  __builtin_amdgcn_sched_barrier(0);\n";
        tp_dep = 0;
        tp_thread = tp_any = false;
    };
    for (uint32_t i = 0; i < count; ++i) {
        const TGate &G = g[i];
        if (tile_wide(G)) {
            snprintf(line, sizeof line, "  tile_phase_acc<%uu>(tp, base, gp[%u]);\n", G.kind, i);
            s += line;
            tp_any = true;
            continue;
        }
        if (!is_diag(G) && G.tpos != kOutside) {
            // a pair gate (or its layout swap) on a thread bit the pending phase depends on
            const int p = cur[G.tpos];
            if (p < kTb && ((tp_dep >> p) & 1u)) flush_phase();
        }
        if (relayout && !is_diag(G) && G.tpos != kOutside) {
            const int p = cur[G.tpos];
            const bool wave = p >= 6 && p < kTb, lane45 = p == 4 || p == 5;
            if (wave || lane45) {
                // the register bit whose qubit is next a target furthest ahead
                int best = kTb;
                uint32_t best_next = 0;
                for (int r = kTb; r < tl; ++r) {
                    uint32_t nx = count;
                    for (uint32_t j = i + 1; j < count; ++j)
                        if (!is_diag(g[j]) && g[j].tpos == (uint32_t)at[r]) {
                            nx = j;
                            break;
                        }
                    if (r == kTb || nx > best_next) {
                        best = r;
                        best_next = nx;
                    }
                }
                if (wave)
                    snprintf(line, sizeof line, "  swap_wave_reg<%d, %d>(v, lds, tid);\n", p, best - kTb);
                else
                    snprintf(line, sizeof line, "  swap_lane_reg<%d, %d>(v);\n", p, best - kTb);
                s += line;
                s += "  __builtin_amdgcn_sched_barrier(0);\n";
                const int a = at[p], b = at[best];
                at[p] = b;
                at[best] = a;
                cur[a] = best;
                cur[b] = p;
                moved = true;
            }
        }
        uint32_t tp, cm, dlo, dhi;
        positions(G, tp, cm, dlo, dhi);
        if (is_diag(G) && dhi == 0 && (G.kind != KRZ || tp == kOutside || tp < (uint32_t)kTb)) {
            // one factor per thread: folded into tp
            snprintf(line, sizeof line,
                     "  { uint32_t tg = tid; asm volatile(\"\" : \"+v\"(tg)); "
                     "thread_phase_acc<%uu, %uu, %uu, %s>(tp, base, tg, gp[%u]); }\n",
                     G.kind, tp, dlo, G.ocm ? "true" : "false", i);
            s += line;
            tp_dep |= dlo;
            if (G.kind == KRZ && tp != kOutside) tp_dep |= 1u << tp;
            tp_thread = tp_any = true;
            continue;
        }
        if (G.kind == KRZ && tp != kOutside && tp >= (uint32_t)kTb && cm == 0 && dhi == 0 && G.ocm == 0) {
            // uncontrolled RZ on a register bit: its |0> phase joins tp (uniform over the tile)
            snprintf(line, sizeof line, "  rz_reg_fold<%uu>(v, tp, gp[%u]);\n  __builtin_amdgcn_sched_barrier(0);\n", tp,
                     i);
            s += line;
            tp_any = true;
            continue;
        }
        if (G.kind == KZ && G.ocm) {
            snprintf(line, sizeof line,
                     "  { uint32_t tg = tid; asm volatile(\"\" : \"+v\"(tg)); "
                     "z_outside_at<%uu, %uu, %uu, %uu>(v, base, tg, gp[%u]); }\n"
                     "  __builtin_amdgcn_sched_barrier(0);\n",
                     tp, cm, dlo, dhi, i);
            s += line;
            continue;
        }
        // the thread index as a new value for every gate: the per-lane masks
        // a gate derives from it are computed where the gate runs, not
        // hoisted to the top of the kernel for every gate at once
        const int k = snprintf(line, sizeof line,
                               "  { uint32_t tg = tid; asm volatile(\"\" : \"+v\"(tg)); const TGate &G = gp[%u]; ", i);
        int m = k;
        if (G.ocm)
            m += snprintf(line + m, sizeof line - m, "if ((base & %lluull) == %lluull) ", (unsigned long long)G.ocm,
                          (unsigned long long)G.ocm);
        if (h_unscaled(G)) {
            snprintf(line + m, sizeof line - m,
                     "gate_at<KH_UNSCALED, %uu, %uu, %uu, %uu>(v, lds, base, tg, G); tp = tp * G.d; }\n", tp, cm, dlo,
                     dhi);
            tp_any = true;
        } else {
            snprintf(line + m, sizeof line - m, "gate_at<%uu, %uu, %uu, %uu, %uu>(v, lds, base, tg, G); }\n", G.kind,
                     tp, cm, dlo, dhi);
        }
        s += line;
        // one gate at a time: scheduled across gates, the temporaries of
        // several gates are live at once
        s += "  __builtin_amdgcn_sched_barrier(0);\n";
    }
    if (tp_thread)
        s += "  thread_phase_apply(v, tp);\n";
    else if (tp_any)
        s += "  tile_phase_apply(v, tp);\n";
    if (flags & kJitPermStore) {
        // Block swaps folded into the pass: the tile goes through tss to
        // addresses that other waves of the workgroup loaded, so every load
        // must have returned before any store is issued
        s += "  __builtin_amdgcn_s_waitcnt(0);\n"
             "  __syncthreads();\n"
             "  tile_store(v, a, thread_offset(base, tid, tss), tss);\n}\n";
    } else if (moved) {
        // the store through the final layout: position p holds the qubit loaded at at[p]
        s += "  tile_store_at<";
        for (int q = 0; q < tl; ++q) s += std::to_string(at[q]) + (q + 1 < tl ? ", " : "");
        s += ">(v, a, base, tid, ts);\n}\n";
    } else {
        s += "  tile_store(v, a, dt, ts);\n}\n";
    }
    return s;
}

// The kernel of one exp_value sweep (11-qubit tiles; see sweep_w in
// tile_dev.h): every workgroup walks tiles blockIdx.x + j gridDim.x, reads
// each tile once, forms w per distinct x of the group's strings and adds each
// string's weighted sum; one partial sum per workgroup.
std::string sweep_source(const SweepKey *k, uint32_t count, const std::string &name) {
    bool wave = false;
    for (uint32_t i = 0; i < count; ++i) wave = wave || (k[i].xt & 0xC0u) != 0;
    std::string s;
    s.reserve(2048 + count * 160);
    // contracted (w's products fused into its adds): equal to the generic
    // sweep to rounding
    s += "#define KBWH_TILE_CONTRACT 1\n#include \"tile_dev.h\"\nusing namespace kbwh::tile;\n";
    s += "extern \"C\" __global__ __launch_bounds__(TT) void " + name +
         "(const dv2 *__restrict__ a, TileSpec ts, const SweepString *__restrict__ ss, uint32_t ntiles,\n"
         "    double *__restrict__ partial) {\n";
    s += wave ? "  __shared__ dv2 lds[TT * TR];\n" : "  const dv2 *lds = nullptr;\n";
    s += "  __shared__ double ws[TT / 64];\n"
         "  const uint32_t tid = threadIdx.x;\n"
         "  uint64_t tmask = 0;\n"
         "#pragma unroll\n"
         "  for (int b = 0; b < TB; ++b)\n"
         "    if ((tid >> b) & 1u) tmask |= ts.qbit[b];\n"
         "  double acc = 0.0;\n"
         "#pragma nounroll\n"
         "  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {\n"
         "    const uint64_t dt = tile_base(t, ts) | tmask;\n"
         "    dv2 v[TR];\n"
         "#pragma unroll\n"
         "    for (int r = 0; r < TR; ++r) v[r] = __builtin_nontemporal_load(a + slot_offset(dt, r, ts));\n";
    if (wave)
        s += "    __syncthreads();  // the previous tile's partner reads are done\n"
             "#pragma unroll\n"
             "    for (int r = 0; r < TR; ++r) lds[tid | (r << TB)] = v[r];\n"
             "    __syncthreads();\n";
    s += "    dv2 w[TR];\n";
    // strings with the same x share w; groups in order of first appearance
    std::vector<bool> done(count, false);
    char line[256];
    for (uint32_t i = 0; i < count; ++i) {
        if (done[i]) continue;
        snprintf(line, sizeof line, "    sweep_w<%uu>(w, v, lds, tid);\n", k[i].xt);
        s += line;
        for (uint32_t j = i; j < count; ++j) {
            if (done[j] || k[j].xt != k[i].xt) continue;
            done[j] = true;
            snprintf(line, sizeof line,
                     "    { const SweepString &S = ss[%u]; const double sk = sweep_sum<%uu, %uu>(w);\n"
                     "      acc += ((__popcll(dt & S.z) & 1) ? -S.coeff : S.coeff) * sk; }\n",
                     j, k[j].zr, k[j].q & 3u);
            s += line;
        }
    }
    s += "  }\n"
         "  const double total = sweep_block_sum(acc, ws);\n"
         "  if (tid == 0) partial[blockIdx.x] = total;\n"
         "}\n";
    return s;
}

std::string sweep_key(const SweepKey *k, uint32_t count) {
    std::string key;
    const uint32_t tag = kJitSweep;
    key.append((const char *)&tag, sizeof tag);
    for (uint32_t i = 0; i < count; ++i) key.append((const char *)&k[i], 3 * sizeof(uint32_t));
    return key;
}

// The reverse sweep's record (adj_dev.h's AGate; TGate does not depend on
// the tile width).
struct AGateRec {
    TGate inv;
    uint32_t slot, gkind, pad[2];
};
static_assert(sizeof(AGateRec) == 160, "AGate layout");
constexpr uint32_t kAdjNoSlot = 0xFFFFFFFFu;

std::string adj_pass_key(const AGateRec *g, uint32_t count) {
    std::string k;
    const uint32_t fl = kJitAdjoint;
    k.append((const char *)&fl, sizeof fl);
    for (uint32_t i = 0; i < count; ++i) {
        const TGate &G = g[i].inv;
        const uint64_t f[9] = {G.kind, G.tpos, G.cm, G.ocm, G.otb, G.dlo, G.dhi, g[i].slot, g[i].gkind};
        k.append((const char *)f, sizeof f);
    }
    return k;
}

// The reverse-sweep kernel of one pass: the generic k_adj_tile's loop over a
// workgroup's tiles (wps workgroups per set, tps tiles per set, the set's own
// records gs + set * count), with each gate's term and un-application
// dispatched at compile time.
std::string adj_pass_source(const AGateRec *g, uint32_t count, const std::string &name) {
    uint32_t nslots = 0;
    for (uint32_t i = 0; i < count; ++i)
        if (g[i].slot != kAdjNoSlot) nslots = std::max(nslots, g[i].slot + 1);
    std::string s;
    s.reserve(2048 + count * 260);
    // contracted like the forward passes (products fused into the following
    // adds): gradients are equal to the generic sweep's to rounding, and
    // uncontracted the reverse kernels took 1.6 times the FP64 instructions
    s += "#define KBWH_TILE_BITS 12\n#define KBWH_TILE_TBITS 8\n#define KBWH_TILE_GATE_RBITS 3\n"
         "#define KBWH_TILE_CONTRACT 1\n#include \"adj_dev.h\"\nusing namespace kbwh::tile;\n";
    s += "extern \"C\" __global__ __launch_bounds__(TT) void " + name +
         "(dv2 *__restrict__ a, TileSpec ts, const AGate *__restrict__ gs, uint32_t tps, uint32_t wps,\n"
         "    double *__restrict__ partial) {\n"
         "  __shared__ dv2 lds[TT * kChunk];\n";
    const uint32_t ns = std::max<uint32_t>(nslots, 1);
    s += "  __shared__ double acc[" + std::to_string(ns) + " * kAccPerSlot];\n"
         "  __shared__ double ws[TT / 64];\n"
         "  const uint32_t tid = threadIdx.x;\n"
         "  const uint32_t set = blockIdx.x / wps, wk = blockIdx.x % wps;\n"
         "  gs += (size_t)set * " + std::to_string(count) + "u;\n";
    s += "  adj_acc_zero(acc, " + std::to_string(ns) + "u, tid);\n"
         "  const uint64_t tmask = thread_offset(0, tid, ts);\n"
         "#pragma nounroll\n"
         "  for (uint32_t tl = wk; tl < tps; tl += wps) {\n"
         "    const uint32_t t = set * tps + tl;\n"
         "    const uint64_t base = tile_base(t, ts);\n"
         "    const uint64_t dt = base | tmask;\n"
         // the thread index as a new value in every iteration: the per-lane masks
         // and shuffle addresses the gates derive from it are then computed where
         // they are used, not hoisted out of the tile loop for every gate at once
         "    uint32_t tidl = tid;\n"
         "    asm volatile(\"\" : \"+v\"(tidl));\n"
         "    dv2 v[TR];\n"
         "#pragma unroll\n"
         "    for (int r = 0; r < TR; ++r) v[r] = a[slot_offset(dt, r, ts)];\n";
    char line[512];
    // The pending per-thread phase tp of the forward passes (thread_phase_acc):
    // a diagonal gate on thread bits and outside bits only is folded into tp,
    // which multiplies phi and lambda at the same index alike.  Every term is
    // a sum of conj(lambda_i) (G phi)_i, so a common factor of lambda_i and
    // phi_i drops out (a unit factor, to rounding), and so does the factor of
    // a partner phi across a bit tp does not depend on; tp is applied to the
    // slots before a pair gate (and its term) on a thread bit in tp_dep, and
    // before the store.  It restarts at 1 for every tile.
    const uint32_t adj_tb = 8;  // thread bits of the reverse sweep's tile
    auto diag_kind = [](uint32_t k) { return k == KZ || k == KP || k == KRZ; };
    uint32_t tp_dep = 0;
    bool tp_live = false;
    s += "    dv2 tp = mk2(1.0, 0.0);\n";
    for (uint32_t i = 0; i < count; ++i) {
        const TGate &G = g[i].inv;
        const bool diag = diag_kind(G.kind);
        if (!diag && G.tpos < adj_tb && ((tp_dep >> G.tpos) & 1u)) {
            s += "    thread_phase_apply(v, tp); tp = mk2(1.0, 0.0);\n    __builtin_amdgcn_sched_barrier(0);\n";
            tp_dep = 0;
        }
        const bool fold = diag && G.dhi == 0 && (G.kind != KRZ || G.tpos == kOutside || G.tpos < adj_tb);
        // uncontrolled RZ on a register bit: its |0> phase joins tp (rz_reg_fold)
        const bool rz_reg = G.kind == KRZ && G.tpos != kOutside && G.tpos >= adj_tb && G.cm == 0 && G.dhi == 0 &&
                            G.ocm == 0;
        int m = snprintf(line, sizeof line, "    { const AGate &A = gs[%u]; ", i);
        if (G.ocm)
            m += snprintf(line + m, sizeof line - m, "if ((base & %lluull) == %lluull) ", (unsigned long long)G.ocm,
                          (unsigned long long)G.ocm);
        m += snprintf(line + m, sizeof line - m, "{ ");
        if (g[i].slot != kAdjNoSlot)
            m += snprintf(line + m, sizeof line - m,
                          "adj_acc_add(acc, %uu, tid, adj_term_at<%uu, %uu, %uu>(v, lds, tidl, base, A.inv)); ", g[i].slot,
                          g[i].gkind, G.tpos, G.cm);
        if (rz_reg) {
            snprintf(line + m, sizeof line - m, "rz_reg_fold<%uu>(v, tp, A.inv); } }\n", G.tpos);
            tp_live = true;
        } else if (fold) {
            // the outside bits are tested by the branch around it
            snprintf(line + m, sizeof line - m, "thread_phase_acc<%uu, %uu, %uu, false>(tp, base, tidl, A.inv); } }\n",
                     G.kind, G.tpos, G.dlo);
            tp_dep |= G.dlo;
            if (G.kind == KRZ && G.tpos != kOutside) tp_dep |= 1u << G.tpos;
            tp_live = true;
        } else {
            snprintf(line + m, sizeof line - m, "gate_at<%uu, %uu, %uu, %uu, %uu>(v, lds, base, tidl, A.inv); } }\n",
                     G.kind, G.tpos, G.cm, G.dlo, G.dhi);
        }
        s += line;
        // one gate's term and update at a time: scheduled across gates, a
        // pass of 30-40 gates kept 230-300 registers live and spilled
        s += "    __builtin_amdgcn_sched_barrier(0);\n";
    }
    if (tp_live) s += "    thread_phase_apply(v, tp);\n";
    s += "#pragma unroll\n"
         "    for (int r = 0; r < TR; ++r) a[slot_offset(dt, r, ts)] = v[r];\n"
         "  }\n";
    s += "  for (uint32_t s = 0; s < " + std::to_string(nslots) + "u; ++s) {\n"
         "    const double x = adj_block_sum(adj_acc_get(acc, s, tid), ws);\n"
         "    if (tid == 0) partial[(size_t)s * gridDim.x + blockIdx.x] = x;\n"
         "  }\n}\n";
    return s;
}

// The kernel's fixed scratch (private segment) size from the code object's
// metadata note (msgpack: the key string, then an unsigned integer); -1 if
// the key is not found or its value is not an unsigned integer.
int64_t scratch_bytes(const std::vector<char> &code) {
    static const char key[] = ".private_segment_fixed_size";
    const size_t kl = sizeof key - 1;
    for (size_t i = 0; i + kl + 2 < code.size(); ++i) {
        if (memcmp(&code[i], key, kl) != 0) continue;
        const unsigned char *v = (const unsigned char *)&code[i + kl];
        if (v[0] <= 0x7f) return v[0];
        if (v[0] == 0xcc && i + kl + 2 <= code.size()) return v[1];
        if (v[0] == 0xcd && i + kl + 3 <= code.size()) return (int64_t)v[1] << 8 | v[2];
        if (v[0] == 0xce && i + kl + 5 <= code.size())
            return (int64_t)v[1] << 24 | (int64_t)v[2] << 16 | (int64_t)v[3] << 8 | v[4];
        return -1;
    }
    return -1;
}

const char *const kOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                "-Xclang", "-target-feature", "-Xclang",
                                "-packed-fp32-ops",
                                // forward passes only (see compile_source)
                                "-mllvm", "-structurizecfg-skip-uniform-regions=true"};
constexpr int kNumOptions = (int)(sizeof kOptions / sizeof kOptions[0]);

// compiles one pass kernel to a code object; false with a log on failure
// (also when the kernel uses scratch memory, unless allow_scratch).
// uniform_regions: the kernel is straight-line code whose only branches are
// uniform per-tile tests (a gate's controls outside the tile), which keep
// their branch instead of being structurized into guarded blocks: guarded,
// each such gate kept a second copy of the register tile live and long
// passes spilled at 3 waves per SIMD.  The option mistreats a loop whose
// lanes leave at different iterations followed by code that depends on how
// it was left; a forward pass kernel has no loop left after unrolling, the
// reverse sweep's kernels (a loop over tiles, block sums) are built without.
bool compile_source(const std::string &src, std::vector<char> &code, std::string &log, bool allow_scratch = false,
                    bool uniform_regions = false) {
    const Rtc &r = rtc();
    if (!r.ok) {
        log = r.why;
        return false;
    }
    hiprtcProgram p;
    const char *hdr[] = {kTileDevSource, kAdjDevSource};
    const char *inc[] = {"tile_dev.h", "adj_dev.h"};
    if (r.create(&p, src.c_str(), "kbwh_pass.hip", 2, hdr, inc) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    const hiprtcResult rc = r.compile(p, uniform_regions ? kNumOptions : kNumOptions - 2, kOptions);
    size_t n = 0;
    if (rc != HIPRTC_SUCCESS) {
        if (r.log_size(p, &n) == HIPRTC_SUCCESS && n > 1) {
            log.resize(n);
            r.log(p, &log[0]);
        } else {
            log = "hiprtc compile failed";
        }
        r.destroy(&p);
        return false;
    }
    bool ok = r.code_size(p, &n) == HIPRTC_SUCCESS && n > 0;
    if (ok) {
        code.resize(n);
        ok = r.code(p, code.data()) == HIPRTC_SUCCESS;
    }
    r.destroy(&p);
    if (!ok) {
        log = "hiprtc returned no code";
        return false;
    }
    // a register tile pushed to scratch memory keeps every result and runs
    // several times slower: such a kernel is not used
    if (!allow_scratch && scratch_bytes(code) != 0) {
        log = "the kernel uses scratch memory";
        return false;
    }
    return true;
}

// The kernel of one pass under `name`.  A wide pass is built for 4 waves per
// SIMD (128 VGPRs), again for 3 (168) if that spills, then for 2 (256);
// allow_scratch keeps the last attempt's code even if it spills (inspection).
bool compile_pass(const JitPass &p, const std::string &name, std::vector<char> &code, std::string &log,
                  bool allow_scratch) {
    if (p.flags & kJitSweep)
        return compile_source(sweep_source((const SweepKey *)p.gates, p.count, name), code, log, allow_scratch);
    if (p.flags & kJitAdjoint)
        return compile_source(adj_pass_source((const AGateRec *)p.gates, p.count, name), code, log, allow_scratch);
    const TGate *g = (const TGate *)p.gates;
    if (!(p.flags & kJitWide))
        return compile_source(pass_source(g, p.count, p.flags, name), code, log, allow_scratch, true);
    if (compile_source(pass_source(g, p.count, p.flags, name, 4), code, log, false, true)) return true;
    if (compile_source(pass_source(g, p.count, p.flags, name, 3), code, log, false, true)) return true;
    return compile_source(pass_source(g, p.count, p.flags, name, 2), code, log, allow_scratch, true);
}

constexpr int kMaxDevices = 64;

struct Kernel {
    std::string name;
    std::vector<char> code;
    bool failed = false;
    std::string log;
    hipModule_t mod[kMaxDevices] = {};
    hipFunction_t fn[kMaxDevices] = {};
};

struct Cache {
    std::mutex mu;
    std::map<std::string, std::shared_ptr<Kernel>> kernels;
    std::atomic<uint64_t> compiled{0}, hits{0}, failures{0}, generic{0};
    std::atomic<uint64_t> compile_us{0};
};

Cache &cache() {
    static Cache *c = new Cache;  // never destroyed: modules stay loaded until the process ends
    return *c;
}

}  // namespace

// Passes longer than this run the generic kernel: a longer straight line
// compiles slowly and gains little once a pass is bound by its gate work.
constexpr uint32_t kMaxSpecialisedGates = 256;

// Compiles (in parallel) every pass of the list not compiled yet, loads the
// kernels on `device` and returns one function per pass (null: run the
// generic kernel for it).  passes[i] = {host records, count}.
int32_t jit_prepare(int device, const std::vector<JitPass> &passes, std::vector<void *> &fns) {
    Cache &C = cache();
    fns.assign(passes.size(), nullptr);
    if (device < 0 || device >= kMaxDevices) return KBWH_OK;
    // one prepare at a time: a kernel in the cache is compiled (or failed)
    std::lock_guard<std::mutex> lk(C.mu);
    std::vector<std::shared_ptr<Kernel>> ks(passes.size());
    struct Todo {
        std::shared_ptr<Kernel> k;
        JitPass p;
    };
    std::vector<Todo> todo;
    {
        for (size_t i = 0; i < passes.size(); ++i) {
            const uint32_t cnt = passes[i].count, fl = passes[i].flags;
            if (cnt == 0 || cnt > kMaxSpecialisedGates) {
                C.generic++;
                continue;
            }
            const bool adj = (fl & kJitAdjoint) != 0;
            const TGate *g = (const TGate *)passes[i].gates;
            const AGateRec *ag = (const AGateRec *)passes[i].gates;
            std::string key = (fl & kJitSweep) ? sweep_key((const SweepKey *)passes[i].gates, cnt)
                              : adj ? adj_pass_key(ag, cnt) : pass_key(g, cnt, fl);
            auto it = C.kernels.find(key);
            if (it != C.kernels.end()) {
                ks[i] = it->second;
                C.hits++;
                continue;
            }
            auto k = std::make_shared<Kernel>();
            char nm[64];
            snprintf(nm, sizeof nm, "kbwh_pass_%016llx_%u", (unsigned long long)fnv1a(key), cnt);
            k->name = nm;
            C.kernels.emplace(key, k);
            ks[i] = k;
            todo.push_back({k, passes[i]});
        }
    }
    if (!todo.empty()) {
        const auto t0 = std::chrono::steady_clock::now();
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t j; (j = next.fetch_add(1)) < todo.size();) {
                Kernel &k = *todo[j].k;
                k.failed = !compile_pass(todo[j].p, k.name, k.code, k.log, false);
            }
        };
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>(todo.size(), std::min(16u, hw));
        std::vector<std::thread> th;
        for (size_t t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto &t : th) t.join();
        C.compile_us += (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
        for (auto &w : todo) (w.k->failed ? C.failures : C.compiled)++;
    }
    // modules load into the current device: make the state's device current
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    for (size_t i = 0; e == hipSuccess && i < passes.size(); ++i) {
        Kernel *k = ks[i].get();
        if (!k || k->failed) continue;
        if (!k->fn[device]) {
            e = hipModuleLoadData(&k->mod[device], k->code.data());
            if (e == hipSuccess) e = hipModuleGetFunction(&k->fn[device], k->mod[device], k->name.c_str());
            if (e != hipSuccess) break;
        }
        fns[i] = (void *)k->fn[device];
    }
    if (cur >= 0 && cur != device) {
        const hipError_t r = hipSetDevice(cur);
        if (e == hipSuccess) e = r;
    }
    if (e != hipSuccess) return hip_status(e, "loading a compiled pass kernel");
    return KBWH_OK;
}

hipError_t jit_launch(void *fn, double2 *amp, const void *spec, const void *gates, uint32_t ntiles,
                      hipStream_t stream, uint32_t set_shift, uint64_t set_mask, const void *store_spec,
                      uint32_t order) {
    TileSpec ts, tss;
    memcpy(&ts, spec, sizeof ts);
    if (store_spec) memcpy(&tss, store_spec, sizeof tss);
    dv2 *a = (dv2 *)amp;
    const TGate *g = (const TGate *)gates;
    void *args[] = {&a, &ts, &g, &ntiles, &set_shift, &set_mask, &order, &tss};
    return hipModuleLaunchKernel((hipFunction_t)fn, ntiles, 1, 1, TT, 1, 1, 0, stream, args, nullptr);
}

// ---- tile-order tuner ----
// Which tile order streams a pass best depends on its tile qubits (how the
// tile's amplitudes and those of the tiles in flight beside it fall on the
// HBM channels) and on how long its gates keep each workgroup: on the
// 30-qubit brickwork circuit single passes moved from 10.1 to 7.2 ms and
// from 9.3 to 8.2 ms between orders, others not at all, and no fixed order
// was best for most passes.  So each pass layout measures the candidates on
// its first runs (one per run, on the state it is run on: the order does not
// change the results) and keeps the fastest.
namespace {

constexpr uint32_t kOrders[] = {3, 5, 9, 0};
constexpr uint32_t kNumOrders = (uint32_t)(sizeof kOrders / sizeof kOrders[0]);
// smaller states are not tuned (fewer tiles than a few waves of workgroups)
constexpr uint32_t kMinTunedTiles = 1u << 12;
// layouts kept for the process; beyond this, new layouts run the default order
constexpr size_t kMaxTunedLayouts = 16384;

struct OrderSlot {
    float ms[kNumOrders];
    uint32_t next = 0, reported = 0;
    uint32_t best = kDefaultTileOrder;
    OrderSlot() {
        for (float &x : ms) x = -1.0f;
    }
};

struct OrderTuner {
    std::mutex mu;
    std::map<std::string, OrderSlot> slots;
    std::atomic<bool> enabled{true};
    std::atomic<uint64_t> trials{0};
};

OrderTuner &tuner() {
    static OrderTuner *t = new OrderTuner;  // kept until the process ends
    return *t;
}

}  // namespace

OrderTrial order_pick(int device, void *fn, const void *spec, size_t spec_bytes, uint32_t ntiles) {
    OrderTrial t{nullptr, kDefaultTileOrder, {nullptr, nullptr}};
    OrderTuner &T = tuner();
    if (ntiles < kMinTunedTiles) return t;
    std::string key;
    key.append((const char *)&device, sizeof device);
    key.append((const char *)&fn, sizeof fn);
    key.append((const char *)&ntiles, sizeof ntiles);
    key.append((const char *)spec, spec_bytes);
    std::lock_guard<std::mutex> lk(T.mu);
    auto it = T.slots.find(key);
    if (it == T.slots.end()) {
        if (T.slots.size() >= kMaxTunedLayouts) return t;
        it = T.slots.emplace(key, OrderSlot()).first;
    }
    OrderSlot &o = it->second;
    if (!T.enabled.load() || o.next >= kNumOrders) {
        t.order = T.enabled.load() ? o.best : kDefaultTileOrder;
        return t;
    }
    if (hipEventCreate(&t.ev[0]) != hipSuccess) {
        (void)hipGetLastError();
        return t;
    }
    if (hipEventCreate(&t.ev[1]) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(t.ev[0]);
        t.ev[0] = nullptr;
        return t;
    }
    t.slot = &o;  // map nodes stay where they are
    t.order = kOrders[o.next++];
    T.trials++;
    return t;
}

void order_report(OrderTrial &t) {
    if (!t.slot) return;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, t.ev[0], t.ev[1]) != hipSuccess) {
        (void)hipGetLastError();
        ms = -1.0f;
    }
    (void)hipEventDestroy(t.ev[0]);
    (void)hipEventDestroy(t.ev[1]);
    OrderTuner &T = tuner();
    std::lock_guard<std::mutex> lk(T.mu);
    OrderSlot &o = *(OrderSlot *)t.slot;
    t.slot = nullptr;
    uint32_t idx = 0;
    while (idx < kNumOrders && kOrders[idx] != t.order) ++idx;
    if (idx == kNumOrders || ms < 0.0f) return;
    o.ms[idx] = ms;
    if (++o.reported < kNumOrders) return;
    uint32_t b = 0;  // ties keep the earlier candidate (the default first)
    for (uint32_t i = 1; i < kNumOrders; ++i)
        if (o.ms[i] >= 0.0f && (o.ms[b] < 0.0f || o.ms[i] < o.ms[b])) b = i;
    o.best = kOrders[b];
}

hipError_t jit_launch_sweep(void *fn, const double2 *amp, const void *spec, const void *strings, uint32_t ntiles,
                            double *partial, uint32_t grid, hipStream_t stream) {
    TileSpec ts;
    memcpy(&ts, spec, sizeof ts);
    const dv2 *a = (const dv2 *)amp;
    const void *ss = strings;
    void *args[] = {&a, &ts, &ss, &ntiles, &partial};
    return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, TT, 1, 1, 0, stream, args, nullptr);
}

hipError_t jit_launch_spec(void *fn, double2 *amp, const void *spec, size_t spec_bytes, const void *gates,
                           uint32_t ntiles, hipStream_t stream, uint32_t order, uint32_t threads) {
    alignas(8) unsigned char ts[256];
    if (spec_bytes > sizeof ts) return hipErrorInvalidValue;
    memcpy(ts, spec, spec_bytes);
    dv2 *a = (dv2 *)amp;
    const void *g = gates;
    uint32_t set_shift = 32;
    uint64_t set_mask = ~0ull;
    void *args[] = {&a, ts, &g, &ntiles, &set_shift, &set_mask, &order};
    return hipModuleLaunchKernel((hipFunction_t)fn, ntiles, 1, 1, threads, 1, 1, 0, stream, args, nullptr);
}

int32_t jit_code(const JitPass &p, std::vector<char> &code) {
    std::string log;
    if (!compile_pass(p, "kbwh_pass", code, log, true))
        return set_error(KBWH_ERR_EXECUTION_FAILED, "pass kernel did not compile: " + log.substr(0, 400));
    return KBWH_OK;
}

hipError_t jit_launch_adj(void *fn, double2 *amp, const void *spec, size_t spec_bytes, const void *gates, uint32_t tps,
                          uint32_t wps, double *partial, uint32_t grid, hipStream_t stream) {
    alignas(8) unsigned char ts[256];
    if (spec_bytes > sizeof ts) return hipErrorInvalidValue;
    memcpy(ts, spec, spec_bytes);
    dv2 *a = (dv2 *)amp;
    const void *g = gates;
    void *args[] = {&a, ts, &g, &tps, &wps, &partial};
    return hipModuleLaunchKernel((hipFunction_t)fn, grid, 1, 1, TT, 1, 1, 0, stream, args, nullptr);
}

}  // namespace kbwh

extern "C" int32_t kbwh_set_order_tuning(int32_t enable) {
    kbwh::tuner().enabled = enable != 0;
    return KBWH_OK;
}

extern "C" int32_t kbwh_order_tuning_stats(uint64_t *layouts_tuned, uint64_t *layouts_pending, uint64_t *trials,
                                           uint64_t *non_default) {
    using namespace kbwh;
    OrderTuner &T = tuner();
    std::lock_guard<std::mutex> lk(T.mu);
    uint64_t tuned = 0, pending = 0, moved = 0;
    for (const auto &kv : T.slots) {
        if (kv.second.reported >= kNumOrders) {
            ++tuned;
            moved += kv.second.best != kDefaultTileOrder;
        } else {
            ++pending;
        }
    }
    if (layouts_tuned) *layouts_tuned = tuned;
    if (layouts_pending) *layouts_pending = pending;
    if (trials) *trials = T.trials.load();
    if (non_default) *non_default = moved;
    return KBWH_OK;
}

extern "C" int32_t kbwh_compiled_stats(uint64_t *kernels_compiled, uint64_t *cache_hits, uint64_t *failures,
                                       uint64_t *generic_passes, double *compile_seconds) {
    using namespace kbwh;
    Cache &C = cache();
    if (kernels_compiled) *kernels_compiled = C.compiled;
    if (cache_hits) *cache_hits = C.hits;
    if (failures) *failures = C.failures;
    if (generic_passes) *generic_passes = C.generic;
    if (compile_seconds) *compile_seconds = (double)C.compile_us * 1e-6;
    return KBWH_OK;
}

// The code object of the specialised kernel of pass `pass` of the
// dependency-ordered plan of ops on a flat state of num_qubits qubits (host
// only, no device).  With code = NULL, *size receives the size.
extern "C" int32_t kbwh_compiled_pass_code(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits,
                                           uint32_t pass, void *code, uint64_t *size) {
    using namespace kbwh;
    if (!ops || !size) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (num_qubits < TL || num_qubits > 40) return set_error(KBWH_ERR_INVALID_ARGUMENT, "11..40 qubits");
    if (num_qubits >= kWideTileQubits) {
        // the wide tiles of KBWH_RUN_COMPILED
        std::vector<unsigned char> recs;
        uint32_t count = 0;
        int32_t rc = wide_pass_records(ops, num_ops, num_qubits, pass, recs, count);
        if (rc) return rc;
        std::vector<char> obj;
        if ((rc = jit_code(JitPass{recs.data(), count, kJitWide}, obj))) return rc;
        if (code) {
            if (*size < obj.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "buffer too small");
            memcpy(code, obj.data(), obj.size());
        }
        *size = obj.size();
        return KBWH_OK;
    }
    std::vector<PhysOp> phys(num_ops);
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ || op.target >= num_qubits || (op.ctrl_mask >> num_qubits))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "op out of range");
        phys[i] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.angle};
    }
    fuse_diagonals(phys, num_qubits);
    std::vector<uint32_t> order, ends;
    schedule_passes(phys, num_qubits, TL - 3, order, ends);
    if (pass >= ends.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "no such pass");
    Pass cur;
    for (size_t k = pass ? ends[pass - 1] : 0; k < ends[pass]; ++k) add_op(cur, phys[order[k]]);
    int pos_of[64];
    uint32_t q_at[TL];
    assign_positions(num_qubits, cur, pos_of, q_at);
    std::vector<TGate> g;
    for (const PhysOp &o : cur.ops) g.push_back(make_tgate(o, pos_of, num_qubits));
    std::vector<char> obj;
    const int32_t rc = jit_code(JitPass{g.data(), (uint32_t)g.size(), 0u}, obj);
    if (rc) return rc;
    if (code) {
        if (*size < obj.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "buffer too small");
        memcpy(code, obj.data(), obj.size());
    }
    *size = obj.size();
    return KBWH_OK;
}
