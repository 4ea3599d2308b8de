// Tile machinery shared by the tiled passes (planned in tiled.hip, built in
// tiled_pass.hip and tiled_fma.hip), the batched persistent kernel (batch.hip), the gate queue (live.hip), the tiled
// adjoint sweep (adjoint_tiled.hip) and the fused-block planner (fused.hip):
// the tile layout, the gate record, the exact per-amplitude gate application
// in registers, and the host-side planners (program order and dependency
// order).  Each unit gets its own copy (internal linkage) and may pick its own
// tile shape.
#pragma once

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "state.h"
#include "tile_dev.h"

namespace kbwh {
namespace tile {
namespace {

inline bool is_diagonal(int kind) { return kind == KZ || kind == KP || kind == KRZ; }

// A gate with its physical target and controls, ready for a pass.
struct PhysOp {
    int kind;
    uint32_t target;
    uint64_t ctrl;  // physical control mask without the target bit
    double angle;
    // a fused diagonal gate carries its phases {|0> re, im, |1> re, im}
    bool phased = false;
    double ph[4] = {0.0, 0.0, 0.0, 0.0};
};

// The phases of an uncontrolled diagonal op, as set_coefficients gives them.
inline void diag_phases(const PhysOp &o, double ph[4]) {
    if (o.phased) {
        memcpy(ph, o.ph, sizeof o.ph);
        return;
    }
    const Coef k = make_coef(o.kind, o.angle);
    if (o.kind == KZ) {
        ph[0] = 1.0, ph[1] = 0.0, ph[2] = -1.0, ph[3] = -0.0;
    } else if (o.kind == KP) {
        ph[0] = 1.0, ph[1] = 0.0, ph[2] = k.ph1.x, ph[3] = k.ph1.y;
    } else {
        ph[0] = k.ph0.x, ph[1] = k.ph0.y, ph[2] = k.ph1.x, ph[3] = k.ph1.y;
    }
}

// Fuses uncontrolled diagonal gates (dependency-ordered runs only: the fused
// phases are products of the gates' phases, equal to rounding).  A Z, P or RZ
// on qubit q merges into the last earlier uncontrolled diagonal gate on q
// when every op between them that touches q is diagonal on q: a diagonal gate
// on q, or a gate with q among its controls; those commute with it.  The
// result is a P when its |0> phase is exactly 1, nothing when both phases
// are exactly 1, else a general diagonal (the RZ form with its own phases).
inline void fuse_diagonals(std::vector<PhysOp> &ops, uint32_t n) {
    std::vector<PhysOp> out;
    out.reserve(ops.size());
    std::vector<std::vector<uint32_t>> touch(n);  // ops of out touching qubit q
    for (const PhysOp &o : ops) {
        if (is_diagonal(o.kind) && o.ctrl == 0) {
            const std::vector<uint32_t> &lst = touch[o.target];
            int64_t j = -1;
            for (auto it = lst.rbegin(); it != lst.rend(); ++it) {
                const PhysOp &p = out[*it];
                if (p.kind < 0) continue;  // fused away
                if (is_diagonal(p.kind) && p.ctrl == 0 && p.target == o.target) {
                    j = (int64_t)*it;
                    break;
                }
                if (!is_diagonal(p.kind) && p.target == o.target) break;
            }
            if (j >= 0) {
                PhysOp &m = out[(size_t)j];
                double a[4], b[4], r[4];
                diag_phases(m, a);
                diag_phases(o, b);
                r[0] = a[0] * b[0] - a[1] * b[1];
                r[1] = a[0] * b[1] + a[1] * b[0];
                r[2] = a[2] * b[2] - a[3] * b[3];
                r[3] = a[2] * b[3] + a[3] * b[2];
                const bool one0 = r[0] == 1.0 && r[1] == 0.0;
                m.kind = one0 && r[2] == 1.0 && r[3] == 0.0 ? -1 : (one0 ? KP : KRZ);
                m.phased = true;
                memcpy(m.ph, r, sizeof r);
                continue;
            }
        }
        out.push_back(o);
        const uint32_t idx = (uint32_t)(out.size() - 1);
        touch[o.target].push_back(idx);
        for (uint64_t c = o.ctrl; c; c &= c - 1) touch[__builtin_ctzll(c)].push_back(idx);
    }
    ops.clear();
    for (const PhysOp &o : out)
        if (o.kind >= 0) ops.push_back(o);
}

struct Pass {
    std::vector<PhysOp> ops;
    std::vector<uint32_t> targets;  // distinct non-diagonal targets, in first-use order
};

// Tile positions for a pass: qubits 0..2 on lane bits 0..2; the most used
// targets on register bits, then lane bits 3..5, then wave bits; free
// positions take the lowest qubits not used yet.
inline void assign_positions(uint32_t n, const Pass &p, int pos_of[64], uint32_t q_at[TL]) {
    for (int i = 0; i < 64; ++i) pos_of[i] = -1;
    for (int b = 0; b < TL; ++b) q_at[b] = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < 3; ++q) {
        q_at[q] = q;
        pos_of[q] = (int)q;
    }
    std::vector<std::pair<int, uint32_t>> freq;
    for (uint32_t t : p.targets) {
        if (t < 3) continue;
        int f = 0;
        for (const PhysOp &o : p.ops)
            if (!is_diagonal(o.kind) && o.target == t) ++f;
        freq.push_back({-f, t});
    }
    std::stable_sort(freq.begin(), freq.end(),
                     [](const std::pair<int, uint32_t> &x, const std::pair<int, uint32_t> &y) {
                         return x.first < y.first;
                     });
    int order[TL - 3], fill_order[TL - 3];
    {
        int o = 0;
        for (int b = TB; b < TL; ++b) order[o++] = b;  // register bits first
        for (int b = 3; b < TB; ++b) order[o++] = b;   // then lane bits 3..5, then wave bits
        for (int b = 0; b < TL - 3; ++b) fill_order[b] = 3 + b;
    }
    size_t k = 0;
    for (; k < freq.size(); ++k) {
        q_at[order[k]] = freq[k].second;
        pos_of[freq[k].second] = order[k];
    }
    // free positions take the lowest unused qubits (n >= TL leaves enough)
    uint32_t filler = 3;
    for (int i = 0; i < TL - 3; ++i) {
        const int b = fill_order[i];
        if (q_at[b] != 0xFFFFFFFFu) continue;
        while (pos_of[filler] >= 0) ++filler;
        q_at[b] = filler;
        pos_of[filler] = b;
        ++filler;
    }
}


// The generic pair coefficients / diagonal phases of a gate (see pair_out).
inline void set_coefficients(TGate &g, int kind, double angle) {
    const Coef k = make_coef(kind, angle);
    g.d = 1.0;
    switch (kind) {
    case KX: g.bu = 1.0; break;
    case KY: g.crs = -1.0; g.cis = 1.0; break;
    case KH: g.as = -1.0; g.bu = 1.0; g.d = 0.70710678118654752440; break;
    case KRX: g.au = k.c; g.cru = k.s; g.ciu = -k.s; break;
    case KRY: g.au = k.c; g.bs = k.s; break;
    case KZ: g.ph0x = 1.0; g.ph1x = -1.0; g.ph1y = -0.0; break;  // dense_v2.rs:191
    case KP: g.ph0x = 1.0; g.ph1x = k.ph1.x; g.ph1y = k.ph1.y; break;
    default: g.ph0x = k.ph0.x; g.ph0y = k.ph0.y; g.ph1x = k.ph1.x; g.ph1y = k.ph1.y; break;
    }
}

// The coefficients of a parameterised RX, RY, RZ or P record from its angle,
// on the device: the forms of set_coefficients with the device library's
// sincos, so they agree with the host's libm values to rounding.
__device__ __forceinline__ void set_coefficients_dev(TGate &g, double th) {
    double s, c;
    switch (g.kind) {
    case KRX:
        sincos(th / 2.0, &s, &c);
        g.au = c;
        g.cru = s;
        g.ciu = -s;
        break;
    case KRY:
        sincos(th / 2.0, &s, &c);
        g.au = c;
        g.bs = s;
        break;
    case KP:
        sincos(th, &s, &c);
        g.ph0x = 1.0;
        g.ph0y = 0.0;
        g.ph1x = c;
        g.ph1y = s;
        break;
    default:  // KRZ
        sincos(-th / 2.0, &s, &c);
        g.ph0x = c;
        g.ph0y = s;
        sincos(th / 2.0, &s, &c);
        g.ph1x = c;
        g.ph1y = s;
        break;
    }
}

// The device record of one gate of a pass, relative to the pass's tile positions.
inline TGate make_tgate(const PhysOp &o, const int pos_of[64], uint32_t n) {
    TGate g;
    memset(&g, 0, sizeof g);
    g.kind = (uint32_t)o.kind;
    g.tpos = pos_of[o.target] >= 0 ? (uint32_t)pos_of[o.target] : kOutside;
    g.otb = 1ull << o.target;
    for (uint32_t q = 0; q < n; ++q) {
        if (!(o.ctrl >> q & 1ull)) continue;
        if (pos_of[q] >= 0)
            g.cm |= 1u << pos_of[q];
        else
            g.ocm |= 1ull << q;
    }
    set_coefficients(g, o.kind, o.angle);
    if (o.phased) {
        g.ph0x = o.ph[0];
        g.ph0y = o.ph[1];
        g.ph1x = o.ph[2];
        g.ph1y = o.ph[3];
    }
    if (is_diagonal(o.kind)) {
        g.dlo = g.cm & (uint32_t)(TT - 1);
        g.dhi = g.cm >> TB;
        if (o.kind != KRZ) {
            // Z and P change the amplitudes whose target and controls are all
            // 1: a target outside the tile acts as one more outside control
            if (g.tpos == kOutside)
                g.ocm |= g.otb;
            else if (g.tpos < (uint32_t)TB)
                g.dlo |= 1u << g.tpos;
            else
                g.dhi |= 1u << (g.tpos - TB);
        }
    }
    return g;
}

inline TileSpec make_tile_spec(const uint32_t q_at[TL]) {
    TileSpec ts;
    uint64_t qmask = 0;
    for (int b = 0; b < TL; ++b) {
        ts.qbit[b] = 1ull << q_at[b];
        qmask |= 1ull << q_at[b];
    }
    const Insert ins = make_insert(qmask);
    for (int j = 0; j < TL; ++j) ts.qlo[j] = ins.mlo[j];
    return ts;
}

// Greedy in-order pass building: a new non-diagonal target that does not fit
// the free tile positions (at most `cap` of the TL - 3) closes the pass.
inline bool fits(const Pass &p, const PhysOp &o, uint32_t cap = TL - 3) {
    if (is_diagonal(o.kind) || o.target < 3) return true;
    if (std::find(p.targets.begin(), p.targets.end(), o.target) != p.targets.end()) return true;
    return p.targets.size() < (size_t)std::min<uint32_t>(cap, TL - 3);
}

inline void add_op(Pass &p, const PhysOp &o) {
    if (!is_diagonal(o.kind) && o.target >= 3 &&
        std::find(p.targets.begin(), p.targets.end(), o.target) == p.targets.end())
        p.targets.push_back(o.target);
    p.ops.push_back(o);
}

// The plan search's random source (xorshift64; the same plan for the same seed).
struct PlanRng {
    uint64_t s;
    explicit PlanRng(uint64_t seed) {
        // splitmix64 of the seed: neighbouring seeds start far apart
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        s = (z ^ (z >> 31)) | 1ull;
    }
    double uniform() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
};

// Dependency-ordered pass planning (KBWH_RUN_REORDER).  Ops that share no
// qubit commute, so any order that keeps, on every qubit, the order of the ops
// touching it is the same unitary; only the rounding of the amplitudes can
// change.  Op i touches the qubits of qm[i] and needs the qubits of req[i] in
// the tile (a non-diagonal gate its target, a Block swap both its qubits).
// An op is ready when it is the first pending op on each of its qubits.  A
// pass starts with tile qubits 0..2 and runs every ready op whose qubits are
// in the tile, again and again; while target positions are left, the missing
// qubits of the ready op that lets the most pending ops run (the lowest such
// set on ties) join the tile.
// order: op indices in execution order; ends[p]: end of pass p in order.
// limited (optional): ops of which a pass may hold at most max_limited.
// rng (optional, the plan search): a candidate's score is the ops it lets run
// plus half of what the best next candidate would add, times a random factor
// in [1, 1.3).
inline void schedule_ops(const std::vector<uint64_t> &qm, const std::vector<uint64_t> &req, uint32_t n,
                         uint32_t free_targets, std::vector<uint32_t> &order, std::vector<uint32_t> &ends,
                         const std::vector<uint8_t> *limited = nullptr, uint32_t max_limited = 0,
                         PlanRng *rng = nullptr) {
    const size_t m = qm.size();
    std::vector<uint32_t> start(n + 1, 0);
    for (size_t i = 0; i < m; ++i)
        for (uint64_t r = qm[i]; r; r &= r - 1) ++start[__builtin_ctzll(r) + 1];
    for (uint32_t q = 0; q < n; ++q) start[q + 1] += start[q];
    std::vector<uint32_t> list(start[n]);
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < m; ++i)
            for (uint64_t r = qm[i]; r; r &= r - 1) list[fill[__builtin_ctzll(r)]++] = (uint32_t)i;
    }
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    auto front = [&](uint32_t q, const std::vector<uint32_t> &h) {
        return h[q] < start[q + 1] ? list[h[q]] : kNone;
    };
    auto ready = [&](uint32_t i, const std::vector<uint32_t> &h) {
        for (uint64_t r = qm[i]; r; r &= r - 1)
            if (front(__builtin_ctzll(r), h) != i) return false;
        return true;
    };
    std::vector<uint32_t> work;
    uint32_t pass_limited = 0;  // limited ops in the pass being built
    // run every op that can run in `tile`; appends them to out when given
    // (and then counts the limited ones into the pass)
    auto closure = [&](std::vector<uint32_t> &h, uint64_t tile, std::vector<uint32_t> *out) {
        size_t count = 0;
        uint32_t nl = pass_limited;
        work.clear();
        for (uint32_t q = 0; q < n; ++q) work.push_back(q);
        while (!work.empty()) {
            const uint32_t q = work.back();
            work.pop_back();
            const uint32_t i = front(q, h);
            if (i == kNone || (req[i] & ~tile) || !ready(i, h)) continue;
            const bool lim = limited && (*limited)[i];
            if (lim && nl >= max_limited) continue;
            nl += lim;
            for (uint64_t r = qm[i]; r; r &= r - 1) {
                const uint32_t p = __builtin_ctzll(r);
                ++h[p];
                work.push_back(p);
            }
            ++count;
            if (out) out->push_back(i);
        }
        if (out) pass_limited = nl;
        return count;
    };
    std::vector<uint32_t> head(start.begin(), start.end() - 1), trial, trial2;
    std::vector<uint64_t> cands, cands2;
    // the missing qubits of each ready op that still fit, ascending
    auto candidates = [&](const std::vector<uint32_t> &h, uint64_t tile, uint32_t used, std::vector<uint64_t> &out) {
        out.clear();
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t i = front(q, h);
            if (i == kNone) continue;
            const uint64_t miss = req[i] & ~tile;
            if (!miss || used + (uint32_t)__builtin_popcountll(miss) > free_targets || !ready(i, h)) continue;
            if (std::find(out.begin(), out.end(), miss) == out.end()) out.push_back(miss);
        }
        std::sort(out.begin(), out.end());
    };
    size_t left = m;
    order.clear();
    ends.clear();
    while (left) {
        uint64_t tile = 7;
        uint32_t used = 0;
        pass_limited = 0;
        const size_t before = order.size();
        left -= closure(head, tile, &order);
        while (used < free_targets && left) {
            candidates(head, tile, used, cands);
            if (cands.empty()) break;
            uint64_t best = cands[0];
            double best_score = 0.0;
            for (const uint64_t c : cands) {
                trial = head;
                double k = (double)closure(trial, tile | c, nullptr);
                if (rng) {
                    const uint32_t u2 = used + (uint32_t)__builtin_popcountll(c);
                    size_t next = 0;
                    if (u2 < free_targets) {
                        candidates(trial, tile | c, u2, cands2);
                        for (const uint64_t d : cands2) {
                            trial2 = trial;
                            next = std::max(next, closure(trial2, tile | c | d, nullptr));
                        }
                    }
                    k = (k + 0.5 * (double)next) * (1.0 + 0.3 * rng->uniform());
                }
                if (k > best_score) {
                    best_score = k;
                    best = c;
                }
            }
            tile |= best;
            used += (uint32_t)__builtin_popcountll(best);
            left -= closure(head, tile, &order);
        }
        // an op needing more tile positions than there are stops the plan
        // short (order then holds fewer than m ops; callers check)
        if (order.size() == before) break;
        ends.push_back((uint32_t)order.size());
    }
}


// Plan search: the greedy plan of schedule_ops and `trials` randomised plans
// (seeds 1..trials, on up to 16 host threads); the plan with the fewest
// passes wins, ties going to the greedy plan, then to the lowest seed, so the
// result does not depend on the thread count.  On the 30-qubit, 20-layer
// brickwork circuit a few hundred trials take 8 target positions from 21
// passes to 19-20 and 9 positions from 18 to 16-17.
inline void schedule_ops_search(const std::vector<uint64_t> &qm, const std::vector<uint64_t> &req, uint32_t n,
                                uint32_t free_targets, uint32_t trials, std::vector<uint32_t> &order,
                                std::vector<uint32_t> &ends, const std::vector<uint8_t> *limited = nullptr,
                                uint32_t max_limited = 0) {
    schedule_ops(qm, req, n, free_targets, order, ends, limited, max_limited);
    if (trials == 0) return;
    const size_t m = qm.size();
    auto passes_of = [&](const std::vector<uint32_t> &o, const std::vector<uint32_t> &e) {
        return o.size() == m ? e.size() : (size_t)-1;
    };
    struct Best {
        size_t passes = (size_t)-1;
        uint32_t seed = 0;
        std::vector<uint32_t> order, ends;
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t nt = std::max<uint32_t>(1u, std::min<uint32_t>(trials, std::min(16u, hw)));
    std::vector<Best> best(nt);
    std::atomic<uint32_t> next{1};
    auto work = [&](uint32_t t) {
        std::vector<uint32_t> o, e;
        for (uint32_t seed; (seed = next.fetch_add(1)) <= trials;) {
            PlanRng rng(seed);
            schedule_ops(qm, req, n, free_targets, o, e, limited, max_limited, &rng);
            const size_t p = passes_of(o, e);
            Best &b = best[t];
            if (p < b.passes || (p == b.passes && seed < b.seed)) {
                b.passes = p;
                b.seed = seed;
                b.order.swap(o);
                b.ends.swap(e);
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    const Best *w = nullptr;
    for (const Best &b : best)
        if (b.passes != (size_t)-1 && (!w || b.passes < w->passes || (b.passes == w->passes && b.seed < w->seed)))
            w = &b;
    if (w && w->passes < passes_of(order, ends)) {
        order = w->order;
        ends = w->ends;
    }
}

// schedule_ops for gates: a gate touches its target and controls and needs
// its target in the tile unless it is diagonal.
// trials > 0: with the plan search.
inline void schedule_passes(const std::vector<PhysOp> &ops, uint32_t n, uint32_t free_targets,
                            std::vector<uint32_t> &order, std::vector<uint32_t> &ends,
                            const std::vector<uint8_t> *limited = nullptr, uint32_t max_limited = 0,
                            uint32_t trials = 0) {
    std::vector<uint64_t> qm(ops.size()), req(ops.size());
    for (size_t i = 0; i < ops.size(); ++i) {
        qm[i] = ops[i].ctrl | (1ull << ops[i].target);
        req[i] = is_diagonal(ops[i].kind) ? 0ull : (1ull << ops[i].target);
    }
    schedule_ops_search(qm, req, n, free_targets, trials, order, ends, limited, max_limited);
}

// Searched plans of dependency-ordered runs, by the structure of their ops
// (kind class, target and controls of each op in order, the qubit count, the
// target cap and the trial count): a circuit run again, with the same or
// other angles, reuses its plan.  At most 64 plans, the oldest dropped first.
inline void schedule_passes_cached(const std::vector<PhysOp> &ops, uint32_t n, uint32_t free_targets, uint32_t trials,
                                   std::vector<uint32_t> &order, std::vector<uint32_t> &ends) {
    if (trials == 0) {
        schedule_passes(ops, n, free_targets, order, ends);
        return;
    }
    std::string key;
    key.reserve(16 + ops.size() * 16);
    const uint32_t head[4] = {n, free_targets, trials, (uint32_t)TL};
    key.append((const char *)head, sizeof head);
    for (const PhysOp &o : ops) {
        const uint64_t f[2] = {(uint64_t)o.target | (is_diagonal(o.kind) ? 1ull << 32 : 0ull), o.ctrl};
        key.append((const char *)f, sizeof f);
    }
    struct Cache {
        std::mutex mu;
        std::map<std::string, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> plans;
        std::vector<std::string> fifo;
    };
    static Cache *cache = new Cache;  // kept until the process ends
    {
        std::lock_guard<std::mutex> lk(cache->mu);
        auto it = cache->plans.find(key);
        if (it != cache->plans.end()) {
            order = it->second.first;
            ends = it->second.second;
            return;
        }
    }
    schedule_passes(ops, n, free_targets, order, ends, nullptr, 0, trials);
    std::lock_guard<std::mutex> lk(cache->mu);
    if (cache->plans.emplace(key, std::make_pair(order, ends)).second) {
        cache->fifo.push_back(key);
        if (cache->fifo.size() > 64) {
            cache->plans.erase(cache->fifo.front());
            cache->fifo.erase(cache->fifo.begin());
        }
    }
}

}  // namespace
}  // namespace tile
}  // namespace kbwh
