/*
 * Oracle restatement of the QubitManager measure loop and the
 * QuantumExecution trait defaults (src/quantum_execution.rs, src/convert.rs).
 * TEST INFRASTRUCTURE.
 *
 * Random numbers are not drawn here: every routine takes host-drawn uniforms
 * u in [0, 1), one per draw, in the order the reference draws them.  The
 * reference draws through rand 0.10.2 (Cargo.toml:23, not vendored); its
 * published algorithm is restated in orc_uniform_scale / the WeightedIndex
 * pick below.  A Rust caller would pass value0_1 = (next_u64() >> 12) * 2^-52
 * from its StdRng.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kbw_oracle.h"
#include "oracle_internal.h"

/*
 * rand's UniformFloat::new(low = 0, high = total): scale = high - low, then
 * decreased one ulp at a time while max_rand * scale + low >= high, with
 * max_rand = 1 - 2^-52 (the largest value0_1).  A draw is u * scale + low.
 */
double orc_uniform_scale(double total) {
    const double max_rand = 1.0 - DBL_EPSILON;
    double scale = total - 0.0;
    while (scale * max_rand + 0.0 >= total) {
        uint64_t bits;
        memcpy(&bits, &scale, sizeof bits);
        bits -= 1;
        memcpy(&scale, &bits, sizeof bits);
    }
    return scale;
}

/*
 * quantum_execution.rs:272-292 (do_measure).  Qubits are measured in reverse
 * list order; the k-th measured qubit consumes uniforms[k].
 * WeightedIndex::new([p0, p1]) has cumulative_weights = [p0] and
 * total = p0 + p1; sample() = #{w in cumulative : w <= x} = (p0 <= x).
 */
uint64_t orc_measure(orc_v2 *s, const uint32_t *qubits, uint32_t nq, const double *uniforms) {
    uint64_t out = 0;
    for (uint32_t idx = 0; idx < nq; idx++) {
        const uint32_t q = qubits[nq - 1 - idx];
        const double p1 = orc_v2_measure_p1(s, q);
        const double p0 = fmax(1.0 - p1, 0.0);
        const double total = p0 + p1;
        const double x = uniforms[idx] * orc_uniform_scale(total) + 0.0;
        const int result = (p0 <= x) ? 1 : 0;
        const double prob = result ? p1 : p0;
        const double p = 1.0 / sqrt(fmax(prob, DBL_MIN));
        orc_v2_measure_collapse(s, q, result, p);
        out |= (uint64_t)result << idx;
    }
    return out;
}

/* Basis changes of quantum_execution.rs:114-121 and their undo at :144-151. */
static void basis_change(orc_v2 *s, const uint8_t *paulis, const uint32_t *qubits, uint32_t lo,
                         uint32_t hi, int undo) {
    const double frac_pi_2 = 1.57079632679489661923; /* std::f64::consts::FRAC_PI_2 */
    for (uint32_t k = lo; k < hi; k++) {
        if (paulis[k] == ORC_PAULI_X) {
            orc_v2_gate_mask(s, ORC_H, 0.0, qubits[k], 0);
        } else if (paulis[k] == ORC_PAULI_Y) {
            if (!undo) {
                orc_v2_gate_mask(s, ORC_P, -frac_pi_2, qubits[k], 0);
                orc_v2_gate_mask(s, ORC_H, 0.0, qubits[k], 0);
            } else {
                orc_v2_gate_mask(s, ORC_H, 0.0, qubits[k], 0);
                orc_v2_gate_mask(s, ORC_P, frac_pi_2, qubits[k], 0);
            }
        }
    }
}

/*
 * Default exp_value (quantum_execution.rs:109-158) with dump (dense_v2.rs:
 * 308-336) and from_dump_to_prob (convert.rs:30-42) streamed instead of
 * materialised: the same entries (hypot > 1e-15, ascending index), the same
 * probability hypot^2, the same parity of the dump key and the same
 * sequential summation, without the O(2^n) DumpData allocation.
 * String k is paulis/qubits[offsets[k] .. offsets[k+1]).
 */
double orc_exp_value(orc_v2 *s, const uint8_t *paulis, const uint32_t *qubits,
                     const uint32_t *offsets, const double *coeffs, uint32_t nstrings) {
    const orc_c64 *a = orc_v2_data(s);
    const uint64_t size = (uint64_t)1 << orc_v2_num_qubits(s);
    double total = 0.0;
    for (uint32_t k = 0; k < nstrings; k++) {
        const uint32_t lo = offsets[k], hi = offsets[k + 1];
        basis_change(s, paulis, qubits, lo, hi, 0);
        double result = 0.0;
        for (uint64_t i = 0; i < size; i++) {
            const double h = hypot(a[i].re, a[i].im);
            if (!(h > ORC_SMALL_EPSILON)) continue;
            const double prob = h * h;
            const uint64_t key = orc_dump_key(i, qubits + lo, hi - lo);
            const double parity = (__builtin_popcountll(key) % 2 == 0) ? 1.0 : -1.0;
            result += prob * parity;
        }
        basis_change(s, paulis, qubits, lo, hi, 1);
        total += result * coeffs[k];
    }
    return total;
}

/*
 * orc_exp_value with the per-string sum split into chunks of 2^20 dump
 * positions: each chunk is summed sequentially in ascending index order (the
 * terms of orc_exp_value, unchanged), the chunks in parallel, and the chunk
 * partials sequentially in ascending order.  Only the association of the
 * additions differs from the single running sum of from_dump_to_prob
 * (convert.rs:30-42); for n <= 20 there is one chunk and the result is the
 * same bits.  It lets large states (30 qubits) be checked in seconds.
 */
#define EXP_CHUNK_BITS 20u

double orc_exp_value_chunked(orc_v2 *s, const uint8_t *paulis, const uint32_t *qubits,
                             const uint32_t *offsets, const double *coeffs, uint32_t nstrings) {
    const orc_c64 *a = orc_v2_data(s);
    const uint32_t n = orc_v2_num_qubits(s);
    const uint64_t size = (uint64_t)1 << n;
    const uint64_t chunk = n > EXP_CHUNK_BITS ? ((uint64_t)1 << EXP_CHUNK_BITS) : size;
    const uint64_t nchunks = size / chunk;
    double *part = (double *)malloc(nchunks * sizeof(double));
    double total = 0.0;
    for (uint32_t k = 0; k < nstrings; k++) {
        const uint32_t lo = offsets[k], hi = offsets[k + 1];
        basis_change(s, paulis, qubits, lo, hi, 0);
#pragma omp parallel for schedule(static) if (nchunks > 1)
        for (uint64_t c = 0; c < nchunks; c++) {
            double r = 0.0;
            for (uint64_t i = c * chunk; i < (c + 1) * chunk; i++) {
                const double h = hypot(a[i].re, a[i].im);
                if (!(h > ORC_SMALL_EPSILON)) continue;
                const double prob = h * h;
                const uint64_t key = orc_dump_key(i, qubits + lo, hi - lo);
                const double parity = (__builtin_popcountll(key) % 2 == 0) ? 1.0 : -1.0;
                r += prob * parity;
            }
            part[c] = r;
        }
        double result = 0.0;
        for (uint64_t c = 0; c < nchunks; c++) result += part[c];
        basis_change(s, paulis, qubits, lo, hi, 1);
        total += result * coeffs[k];
    }
    free(part);
    return total;
}

/*
 * Default sample (quantum_execution.rs:99-102) through from_dump_to_prob and
 * from_prob_to_shots (convert.rs:30-73) with rand's WeightedIndex: weights in
 * dump order, cumulative_weights[j] = w_0 + ... + w_j for j < m-1 (one
 * sequential running sum), total = the full sum, index = #{j : cum_j <= x}.
 * Writes the dump key of each shot instead of the HashMap histogram, whose
 * iteration order is random.  Returns -1 when all weights are zero (the
 * reference panics there, convert.rs:60).
 */
int orc_sample_sequential(orc_v2 *s, const uint32_t *qubits, uint32_t nq, uint64_t shots,
                          const double *uniforms, uint64_t *outcomes) {
    const uint64_t m = orc_v2_dump(s, qubits, nq, NULL, NULL, NULL);
    if (m == 0) return -1;
    uint64_t *keys = (uint64_t *)malloc(m * sizeof(uint64_t));
    double *re = (double *)malloc(m * sizeof(double));
    double *im = (double *)malloc(m * sizeof(double));
    double *cum = (double *)malloc(m * sizeof(double));
    orc_v2_dump(s, qubits, nq, keys, re, im);
    double total = 0.0;
    for (uint64_t j = 0; j < m; j++) {
        const double h = hypot(re[j], im[j]);
        const double w = h * h;
        if (j == 0) {
            total = w;
        } else {
            cum[j - 1] = total;
            total += w;
        }
    }
    int rc = 0;
    if (total == 0.0) {
        rc = -1;
    } else {
        const double scale = orc_uniform_scale(total);
        for (uint64_t t = 0; t < shots; t++) {
            const double x = uniforms[t] * scale + 0.0;
            /* partition_point(|w| w <= x) over cum[0 .. m-1) */
            uint64_t lo = 0, hi = m - 1;
            while (lo < hi) {
                uint64_t mid = lo + (hi - lo) / 2;
                if (cum[mid] <= x)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            outcomes[t] = keys[lo];
        }
    }
    free(keys);
    free(re);
    free(im);
    free(cum);
    return rc;
}

/*
 * Sampling with a fixed hierarchical cumulative order, the order the GPU
 * reproduces bit for bit.  It replaces the single 2^n-long sequential
 * running sum of WeightedIndex (which no parallel machine can reproduce)
 * by nested sequential sums over groups of 64:
 *   w_i     = re*re + im*im, set to 0 when <= 1e-30 (|a| <= 1e-15, the dump
 *             filter of dense_v2.rs:313);
 *   T^0     = w;  T^(l+1)_j = (((0 + T^l_64j) + T^l_64j+1) + ... + T^l_64j+63)
 *             until at most 64 entries remain (the top level);
 *   total   = sequential sum of the top level;
 *   x       = u * scale(total)  (rand's UniformFloat, as in orc_measure);
 *   descent = at each level, starting from the parent's running base, scan
 *             the 64 children: pick the first child e with x < base + T_e
 *             (base += T_e otherwise), or else the last child of nonzero
 *             weight (with the running base before it).  The child level
 *             sums from the parent's base, the parent summed its children
 *             from 0, so rounding can leave x past every child's running sum
 *             although the parent was picked; the last child may then have
 *             weight 0, an entry the dump filter of the reference excludes.
 * The picked index is the physical amplitude index; the outcome is its dump
 * key.  The reference's rule (first entry whose running sum exceeds x, else
 * the last) is kept; only the association of the additions changes.
 */
#define HIER_FAN 64u

int orc_sample_hier(const orc_c64 *amp, uint32_t num_qubits, const uint32_t *qubits, uint32_t nq,
                    uint64_t shots, const double *uniforms, uint64_t *outcomes) {
    const uint64_t size = (uint64_t)1 << num_qubits;
    /* level 0: weights */
    double *levels[16];
    uint64_t counts[16];
    int nlev = 0;
    double *w = (double *)malloc(size * sizeof(double));
#pragma omp parallel for schedule(static) if (size > 65536)
    for (uint64_t i = 0; i < size; i++) {
        double v = amp[i].re * amp[i].re + amp[i].im * amp[i].im;
        w[i] = (v > 1e-30) ? v : 0.0;
    }
    levels[0] = w;
    counts[0] = size;
    nlev = 1;
    while (counts[nlev - 1] > HIER_FAN) {
        const uint64_t cnt = counts[nlev - 1] / HIER_FAN;
        double *prev = levels[nlev - 1];
        double *t = (double *)malloc(cnt * sizeof(double));
#pragma omp parallel for schedule(static) if (cnt > 4096)
        for (uint64_t j = 0; j < cnt; j++) {
            double acc = 0.0;
            for (uint32_t e = 0; e < HIER_FAN; e++) acc = acc + prev[j * HIER_FAN + e];
            t[j] = acc;
        }
        levels[nlev] = t;
        counts[nlev] = cnt;
        nlev++;
    }
    const int top = nlev - 1;
    double total = 0.0;
    for (uint64_t e = 0; e < counts[top]; e++) total = total + levels[top][e];
    int rc = 0;
    if (total == 0.0) {
        rc = -1;
    } else {
        const double scale = orc_uniform_scale(total);
#pragma omp parallel for schedule(static) if (shots > 1024)
        for (uint64_t t = 0; t < shots; t++) {
            const double x = uniforms[t] * scale + 0.0;
            double base = 0.0;
            uint64_t start = 0, cnt = counts[top];
            uint64_t pick = 0;
            for (int l = top; l >= 0; l--) {
                const double *v = levels[l];
                uint64_t e = 0, nz = cnt;
                double acc = base, acc_nz = base;
                int hit = 0;
                for (e = 0; e + 1 < cnt; e++) {
                    const double w = v[start + e];
                    const double nxt = acc + w;
                    if (x < nxt) {
                        hit = 1;
                        break;
                    }
                    if (w > 0.0) {
                        nz = e;
                        acc_nz = acc;
                    }
                    acc = nxt;
                }
                if (!hit && !(v[start + cnt - 1] > 0.0) && nz < cnt) {
                    e = nz;
                    acc = acc_nz;
                }
                base = acc;
                pick = start + e;
                start = pick * HIER_FAN;
                cnt = HIER_FAN;
            }
            outcomes[t] = orc_dump_key(pick, qubits, nq);
        }
    }
    for (int l = 0; l < nlev; l++) free(levels[l]);
    return rc;
}
