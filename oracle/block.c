/*
 * Oracle restatement of Block<(), DenseV2<f64>> (src/block.rs) with the
 * out-of-place global/local swap of DenseV2's BlockSimulator impl
 * (dense_v2.rs:344-396).  TEST INFRASTRUCTURE.
 *
 * The register is split into L local qubits (one DenseV2 per block) and
 * G = n - L global qubits (the block index).  Logical qubits map to physical
 * ones through a bijection (the reference's BiMap); a gate whose target is
 * global first flushes the gate queue and swaps that qubit with the least
 * recently used local one (block.rs:123-156).  Global controls become a
 * per-block execute flag (block.rs:158-177).
 */
#include <stdlib.h>
#include <string.h>

#include "kbw_oracle.h"
#include "oracle_internal.h"

typedef struct {
    int kind;
    double angle;
    uint32_t target;
    uint64_t ctrl_mask; /* local physical controls */
    uint8_t *execute;   /* one flag per block */
} queued_gate;

struct orc_block {
    uint32_t nglobal, nlocal, block_size_arg;
    uint64_t nblocks;
    orc_v2 **sims;
    uint32_t *l2p, *p2l;
    uint64_t *last_access;
    uint64_t counter;
    queued_gate *queue;
    uint64_t qlen, qcap;
    uint64_t swaps;
};

/* dense_v2.rs:346-358: block 0 holds |0>, the others are zero */
orc_block *orc_block_new(uint32_t n, uint32_t block_size) {
    if (n <= 1) return NULL; /* block.rs:218-220 */
    uint32_t bs = block_size == 0 ? 20 : block_size; /* KBW_BLOCK_SIZE default, block.rs:222-228 */
    if (bs > n - 1) bs = n - 1;
    if (bs > 32) return NULL; /* DenseV2::new(num_local) rejects > 32 */
    orc_block *b = (orc_block *)calloc(1, sizeof *b);
    b->block_size_arg = block_size;
    b->nlocal = bs;
    b->nglobal = n - bs;
    b->nblocks = (uint64_t)1 << b->nglobal;
    b->sims = (orc_v2 **)calloc(b->nblocks, sizeof(orc_v2 *));
    for (uint64_t g = 0; g < b->nblocks; g++) {
        b->sims[g] = orc_v2_new(bs);
        if (g != 0) orc_v2_data(b->sims[g])[0].re = 0.0;
    }
    b->l2p = (uint32_t *)malloc(n * sizeof(uint32_t));
    b->p2l = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) b->l2p[i] = b->p2l[i] = i;
    b->last_access = (uint64_t *)calloc(bs, sizeof(uint64_t));
    return b;
}

static void clear_queue(orc_block *b) {
    for (uint64_t i = 0; i < b->qlen; i++) free(b->queue[i].execute);
    b->qlen = 0;
}

void orc_block_free(orc_block *b) {
    if (!b) return;
    clear_queue(b);
    free(b->queue);
    for (uint64_t g = 0; g < b->nblocks; g++) orc_v2_free(b->sims[g]);
    free(b->sims);
    free(b->l2p);
    free(b->p2l);
    free(b->last_access);
    free(b);
}

/* block.rs:104-109 */
static void touch(orc_block *b, uint32_t phys) {
    if (phys < b->nlocal) {
        b->counter += 1;
        b->last_access[phys] = b->counter;
    }
}

/* block.rs:111-117: the first minimum timestamp (lowest index on ties) */
static uint32_t next_local_qubit(const orc_block *b) {
    uint32_t best = 0;
    for (uint32_t i = 1; i < b->nlocal; i++)
        if (b->last_access[i] < b->last_access[best]) best = i;
    return best;
}

/* block.rs:179-208: every block runs the whole queue */
static void execute_gates(orc_block *b) {
    for (uint64_t g = 0; g < b->nblocks; g++)
        for (uint64_t k = 0; k < b->qlen; k++) {
            const queued_gate *q = &b->queue[k];
            if (q->execute[g]) orc_v2_gate_mask(b->sims[g], q->kind, q->angle, q->target, q->ctrl_mask);
        }
    clear_queue(b);
}

/* dense_v2.rs:360-396: new[G][L] = old[G with bit g := l-bit of L][L with bit l := g-bit of G] */
static void swap_blocks(orc_block *b, uint32_t gq, uint32_t lq) {
    const uint64_t lsize = (uint64_t)1 << b->nlocal;
    orc_c64 **fresh = (orc_c64 **)malloc(b->nblocks * sizeof(orc_c64 *));
    for (uint64_t G = 0; G < b->nblocks; G++) {
        fresh[G] = (orc_c64 *)malloc(lsize * sizeof(orc_c64));
        for (uint64_t L = 0; L < lsize; L++) {
            const uint64_t g_bit = (G >> gq) & 1, l_bit = (L >> lq) & 1;
            const uint64_t src_g = (G & ~((uint64_t)1 << gq)) | (l_bit << gq);
            const uint64_t src_l = (L & ~((uint64_t)1 << lq)) | (g_bit << lq);
            fresh[G][L] = orc_v2_data(b->sims[src_g])[src_l];
        }
    }
    for (uint64_t G = 0; G < b->nblocks; G++) {
        memcpy(orc_v2_data(b->sims[G]), fresh[G], lsize * sizeof(orc_c64));
        free(fresh[G]);
    }
    free(fresh);
    b->swaps++;
}

/* block.rs:123-156 */
static uint32_t get_local_qubit(orc_block *b, uint32_t logical) {
    const uint32_t phys = b->l2p[logical];
    if (phys < b->nlocal) {
        touch(b, phys);
        return phys;
    }
    execute_gates(b);
    const uint32_t gq = phys - b->nlocal;
    const uint32_t lq = next_local_qubit(b);
    swap_blocks(b, gq, lq);
    const uint32_t old_logical = b->p2l[lq];
    b->l2p[logical] = lq;
    b->p2l[lq] = logical;
    b->l2p[old_logical] = phys;
    b->p2l[phys] = old_logical;
    touch(b, lq);
    return lq;
}

void orc_block_gate(orc_block *b, int kind, double angle, uint32_t target, const uint32_t *control,
                    uint32_t nc) {
    const uint32_t t = get_local_qubit(b, target);
    /* block.rs:158-177 */
    queued_gate q;
    q.kind = kind;
    q.angle = angle;
    q.target = t;
    q.ctrl_mask = 0;
    q.execute = (uint8_t *)malloc(b->nblocks);
    memset(q.execute, 1, b->nblocks);
    for (uint32_t k = 0; k < nc; k++) {
        const uint32_t c = b->l2p[control[k]];
        if (c < b->nlocal) {
            q.ctrl_mask |= (uint64_t)1 << c;
        } else {
            const uint64_t gm = (uint64_t)1 << (c - b->nlocal);
            for (uint64_t g = 0; g < b->nblocks; g++) q.execute[g] &= (g & gm) == gm;
        }
    }
    if (b->qlen == b->qcap) {
        b->qcap = b->qcap ? 2 * b->qcap : 64;
        b->queue = (queued_gate *)realloc(b->queue, b->qcap * sizeof(queued_gate));
    }
    b->queue[b->qlen++] = q;
}

/* block.rs:302-309 */
double orc_block_measure_p1(orc_block *b, uint32_t target) {
    execute_gates(b);
    const uint32_t t = get_local_qubit(b, target);
    double sum = 0.0;
    for (uint64_t g = 0; g < b->nblocks; g++) sum += orc_v2_measure_p1(b->sims[g], t);
    return sum;
}

/* block.rs:311-316 (no flush of the queue, as in the reference) */
void orc_block_measure_collapse(orc_block *b, uint32_t target, int result, double p) {
    const uint32_t t = get_local_qubit(b, target);
    for (uint64_t g = 0; g < b->nblocks; g++) orc_v2_measure_collapse(b->sims[g], t, result, p);
}

/* block.rs:318-376 (release build: block-major physical order, no sort) */
uint64_t orc_block_dump(orc_block *b, const uint32_t *qubits, uint32_t nq, uint64_t *keys,
                        double *re, double *im) {
    execute_gates(b);
    uint32_t *phys = (uint32_t *)malloc((nq ? nq : 1) * sizeof(uint32_t));
    for (uint32_t k = 0; k < nq; k++) phys[k] = b->l2p[qubits[k]];
    const uint64_t lsize = (uint64_t)1 << b->nlocal;
    uint64_t m = 0;
    for (uint64_t g = 0; g < b->nblocks; g++) {
        const orc_c64 *a = orc_v2_data(b->sims[g]);
        for (uint64_t L = 0; L < lsize; L++) {
            if (!(hypot(a[L].re, a[L].im) > ORC_SMALL_EPSILON)) continue;
            if (keys) {
                keys[m] = orc_dump_key((g << b->nlocal) | L, phys, nq);
                re[m] = a[L].re;
                im[m] = a[L].im;
            }
            m++;
        }
    }
    free(phys);
    return m;
}

void orc_block_physical_state(orc_block *b, orc_c64 *out) {
    execute_gates(b);
    const uint64_t lsize = (uint64_t)1 << b->nlocal;
    for (uint64_t g = 0; g < b->nblocks; g++)
        memcpy(out + g * lsize, orc_v2_data(b->sims[g]), lsize * sizeof(orc_c64));
}

void orc_block_logical_state(orc_block *b, orc_c64 *out) {
    execute_gates(b);
    const uint32_t n = b->nglobal + b->nlocal;
    const uint64_t lsize = (uint64_t)1 << b->nlocal;
    for (uint64_t g = 0; g < b->nblocks; g++) {
        const orc_c64 *a = orc_v2_data(b->sims[g]);
        for (uint64_t L = 0; L < lsize; L++) {
            const uint64_t P = (g << b->nlocal) | L;
            uint64_t li = 0;
            for (uint32_t q = 0; q < n; q++) li |= ((P >> b->l2p[q]) & 1u) << q;
            out[li] = a[L];
        }
    }
}

void orc_block_map(const orc_block *b, uint32_t *l2p) {
    memcpy(l2p, b->l2p, (b->nglobal + b->nlocal) * sizeof(uint32_t));
}

uint64_t orc_block_swap_count(const orc_block *b) { return b->swaps; }

/* block.rs:378-380: a fresh Block of the same size */
void orc_block_clear(orc_block *b) {
    const uint32_t n = b->nglobal + b->nlocal;
    orc_block *f = orc_block_new(n, b->block_size_arg);
    orc_block tmp = *b;
    *b = *f;
    *f = tmp;
    orc_block_free(f);
}
