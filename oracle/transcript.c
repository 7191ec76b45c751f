/* transcript.c -- duplex challenger and Fiat-Shamir prover/verifier state (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * DuplexChallenger<F, Poseidon16, 16, 8> (src/examples/poseidon2.rs:27),
 * restated from p3-challenger @ d0c4a36 (parity unpinned):
 *   observe(v): clear output; push v to input; duplex when input holds 8
 *   sample():   duplex if input non-empty or output empty; pop output's last
 *   duplex:     overwrite state[0..k] with input, permute, output = state[0..8]
 *   sample_bits(b) = canonical(sample()) & (2^b - 1)
 *   grind(b): a witness w with observe(w); sample_bits(b) == 0.  The minimal
 *   such w is taken (the reference's rayon find_any equals it single-threaded).
 * ProverState (whir-p3 @ 9caef91, parity unpinned): add_* appends to the
 * proof and observes; hint_* appends without observing; pow_grinding(0) is a
 * no-op, otherwise the witness is appended (and observed by the check).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void or_chal_init(or_chal_t *c, const or_p2_t *p16) {
    memset(c, 0, sizeof *c);
    c->perm = p16;
}

static void duplex(or_chal_t *c) {
    for (int i = 0; i < c->n_in; i++) c->state[i] = c->in_buf[i];
    c->n_in = 0;
    or_p2_permute(c->perm, c->state);
    memcpy(c->out_buf, c->state, 8 * sizeof(f_t));
    c->n_out = 8;
}

void or_chal_observe(or_chal_t *c, f_t v) {
    c->n_out = 0;
    c->in_buf[c->n_in++] = v;
    if (c->n_in == 8) duplex(c);
}

f_t or_chal_sample(or_chal_t *c) {
    if (c->n_in > 0 || c->n_out == 0) duplex(c);
    return c->out_buf[--c->n_out];
}

uint32_t or_chal_sample_bits(or_chal_t *c, int bits) {
    f_t v = or_chal_sample(c);
    return v & (uint32_t)((1ull << bits) - 1);
}

ef_t or_chal_sample_ef(or_chal_t *c) {
    ef_t r;
    for (int i = 0; i < 8; i++) r.c[i] = or_chal_sample(c);
    return r;
}

/* The smallest witness, searched in rounds of consecutive candidates (64, then
 * four times more per round up to 2^16) whose members are tried in parallel: the
 * first round with a hit holds the smallest one. */
f_t or_grind_witness(const or_chal_t *c, int bits) {
    uint64_t base = 0, span = 64;
    while (base < OR_P) {
        uint64_t end = base + span < OR_P ? base + span : OR_P;
        f_t best = OR_P;
#pragma omp parallel for schedule(static) reduction(min : best) if (span > 64)
        for (uint64_t w = base; w < end; w++) {
            or_chal_t t = *c;
            or_chal_observe(&t, (f_t)w);
            if (or_chal_sample_bits(&t, bits) == 0 && (f_t)w < best) best = (f_t)w;
        }
        if (best < OR_P) return best;
        base = end;
        if (span < 65536) span *= 4;
    }
    abort();
}

static void push(or_ts_t *t, f_t v) {
    if (t->len == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->data = (f_t *)realloc(t->data, t->cap * sizeof(f_t));
    }
    t->data[t->len++] = v;
}

void or_ts_init(or_ts_t *t, const or_p2_t *p16, const f_t *domsep, size_t n) {
    memset(t, 0, sizeof *t);
    or_chal_init(&t->ch, p16);
    for (size_t i = 0; i < n; i++) or_chal_observe(&t->ch, domsep[i]);
}

void or_ts_init_verifier(or_ts_t *t, const or_p2_t *p16, const f_t *domsep, size_t n,
                         const f_t *proof, size_t len) {
    or_ts_init(t, p16, domsep, n);
    t->data = (f_t *)malloc((len ? len : 1) * sizeof(f_t));
    memcpy(t->data, proof, len * sizeof(f_t));
    t->len = t->cap = len;
}

void or_ts_free(or_ts_t *t) { free(t->data); t->data = NULL; }

void or_ts_add_base(or_ts_t *t, const f_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) { push(t, v[i]); or_chal_observe(&t->ch, v[i]); }
}
void or_ts_add_ef(or_ts_t *t, const ef_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) or_ts_add_base(t, v[i].c, 8);
}
void or_ts_hint_base(or_ts_t *t, const f_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) push(t, v[i]);
}
ef_t or_ts_sample_ef(or_ts_t *t) { return or_chal_sample_ef(&t->ch); }

void or_ts_pow(or_ts_t *t, int bits) {
    if (bits <= 0) return;
    f_t w = or_grind_witness(&t->ch, bits);
    or_chal_observe(&t->ch, w);
    (void)or_chal_sample_bits(&t->ch, bits);
    push(t, w);
}

int or_ts_next_base(or_ts_t *t, f_t *v, size_t n) {
    if (t->pos + n > t->len) { t->failed = 1; return -1; }
    for (size_t i = 0; i < n; i++) {
        v[i] = t->data[t->pos++];
        if (v[i] >= OR_P) { t->failed = 1; return -1; }
        or_chal_observe(&t->ch, v[i]);
    }
    return 0;
}
int or_ts_next_ef(or_ts_t *t, ef_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) if (or_ts_next_base(t, v[i].c, 8)) return -1;
    return 0;
}
int or_ts_next_hint_base(or_ts_t *t, f_t *v, size_t n) {
    if (t->pos + n > t->len) { t->failed = 1; return -1; }
    for (size_t i = 0; i < n; i++) {
        v[i] = t->data[t->pos++];
        if (v[i] >= OR_P) { t->failed = 1; return -1; }
    }
    return 0;
}
int or_ts_check_pow(or_ts_t *t, int bits) {
    if (bits <= 0) return 0;
    if (t->pos + 1 > t->len) { t->failed = 1; return -1; }
    f_t w = t->data[t->pos++];
    if (w >= OR_P) return -1;
    or_chal_observe(&t->ch, w);
    return or_chal_sample_bits(&t->ch, bits) == 0 ? 0 : -1;
}
