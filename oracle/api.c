/* api.c -- flat entry points of the oracle for the Python test harness.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Pointer-and-size functions so
 * that ctypes can drive the restatement without knowing struct layouts.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void *or_api_bench_new(int log_n, uint64_t seed) {
    or_p2_bench_t *b = (or_p2_bench_t *)malloc(sizeof *b);
    or_p2_bench_setup(b, log_n, seed);
    return b;
}
void or_api_bench_free(void *h) {
    or_p2_bench_free((or_p2_bench_t *)h);
    free(h);
}
const f_t *or_api_bench_trace(void *h) { return ((or_p2_bench_t *)h)->trace; }
const f_t *or_api_bench_inputs(void *h) { return ((or_p2_bench_t *)h)->inputs; }
/* AIR round constants: beginning[64], partial[20], ending[64] */
void or_api_bench_air_consts(void *h, f_t *out) { memcpy(out, &((or_p2_bench_t *)h)->air, sizeof(or_air_consts_t)); }
/* Poseidon2 constants: ext_init[4][W], ext_term[4][W], internal[rounds_p] (packed) */
int or_api_bench_p2_consts(void *h, int width, f_t *out) {
    or_p2_bench_t *b = (or_p2_bench_t *)h;
    const or_p2_t *p = width == 16 ? &b->p16 : &b->p24;
    int k = 0;
    for (int r = 0; r < 4; r++) for (int i = 0; i < width; i++) out[k++] = p->ext_init[r][i];
    for (int r = 0; r < 4; r++) for (int i = 0; i < width; i++) out[k++] = p->ext_term[r][i];
    for (int r = 0; r < p->rounds_p; r++) out[k++] = p->internal[r];
    return k;
}
void or_api_permute(void *h, int width, f_t *state) {
    or_p2_bench_t *b = (or_p2_bench_t *)h;
    or_p2_permute(width == 16 ? &b->p16 : &b->p24, state);
}
void or_api_hash_leaf(void *h, const f_t *row, size_t len, f_t *out) {
    or_hash_leaf(&((or_p2_bench_t *)h)->p24, row, len, out);
}
void or_api_compress(void *h, const f_t *l, const f_t *r, f_t *out) {
    or_compress(&((or_p2_bench_t *)h)->p16, l, r, out);
}

static void settings(or_air_settings_t *s, int sec, int f0, int f, int rate, int skips, int rs0, int sa) {
    s->soundness = sa;
    s->security_bits = sec;
    s->folding_f0 = f0;
    s->folding_f = f;
    s->log_inv_rate = rate;
    s->univariate_skips = skips;
    s->rs_reduction0 = rs0;
}

size_t or_api_prove(void *h, int sec, int f0, int f, int rate, int skips, int rs0, int sa, f_t **out) {
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    return or_air_prove(&s, (or_p2_bench_t *)h, out);
}
int or_api_verify(void *h, int sec, int f0, int f, int rate, int skips, int rs0, int sa, const f_t *proof,
                  size_t len) {
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    return or_air_verify(&s, (or_p2_bench_t *)h, proof, len);
}
void or_api_free(void *p) { free(p); }

/* WHIR schedule as a flat int array: [n_rounds, final_sumcheck_rounds,
 * commit_ood, starting_folding_pow, final_queries, final_pow, final_folding_pow,
 * final_log_domain, then per round 8 ints] */
int or_api_whir_config(int v, int sec, int max_pow, int f0, int f, int rate, int rs0, int sa, int *out) {
    or_whir_cfg_t c;
    if (or_whir_config(&c, v, sec, max_pow, f0, f, rate, rs0, sa)) return -1;
    int k = 0;
    out[k++] = c.n_rounds; out[k++] = c.final_sumcheck_rounds; out[k++] = c.commit_ood_samples;
    out[k++] = c.starting_folding_pow_bits; out[k++] = c.final_queries; out[k++] = c.final_pow_bits;
    out[k++] = c.final_folding_pow_bits; out[k++] = c.final_log_domain_size;
    for (int r = 0; r < c.n_rounds; r++) {
        const or_whir_round_t *p = &c.rounds[r];
        out[k++] = p->pow_bits; out[k++] = p->folding_pow_bits; out[k++] = p->num_queries;
        out[k++] = p->ood_samples; out[k++] = p->log_inv_rate; out[k++] = p->num_variables;
        out[k++] = p->folding_factor; out[k++] = p->log_domain_size;
    }
    return k;
}

/* ---- field helpers ---- */
void or_api_ef_mul(const f_t *a, const f_t *b, f_t *out) {
    ef_t x, y;
    memcpy(x.c, a, 32); memcpy(y.c, b, 32);
    ef_t r = ef_mul(x, y);
    memcpy(out, r.c, 32);
}
void or_api_ef_inv(const f_t *a, f_t *out) {
    ef_t x;
    memcpy(x.c, a, 32);
    ef_t r = ef_inv(x);
    memcpy(out, r.c, 32);
}
f_t or_api_two_adic_generator(int bits) { return f_two_adic_generator(bits); }
void or_api_rng_u32(uint64_t seed, uint32_t *out, size_t n) {
    or_rng_t r;
    or_rng_seed_u64(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = or_rng_next_u32(&r);
}
void or_api_rng_f(uint64_t seed, f_t *out, size_t n) {
    or_rng_t r;
    or_rng_seed_u64(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = or_rng_sample_f(&r);
}
void or_api_eval_eq(const f_t *point, int n, f_t *out) { or_eval_eq((const ef_t *)point, n, (ef_t *)out); }
void or_api_mle_eval_f(const f_t *evals, int n, const f_t *point, f_t *out) {
    ef_t r = or_mle_eval_f(evals, n, (const ef_t *)point);
    memcpy(out, r.c, 32);
}
void or_api_mle_eval_ef(const f_t *evals, int n, const f_t *point, f_t *out) {
    ef_t r = or_mle_eval_ef((const ef_t *)evals, n, (const ef_t *)point);
    memcpy(out, r.c, 32);
}
void or_api_air_eval_f(void *h, const f_t *local, const f_t *alpha_pows, f_t *out) {
    ef_t r = or_air_eval_f(&((or_p2_bench_t *)h)->air, local, (const ef_t *)alpha_pows);
    memcpy(out, r.c, 32);
}
void or_api_air_eval_ef(void *h, const f_t *local, const f_t *alpha_pows, f_t *out) {
    ef_t r = or_air_eval_ef(&((or_p2_bench_t *)h)->air, (const ef_t *)local, (const ef_t *)alpha_pows);
    memcpy(out, r.c, 32);
}
void or_api_selectors_at_ef(int k, const f_t *x, f_t *out) {
    ef_t xx;
    memcpy(xx.c, x, 32);
    or_selectors_at_ef(k, xx, (ef_t *)out);
}
void or_api_lagrange(const f_t *xs, const f_t *ys, int n, f_t *coeffs) {
    or_lagrange_interp(xs, (const ef_t *)ys, n, (ef_t *)coeffs);
}

/* ---- folds (crates/utils/src/multilinear.rs) ---- */
/* fold an F table (2^log_n) by 2^k EF scalars -> EF table (2^(log_n-k)) */
void or_api_fold_f_large(const f_t *in, int log_n, const f_t *scalars, int k, f_t *out) {
    size_t M = (size_t)1 << (log_n - k);
    int K = 1 << k;
    const ef_t *s = (const ef_t *)scalars;
    ef_t *o = (ef_t *)out;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < M; i++) {
        ef_t acc = ef_zero();
        for (int j = 0; j < K; j++) acc = ef_add(acc, ef_mul_f(s[j], in[i + (size_t)j * M]));
        o[i] = acc;
    }
}
void or_api_fold_ef_large(const f_t *in, int log_n, const f_t *scalars, int k, f_t *out) {
    size_t M = (size_t)1 << (log_n - k);
    int K = 1 << k;
    const ef_t *s = (const ef_t *)scalars, *x = (const ef_t *)in;
    ef_t *o = (ef_t *)out;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < M; i++) {
        ef_t acc = ef_zero();
        for (int j = 0; j < K; j++) acc = ef_add(acc, ef_mul(s[j], x[i + (size_t)j * M]));
        o[i] = acc;
    }
}

/* crates/utils/src/multilinear.rs:10-52 fold_multilinear_in_small_field:
 * out[i] = sum_j s_j m[i + j 2^(log_n - k)], F scalars; m and out in F (width 1)
 * or EF (width 8) */
void or_api_fold_small(const f_t *in, int width, int log_n, const f_t *scalars, int k, f_t *out) {
    size_t M = (size_t)1 << (log_n - k);
    int K = 1 << k;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < M; i++)
        for (int c = 0; c < width; c++) {
            f_t acc = 0;
            for (int j = 0; j < K; j++) acc = f_add(acc, f_mul(scalars[j], in[(i + (size_t)j * M) * width + c]));
            out[i * width + c] = acc;
        }
}
/* crates/utils/src/multilinear.rs:125-147 multilinears_linear_combination:
 * out[x] = sum_c scalars[c] col_c[x], F columns, EF scalars and output */
void or_api_lincomb(const f_t *cols, int n_cols, int log_n, const f_t *scalars, f_t *out) {
    size_t N = (size_t)1 << log_n;
    const ef_t *s = (const ef_t *)scalars;
    ef_t *o = (ef_t *)out;
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) {
        ef_t acc = ef_zero();
        for (int c = 0; c < n_cols; c++) acc = ef_add(acc, ef_mul_f(s[c], cols[(size_t)c * N + x]));
        o[x] = acc;
    }
}

/* ---- product sumcheck micro: sumcheck::prove(1, [a, b], a^2 b, 3, [1], Some(z), false, ...)
 * over a transcript whose challenger uses the Poseidon16 of bench handle h with an
 * empty domain separator.  Returns the proof length; proof copied to out. ---- */
size_t or_api_sc_a2b(void *h, const f_t *a, const f_t *b, int n, const f_t *z, f_t *challenges,
                     f_t *final_evals, f_t *final_sum, f_t **proof_out) {
    or_p2_bench_t *bb = (or_p2_bench_t *)h;
    or_ts_t ts;
    or_ts_init(&ts, &bb->p16, NULL, 0);
    /* claimed sum = sum_x eq(z, x) a(x)^2 b(x) */
    size_t N = (size_t)1 << n;
    ef_t *eq = (ef_t *)malloc(N * sizeof(ef_t));
    or_eval_eq((const ef_t *)z, n, eq);
    ef_t sum = ef_zero();
    for (size_t x = 0; x < N; x++) sum = ef_add(sum, ef_mul_f(eq[x], f_mul(f_mul(a[x], a[x]), b[x])));
    free(eq);
    const void *mles[2] = {a, b};
    or_sc_comp_t comp = {OR_SC_A2B, NULL, NULL, 0, NULL};
    ef_t fs;
    or_sc_prove(1, 2, 1, mles, n, &comp, 3, (const ef_t *)z, 0, &ts, sum, 128,
                (ef_t *)challenges, (ef_t *)final_evals, &fs);
    memcpy(final_sum, fs.c, 32);
    *proof_out = ts.data;
    return ts.len;
}

/* ---- transcript handles (whir-p3 ProverState over DuplexChallenger) ---- */
typedef struct {
    or_p2_t p16;
    or_ts_t ts;
} or_api_ts_t;

static void p16_from_packed(or_p2_t *p, const f_t *c) {
    memset(p, 0, sizeof *p);
    p->width = 16;
    p->rounds_f = 8;
    p->rounds_p = 20;
    int k = 0;
    for (int r = 0; r < 4; r++) for (int i = 0; i < 16; i++) p->ext_init[r][i] = c[k++];
    for (int r = 0; r < 4; r++) for (int i = 0; i < 16; i++) p->ext_term[r][i] = c[k++];
    for (int r = 0; r < 20; r++) p->internal[r] = c[k++];
}
void *or_api_ts_new(const f_t *p16c, const f_t *domsep, size_t n) {
    or_api_ts_t *h = (or_api_ts_t *)malloc(sizeof *h);
    p16_from_packed(&h->p16, p16c);
    or_ts_init(&h->ts, &h->p16, domsep, n);
    return h;
}
void or_api_ts_free(void *h) {
    or_ts_free(&((or_api_ts_t *)h)->ts);
    free(h);
}
void or_api_ts_observe(void *h, const f_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) or_chal_observe(&((or_api_ts_t *)h)->ts.ch, v[i]);
}
void or_api_ts_add_base(void *h, const f_t *v, size_t n) { or_ts_add_base(&((or_api_ts_t *)h)->ts, v, n); }
void or_api_ts_hint(void *h, const f_t *v, size_t n) { or_ts_hint_base(&((or_api_ts_t *)h)->ts, v, n); }
void or_api_ts_sample_ef(void *h, f_t *out, size_t n_ef) {
    for (size_t i = 0; i < n_ef; i++) {
        ef_t e = or_ts_sample_ef(&((or_api_ts_t *)h)->ts);
        memcpy(out + 8 * i, e.c, 32);
    }
}
uint32_t or_api_ts_sample_bits(void *h, int bits) { return or_chal_sample_bits(&((or_api_ts_t *)h)->ts.ch, bits); }
void or_api_ts_pow(void *h, int bits) { or_ts_pow(&((or_api_ts_t *)h)->ts, bits); }
size_t or_api_ts_len(void *h) { return ((or_api_ts_t *)h)->ts.len; }
const f_t *or_api_ts_data(void *h) { return ((or_api_ts_t *)h)->ts.data; }
/* AirTable::prove appended to the handle's transcript (bench h for the inputs) */
void or_api_prove_ts(void *bench, int sec, int f0, int f, int rate, int skips, int rs0, int sa, void *h) {
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    or_air_prove_ts(&s, (or_p2_bench_t *)bench, &((or_api_ts_t *)h)->ts);
}
/* the domain-separator words of the default transcript */
size_t or_api_domsep(int log_n, int sec, int f0, int f, int rate, int skips, int rs0, int sa, f_t *out,
                     size_t cap) {
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    or_whir_cfg_t cfg;
    if (or_air_whir_config(&cfg, &s, log_n)) return 0;
    return or_whir_domsep(&cfg, out, cap);
}

/* ---- generic sumcheck (crates/sumcheck/src/prove.rs:22-93) with a constraint
 * program; mles contiguous (n_mles x 2^n_vars, F when base else EF) ---- */
void or_api_sc_prove(void *h, int skips, const f_t *mles, int n_mles, int base, int n_vars, const uint32_t *code,
                     int n_instr, int degree, const f_t *batching, const f_t *eq_factor, int is_zerofier,
                     const f_t *sum, int n_rounds, int grind_kind, int grind_value, const f_t *missing,
                     f_t *challenges, f_t *folded, f_t *final_sum) {
    size_t N = (size_t)1 << n_vars, w = base ? 1 : 8;
    const void **ptrs = (const void **)malloc((size_t)n_mles * sizeof(void *));
    for (int m = 0; m < n_mles; m++) ptrs[m] = mles + (size_t)m * N * w;
    or_prog_t prog = {n_mles, n_instr, code};
    or_sc_comp_t comp = {OR_SC_PROGRAM, NULL, (const ef_t *)batching, 0, &prog};
    ef_t s, fs;
    memcpy(s.c, sum, 32);
    ef_t mm;
    if (missing) memcpy(mm.c, missing, 32);
    or_sc_prove_ex(skips, n_mles, base, ptrs, n_vars, &comp, degree, (const ef_t *)eq_factor, is_zerofier,
                   &((or_api_ts_t *)h)->ts, s, n_rounds, grind_kind, grind_value, missing ? &mm : NULL,
                   (ef_t *)challenges, (ef_t *)folded, &fs);
    memcpy(final_sum, fs.c, 32);
    free(ptrs);
}
int or_api_prog_check(const uint32_t *code, int n_instr, int n_inputs) {
    or_prog_t p = {n_inputs, n_instr, code};
    return or_prog_check(&p);
}
void or_api_prog_eval(const uint32_t *code, int n_instr, int n_inputs, int base, const f_t *pt,
                      const f_t *batching, f_t *out) {
    or_prog_t p = {n_inputs, n_instr, code};
    ef_t r = base ? or_prog_eval_f(&p, pt, (const ef_t *)batching)
                  : or_prog_eval_ef(&p, (const ef_t *)pt, (const ef_t *)batching);
    memcpy(out, r.c, 32);
}

/* ---- generic AIR (crates/air/src/table.rs:32-56, prove.rs, verify.rs) ---- */
static void p24_from_packed(or_p2_t *p, const f_t *c) {
    memset(p, 0, sizeof *p);
    p->width = 24;
    p->rounds_f = 8;
    p->rounds_p = 23;
    int k = 0;
    for (int r = 0; r < 4; r++) for (int i = 0; i < 24; i++) p->ext_init[r][i] = c[k++];
    for (int r = 0; r < 4; r++) for (int i = 0; i < 24; i++) p->ext_term[r][i] = c[k++];
    for (int r = 0; r < 23; r++) p->internal[r] = c[k++];
}
static void gair(or_gair_t *a, int C, int np, int log_n, const f_t *pre, const uint32_t *code, int n_instr,
                 int degree) {
    a->n_columns = C;
    a->n_preprocessed = np;
    a->log_n = log_n;
    a->preprocessed = pre;
    a->prog.n_inputs = 2 * C;
    a->prog.n_instr = n_instr;
    a->prog.code = code;
    a->degree = degree;
}
/* appends the proof to transcript handle h (whose Poseidon2-16 is p16c) */
void or_api_gair_prove(int C, int np, int log_n, const f_t *pre, const uint32_t *code, int n_instr, int degree,
                       int sec, int f0, int f, int rate, int skips, int rs0, int sa, const f_t *witness,
                       const f_t *p16c, const f_t *p24c, void *h) {
    or_gair_t a;
    gair(&a, C, np, log_n, pre, code, n_instr, degree);
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    or_p2_t p16, p24;
    p16_from_packed(&p16, p16c);
    p24_from_packed(&p24, p24c);
    or_gair_prove(&s, &a, witness, &p16, &p24, &((or_api_ts_t *)h)->ts);
}
/* verifies proof against the transcript that first observes domsep; 0 on accept */
int or_api_gair_verify(int C, int np, int log_n, const f_t *pre, const uint32_t *code, int n_instr, int degree,
                       int sec, int f0, int f, int rate, int skips, int rs0, int sa, const f_t *p16c,
                       const f_t *p24c, const f_t *domsep, size_t dn, const f_t *proof, size_t len) {
    or_gair_t a;
    gair(&a, C, np, log_n, pre, code, n_instr, degree);
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    or_p2_t p16, p24;
    p16_from_packed(&p16, p16c);
    p24_from_packed(&p24, p24c);
    or_ts_t ts;
    or_ts_init_verifier(&ts, &p16, domsep, dn, proof, len);
    int rc = or_gair_verify(&s, &a, &p16, &p24, &ts);
    if (!rc && ts.pos != ts.len) rc = 7; /* trailing words */
    or_ts_free(&ts);
    return rc;
}
size_t or_api_gair_domsep(int nw, int log_n, int sec, int f0, int f, int rate, int skips, int rs0, int sa,
                          f_t *out, size_t cap) {
    or_air_settings_t s;
    settings(&s, sec, f0, f, rate, skips, rs0, sa);
    or_whir_cfg_t cfg;
    int v = log_n + (int)or_log2_up((size_t)nw);
    if (or_whir_config(&cfg, v, sec, 16, f0, f, rate, rs0, sa)) return 0;
    return or_whir_domsep(&cfg, out, cap);
}
