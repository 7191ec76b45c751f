/* air.c -- AirTable prove/verify for the Poseidon2 AIR and the benchmark setup (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates crates/air/src/prove.rs:49-226, verify.rs:55-249, utils.rs:12-212,
 * uni_skip_utils.rs:5-34, table.rs:32-98 and src/examples/poseidon2.rs:85-191
 * (first-party code: step order, transcript order and the algebra are the
 * reference's).  Pinned by: verifier acceptance of honest proofs, rejection
 * of tampered proofs, matrix_*_folded against matrix_*_lde, sub_evals check.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NC OR_AIR_NCOLS
#define NCONS OR_AIR_NCONSTRAINTS
#define WHIR_POW_BITS 16 /* crates/air/src/lib.rs:9 */

static int sat_sub(int a, int b) { return a > b ? a - b : 0; }
/* constraints_batching_pow / zerocheck_pow / secondary_sumchecks_batching_pow
 * (crates/air/src/utils.rs:162-212), EF::bits() = 248 */
static int pow_bits_for(int sec, int x) { return sat_sub(sec, sat_sub(248, x)); }

/* src/examples/poseidon2.rs:104-151 */
void or_p2_bench_setup(or_p2_bench_t *b, int log_n, uint64_t seed) {
    memset(b, 0, sizeof *b);
    b->log_n = log_n;
    or_rng_t rng;
    or_rng_seed_u64(&rng, seed);
    or_air_consts_from_rng(&b->air, &rng);
    size_t N = (size_t)1 << log_n;
    b->inputs = (f_t *)malloc(N * 16 * sizeof(f_t));
    for (size_t i = 0; i < N * 16; i++) b->inputs[i] = or_rng_sample_f(&rng);
    b->trace = (f_t *)malloc(N * NC * sizeof(f_t));
    or_air_generate_trace(&b->air, b->inputs, log_n, b->trace);
    or_p2_init_from_rng(&b->p16, 16, &rng);
    or_p2_init_from_rng(&b->p24, 24, &rng);
}

void or_p2_bench_free(or_p2_bench_t *b) {
    free(b->inputs);
    free(b->trace);
    b->inputs = b->trace = NULL;
}

/* AirTable::build_whir_params (crates/air/src/table.rs:72-98) */
int or_air_whir_config(or_whir_cfg_t *c, const or_air_settings_t *s, int log_n) {
    int v = log_n + (int)or_log2_up(NC);
    return or_whir_config(c, v, s->security_bits, WHIR_POW_BITS, s->folding_f0, s->folding_f,
                          s->log_inv_rate, s->rs_reduction0, s->soundness);
}

/* crates/air/src/uni_skip_utils.rs:5-13 */
static void matrix_up_folded(const ef_t *p, int n, ef_t *out) {
    size_t N = (size_t)1 << n;
    or_eval_eq(p, n, out);
    ef_t prod = ef_one();
    for (int i = 0; i < n; i++) prod = ef_mul(prod, p[i]);
    out[N - 1] = ef_sub(out[N - 1], prod);
    out[N - 2] = ef_add(out[N - 2], prod);
}

/* crates/air/src/uni_skip_utils.rs:15-34 */
static void matrix_down_folded(const ef_t *p, int n, ef_t *out) {
    size_t N = (size_t)1 << n;
    for (size_t i = 0; i < N; i++) out[i] = ef_zero();
    ef_t *eq = (ef_t *)malloc(N * sizeof(ef_t));
    for (int k = 0; k < n; k++) {
        ef_t prod = ef_sub(ef_one(), p[n - k - 1]);
        for (int i = n - k; i < n; i++) prod = ef_mul(prod, p[i]);
        or_eval_eq_scaled(p, n - k - 1, prod, eq);
        size_t M = (size_t)1 << (n - k - 1);
#pragma omp parallel for schedule(static) if (M >= 4096)
        for (size_t i = 0; i < M; i++) {
            size_t idx = (i << (k + 1)) + ((size_t)1 << k);
            out[idx] = ef_add(out[idx], eq[i]);
        }
    }
    ef_t prod = ef_one();
    for (int i = 0; i < n; i++) prod = ef_mul(prod, p[i]);
    out[N - 1] = ef_add(out[N - 1], prod);
    free(eq);
}

/* crates/air/src/utils.rs:12-37 */
static ef_t matrix_up_lde(const ef_t *p, int len) {
    int n = len / 2;
    ef_t r = or_eq_poly_outside(p, p + n, n);
    ef_t prod = ef_one();
    for (int i = 0; i < len - 1; i++) prod = ef_mul(prod, p[i]);
    ef_t t = ef_sub(ef_one(), ef_mul_f(p[len - 1], 2));
    return ef_add(r, ef_mul(prod, t));
}

/* crates/air/src/utils.rs:92-138 */
static ef_t next_mle(const ef_t *p, int len) {
    int n = len / 2;
    const ef_t *x = p, *y = p + n;
    ef_t one = ef_one(), total = ef_zero();
    for (int k = 0; k < n; k++) {
        ef_t eqh = one;
        for (int i = k + 1; i < n; i++) {
            ef_t a = x[n - 1 - i], b = y[n - 1 - i];
            eqh = ef_mul(eqh, ef_add(ef_mul(a, b), ef_mul(ef_sub(one, a), ef_sub(one, b))));
        }
        ef_t carry = ef_mul(ef_sub(one, x[n - 1 - k]), y[n - 1 - k]);
        ef_t low = one;
        for (int i = 0; i < k; i++) low = ef_mul(low, ef_mul(x[n - 1 - i], ef_sub(one, y[n - 1 - i])));
        total = ef_add(total, ef_mul(ef_mul(eqh, carry), low));
    }
    return total;
}

/* crates/air/src/utils.rs:39-66 */
static ef_t matrix_down_lde(const ef_t *p, int len) {
    ef_t prod = ef_one();
    for (int i = 0; i < len; i++) prod = ef_mul(prod, p[i]);
    return ef_add(next_mle(p, len), prod);
}

/* crates/air/src/prove.rs:49-226 with the default transcript (the domain
 * separator of src/examples/poseidon2.rs:155-163); returns the proof length */
size_t or_air_prove(const or_air_settings_t *s, const or_p2_bench_t *b, f_t **proof_out) {
    or_whir_cfg_t cfg;
    or_air_whir_config(&cfg, s, b->log_n);
    f_t domsep[4096];
    size_t dn = or_whir_domsep(&cfg, domsep, 4096);
    or_ts_t ts;
    or_ts_init(&ts, &b->p16, domsep, dn);
    or_air_prove_ts(s, b, &ts);
    *proof_out = ts.data;
    return ts.len;
}

/* crates/air/src/prove.rs:49-226: appends the proof to a caller's transcript */
void or_air_prove_ts(const or_air_settings_t *s, const or_p2_bench_t *b, or_ts_t *ts) {
    int n = b->log_n, skips = s->univariate_skips, sec = s->security_bits;
    size_t N = (size_t)1 << n;
    int lc = (int)or_log2_up(NC);
    or_whir_cfg_t cfg;
    or_air_whir_config(&cfg, s, n);

    /* 1) commit: packed_multilinear (crates/utils/src/multilinear.rs:169-181) */
    int v = n + lc;
    f_t *packed = (f_t *)calloc((size_t)1 << v, sizeof(f_t));
    memcpy(packed, b->trace, N * NC * sizeof(f_t));
    or_whir_witness_t w;
    or_whir_commit(&w, &cfg, &b->p16, &b->p24, ts, packed);
    free(packed);

    or_ts_pow(ts, pow_bits_for(sec, (int)or_log2_up(NCONS)));
    ef_t alpha = or_ts_sample_ef(ts);
    ef_t ap[NCONS];
    ap[0] = ef_one();
    for (int i = 1; i < NCONS; i++) ap[i] = ef_mul(ap[i - 1], alpha);
    or_ts_pow(ts, pow_bits_for(sec, n));
    int nz = n + 1 - skips;
    ef_t *zc = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    for (int i = 0; i < nz; i++) zc[i] = or_ts_sample_ef(ts);

    /* columns_up_and_down (crates/air/src/utils.rs:140-160) */
    f_t *ud = (f_t *)malloc(2 * NC * N * sizeof(f_t));
    const void **mles = (const void **)malloc(2 * NC * sizeof(void *));
    for (int c = 0; c < NC; c++) {
        const f_t *col = b->trace + (size_t)c * N;
        f_t *up = ud + (size_t)c * N, *dn2 = ud + (size_t)(NC + c) * N;
        memcpy(up, col, N * sizeof(f_t));
        up[N - 1] = col[N - 2];
        memcpy(dn2, col + 1, (N - 1) * sizeof(f_t));
        dn2[N - 1] = col[N - 1];
        mles[c] = up;
        mles[NC + c] = dn2;
    }
    or_sc_comp_t comp = {OR_SC_AIR, &b->air, ap, NC, NULL};
    ef_t *ch = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    ef_t fe[2 * NC], fsum;
    or_sc_prove(skips, 2 * NC, 1, mles, n, &comp, 3, zc, 1, ts, ef_zero(), sec, ch, fe, &fsum);
    free(ud);
    free(mles);

    or_ts_add_ef(ts, fe, NC);
    or_ts_add_ef(ts, fe + NC, NC);
    or_ts_pow(ts, pow_bits_for(sec, lc));
    ef_t cbs[16];
    for (int i = 0; i < lc; i++) cbs[i] = or_ts_sample_ef(ts);
    ef_t eqc[256];
    or_eval_eq(cbs, lc, eqc);
    /* multilinears_linear_combination (crates/utils/src/multilinear.rs:125-147) */
    ef_t *batched = (ef_t *)malloc(N * sizeof(ef_t));
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) {
        ef_t acc = ef_zero();
        for (int c = 0; c < NC; c++) acc = ef_add(acc, ef_mul_f(eqc[c], b->trace[(size_t)c * N + x]));
        batched[x] = acc;
    }
    ef_t alpha2 = or_ts_sample_ef(ts);
    ef_t *mixed = (ef_t *)malloc(N * sizeof(ef_t));
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) {
        ef_t up = x == N - 1 ? batched[N - 2] : batched[x];
        ef_t dn2 = x == N - 1 ? batched[N - 1] : batched[x + 1];
        mixed[x] = ef_add(up, ef_mul(dn2, alpha2));
    }
    /* fold(zerocheck_challenges[1..]): bind the last n - skips variables */
    int K = 1 << skips;
    size_t M = N / (size_t)K;
    ef_t *eqr = (ef_t *)malloc(M * sizeof(ef_t));
    or_eval_eq(ch + 1, n - skips, eqr);
    ef_t sub[64];
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < K; j++) {
        ef_t acc = ef_zero();
        for (size_t i = 0; i < M; i++) acc = ef_add(acc, ef_mul(mixed[(size_t)j * M + i], eqr[i]));
        sub[j] = acc;
    }
    free(eqr);
    or_ts_add_ef(ts, sub, (size_t)K);
    ef_t *pt = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    for (int i = 0; i < skips; i++) pt[i] = or_ts_sample_ef(ts);
    for (int i = 0; i < n - skips; i++) pt[skips + i] = ch[1 + i];
    ef_t *mu = (ef_t *)malloc(N * sizeof(ef_t));
    ef_t *md = (ef_t *)malloc(N * sizeof(ef_t));
    matrix_up_folded(pt, n, mu);
    matrix_down_folded(pt, n, md);
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) mu[x] = ef_add(mu[x], ef_mul(md[x], alpha2));
    free(md);
    ef_t inner_sum = or_mle_eval_ef(mixed, n, pt);
    free(mixed);
    const void *imles[2] = {mu, batched};
    or_sc_comp_t icomp = {OR_SC_PRODUCT, NULL, NULL, 0, NULL};
    ef_t *ich = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    ef_t ie[2], isum;
    or_sc_prove(1, 2, 0, imles, n, &icomp, 2, NULL, 0, ts, inner_sum, sec, ich, ie, &isum);
    free(mu);
    free(batched);
    ef_t *fp = (ef_t *)malloc((size_t)v * sizeof(ef_t));
    for (int i = 0; i < lc; i++) fp[i] = cbs[i];
    for (int i = 0; i < n; i++) fp[lc + i] = ich[i];
    or_whir_prove(&w, ts, fp, ie[1]);
    or_whir_witness_free(&w);
    free(fp);
    free(ich);
    free(pt);
    free(ch);
    free(zc);
}

/* crates/air/src/verify.rs:55-249.  Error codes follow AirVerifError order:
 * 1 InvalidPcsCommitment, 2 InvalidPcsOpening, 3 Fs, 4 Sumcheck,
 * 5 InvalidBoundaryCondition, 6 SumMismatch. */
int or_air_verify(const or_air_settings_t *s, const or_p2_bench_t *b, const f_t *proof,
                  size_t len) {
    int n = b->log_n, skips = s->univariate_skips, sec = s->security_bits;
    int lc = (int)or_log2_up(NC);
    int K = 1 << skips;
    or_whir_cfg_t cfg;
    or_air_whir_config(&cfg, s, n);
    f_t domsep[4096];
    size_t dn = or_whir_domsep(&cfg, domsep, 4096);
    or_ts_t ts;
    or_ts_init_verifier(&ts, &b->p16, domsep, dn, proof, len);
    int rc = 0;
    ef_t *zc = NULL, *opt = NULL, *ipt = NULL, *lde = NULL, *fp = NULL;
    or_whir_commitment_t cm;
    if (or_whir_parse_commitment(&cfg, &ts, &cm)) { rc = 1; goto out; }
    if (or_ts_check_pow(&ts, pow_bits_for(sec, (int)or_log2_up(NCONS)))) { rc = 3; goto out; }
    ef_t alpha = or_ts_sample_ef(&ts);
    if (or_ts_check_pow(&ts, pow_bits_for(sec, n))) { rc = 3; goto out; }
    int nz = n - skips + 1;
    zc = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    for (int i = 0; i < nz; i++) zc[i] = or_ts_sample_ef(&ts);
    ef_t scsum, oval;
    opt = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    if (or_sc_verify_uskip(&ts, 3 + 1, n, skips, sec, &scsum, opt, &oval)) { rc = 4; goto out; }
    if (!ef_is_zero(scsum)) { rc = 6; goto out; }
    ef_t wud[2 * NC];
    if (or_ts_next_ef(&ts, wud, 2 * NC)) { rc = 3; goto out; }
    ef_t osel[64], zsel[64];
    or_selectors_at_ef(skips, opt[0], osel);
    ef_t ap[NCONS];
    ap[0] = ef_one();
    for (int i = 1; i < NCONS; i++) ap[i] = ef_mul(ap[i - 1], alpha);
    ef_t gce = or_air_eval_ef(&b->air, wud, ap);
    or_selectors_at_ef(skips, zc[0], zsel);
    ef_t dot = ef_zero();
    for (int i = 0; i < K; i++) dot = ef_add(dot, ef_mul(zsel[i], osel[i]));
    ef_t lhs = ef_mul(ef_mul(dot, or_eq_poly_outside(zc + 1, opt + 1, n - skips)), gce);
    if (!ef_eq(lhs, oval)) { rc = 6; goto out; }
    if (or_ts_check_pow(&ts, pow_bits_for(sec, lc))) { rc = 3; goto out; }
    ef_t cbs[16];
    for (int i = 0; i < lc; i++) cbs[i] = or_ts_sample_ef(&ts);
    ef_t alpha2 = or_ts_sample_ef(&ts);
    ef_t sub[64];
    if (or_ts_next_ef(&ts, sub, (size_t)K)) { rc = 3; goto out; }
    ef_t eqc[256];
    or_eval_eq(cbs, lc, eqc);
    ef_t l = ef_zero(), ru = ef_zero(), rd = ef_zero();
    for (int i = 0; i < K; i++) l = ef_add(l, ef_mul(sub[i], osel[i]));
    for (int c = 0; c < NC; c++) {
        ru = ef_add(ru, ef_mul(wud[c], eqc[c]));
        rd = ef_add(rd, ef_mul(wud[NC + c], eqc[c]));
    }
    if (!ef_eq(l, ef_add(ru, ef_mul(rd, alpha2)))) { rc = 6; goto out; }
    ef_t eps[16];
    for (int i = 0; i < skips; i++) eps[i] = or_ts_sample_ef(&ts);
    ef_t isum, ival;
    ipt = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    if (or_sc_verify(&ts, n, 2, sec, &isum, ipt, &ival)) { rc = 4; goto out; }
    if (!ef_eq(isum, or_mle_eval_ef(sub, skips, eps))) { rc = 6; goto out; }
    lde = (ef_t *)malloc((size_t)(2 * n) * sizeof(ef_t));
    for (int i = 0; i < skips; i++) lde[i] = eps[i];
    for (int i = 0; i < n - skips; i++) lde[skips + i] = opt[1 + i];
    for (int i = 0; i < n; i++) lde[n + i] = ipt[i];
    ef_t up = matrix_up_lde(lde, 2 * n), dn2 = matrix_down_lde(lde, 2 * n);
    ef_t expected = ef_mul(ival, ef_inv(ef_add(up, ef_mul(alpha2, dn2))));
    fp = (ef_t *)malloc((size_t)(lc + n) * sizeof(ef_t));
    for (int i = 0; i < lc; i++) fp[i] = cbs[i];
    for (int i = 0; i < n; i++) fp[lc + i] = ipt[i];
    if (or_whir_verify(&cfg, &b->p16, &b->p24, &ts, &cm, fp, expected)) { rc = 2; goto out; }
out:
    free(zc);
    free(opt);
    free(ipt);
    free(lde);
    free(fp);
    or_ts_free(&ts);
    return rc;
}

/* ======================================================== generic AIR ======
 * AirTable::prove / verify for any AIR whose constraints are a straight-line
 * program over the point [up(0..C), down(0..C)] (SumcheckComputation for
 * A: Air, crates/sumcheck/src/sc_computation.rs:17-36), with the first
 * n_preprocessed columns preprocessed (crates/air/src/table.rs:18,60-66):
 * they join the zerocheck (prove.rs:109-113) but are neither committed nor sent;
 * the verifier evaluates them itself (verify.rs:114-136). */

int or_gair_whir_config(or_whir_cfg_t *c, const or_air_settings_t *s, const or_gair_t *a) {
    int v = a->log_n + (int)or_log2_up((size_t)(a->n_columns - a->n_preprocessed));
    return or_whir_config(c, v, s->security_bits, WHIR_POW_BITS, s->folding_f0, s->folding_f,
                          s->log_inv_rate, s->rs_reduction0, s->soundness);
}

void or_gair_prove(const or_air_settings_t *s, const or_gair_t *a, const f_t *witness, const or_p2_t *p16,
                   const or_p2_t *p24, or_ts_t *ts) {
    int n = a->log_n, skips = s->univariate_skips, sec = s->security_bits;
    int C = a->n_columns, np = a->n_preprocessed, nw = C - np;
    int ncons = or_prog_n_asserts(&a->prog);
    size_t N = (size_t)1 << n;
    int lc = (int)or_log2_up((size_t)nw);
    or_whir_cfg_t cfg;
    or_gair_whir_config(&cfg, s, a);
    int v = n + lc;
    f_t *packed = (f_t *)calloc((size_t)1 << v, sizeof(f_t));
    memcpy(packed, witness, N * (size_t)nw * sizeof(f_t));
    or_whir_witness_t w;
    or_whir_commit(&w, &cfg, p16, p24, ts, packed);
    free(packed);

    or_ts_pow(ts, pow_bits_for(sec, (int)or_log2_up((size_t)ncons)));
    ef_t alpha = or_ts_sample_ef(ts);
    ef_t *ap = (ef_t *)malloc((size_t)(ncons ? ncons : 1) * sizeof(ef_t));
    if (ncons) ap[0] = ef_one();
    for (int i = 1; i < ncons; i++) ap[i] = ef_mul(ap[i - 1], alpha);
    or_ts_pow(ts, pow_bits_for(sec, n));
    int nz = n + 1 - skips;
    ef_t *zc = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    for (int i = 0; i < nz; i++) zc[i] = or_ts_sample_ef(ts);

    /* columns_up_and_down of preprocessed ++ witness (crates/air/src/utils.rs:140-160) */
    f_t *ud = (f_t *)malloc(2 * (size_t)C * N * sizeof(f_t));
    const void **mles = (const void **)malloc(2 * (size_t)C * sizeof(void *));
    for (int c = 0; c < C; c++) {
        const f_t *col = c < np ? a->preprocessed + (size_t)c * N : witness + (size_t)(c - np) * N;
        f_t *up = ud + (size_t)c * N, *dn = ud + (size_t)(C + c) * N;
        memcpy(up, col, N * sizeof(f_t));
        up[N - 1] = col[N - 2];
        memcpy(dn, col + 1, (N - 1) * sizeof(f_t));
        dn[N - 1] = col[N - 1];
        mles[c] = up;
        mles[C + c] = dn;
    }
    or_sc_comp_t comp = {OR_SC_PROGRAM, NULL, ap, C, &a->prog};
    ef_t *ch = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    ef_t *fe = (ef_t *)malloc(2 * (size_t)C * sizeof(ef_t));
    ef_t fsum;
    or_sc_prove(skips, 2 * C, 1, mles, n, &comp, a->degree, zc, 1, ts, ef_zero(), sec, ch, fe, &fsum);
    free(ud);
    free(mles);

    or_ts_add_ef(ts, fe + np, (size_t)nw);
    or_ts_add_ef(ts, fe + C + np, (size_t)nw);
    or_ts_pow(ts, pow_bits_for(sec, lc));
    ef_t cbs[32];
    for (int i = 0; i < lc; i++) cbs[i] = or_ts_sample_ef(ts);
    ef_t *eqc = (ef_t *)malloc(((size_t)1 << lc) * sizeof(ef_t));
    or_eval_eq(cbs, lc, eqc);
    ef_t *batched = (ef_t *)malloc(N * sizeof(ef_t));
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) {
        ef_t acc = ef_zero();
        for (int c = 0; c < nw; c++) acc = ef_add(acc, ef_mul_f(eqc[c], witness[(size_t)c * N + x]));
        batched[x] = acc;
    }
    ef_t alpha2 = or_ts_sample_ef(ts);
    ef_t *mixed = (ef_t *)malloc(N * sizeof(ef_t));
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) {
        ef_t up = x == N - 1 ? batched[N - 2] : batched[x];
        ef_t dn = x == N - 1 ? batched[N - 1] : batched[x + 1];
        mixed[x] = ef_add(up, ef_mul(dn, alpha2));
    }
    int K = 1 << skips;
    size_t M = N / (size_t)K;
    ef_t *eqr = (ef_t *)malloc(M * sizeof(ef_t));
    or_eval_eq(ch + 1, n - skips, eqr);
    ef_t sub[64];
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < K; j++) {
        ef_t acc = ef_zero();
        for (size_t i = 0; i < M; i++) acc = ef_add(acc, ef_mul(mixed[(size_t)j * M + i], eqr[i]));
        sub[j] = acc;
    }
    free(eqr);
    or_ts_add_ef(ts, sub, (size_t)K);
    ef_t *pt = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    for (int i = 0; i < skips; i++) pt[i] = or_ts_sample_ef(ts);
    for (int i = 0; i < n - skips; i++) pt[skips + i] = ch[1 + i];
    ef_t *mu = (ef_t *)malloc(N * sizeof(ef_t));
    ef_t *md = (ef_t *)malloc(N * sizeof(ef_t));
    matrix_up_folded(pt, n, mu);
    matrix_down_folded(pt, n, md);
#pragma omp parallel for schedule(static)
    for (size_t x = 0; x < N; x++) mu[x] = ef_add(mu[x], ef_mul(md[x], alpha2));
    free(md);
    ef_t inner_sum = or_mle_eval_ef(mixed, n, pt);
    free(mixed);
    const void *imles[2] = {mu, batched};
    or_sc_comp_t icomp = {OR_SC_PRODUCT, NULL, NULL, 0, NULL};
    ef_t *ich = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    ef_t ie[2], isum;
    or_sc_prove(1, 2, 0, imles, n, &icomp, 2, NULL, 0, ts, inner_sum, sec, ich, ie, &isum);
    free(mu);
    free(batched);
    ef_t *fp = (ef_t *)malloc((size_t)v * sizeof(ef_t));
    for (int i = 0; i < lc; i++) fp[i] = cbs[i];
    for (int i = 0; i < n; i++) fp[lc + i] = ich[i];
    or_whir_prove(&w, ts, fp, ie[1]);
    or_whir_witness_free(&w);
    free(fp);
    free(ich);
    free(pt);
    free(ch);
    free(fe);
    free(zc);
    free(eqc);
    free(ap);
}

/* column_up / column_down of a preprocessed column folded by the outer selectors
 * and evaluated at the rest of the outer point (crates/air/src/verify.rs:114-136) */
static ef_t pre_eval(const f_t *col, int n, int skips, int down, const ef_t *osel, const ef_t *rest) {
    size_t N = (size_t)1 << n, M = N >> skips;
    int K = 1 << skips;
    ef_t *folded = (ef_t *)malloc(M * sizeof(ef_t));
    for (size_t i = 0; i < M; i++) {
        ef_t acc = ef_zero();
        for (int j = 0; j < K; j++) {
            size_t x = (size_t)j * M + i;
            f_t val = down ? (x == N - 1 ? col[N - 1] : col[x + 1]) : (x == N - 1 ? col[N - 2] : col[x]);
            acc = ef_add(acc, ef_mul_f(osel[j], val));
        }
        folded[i] = acc;
    }
    ef_t r = or_mle_eval_ef(folded, n - skips, rest);
    free(folded);
    return r;
}

int or_gair_verify(const or_air_settings_t *s, const or_gair_t *a, const or_p2_t *p16, const or_p2_t *p24,
                   or_ts_t *ts) {
    int n = a->log_n, skips = s->univariate_skips, sec = s->security_bits;
    int C = a->n_columns, np = a->n_preprocessed, nw = C - np;
    int ncons = or_prog_n_asserts(&a->prog);
    int lc = (int)or_log2_up((size_t)nw);
    int K = 1 << skips;
    or_whir_cfg_t cfg;
    if (or_gair_whir_config(&cfg, s, a)) return 1;
    int rc = 0;
    ef_t *zc = NULL, *opt = NULL, *ipt = NULL, *lde = NULL, *fp = NULL, *wud = NULL, *gp = NULL, *ap = NULL,
         *eqc = NULL;
    or_whir_commitment_t cm;
    if (or_whir_parse_commitment(&cfg, ts, &cm)) { rc = 1; goto out; }
    if (or_ts_check_pow(ts, pow_bits_for(sec, (int)or_log2_up((size_t)ncons)))) { rc = 3; goto out; }
    ef_t alpha = or_ts_sample_ef(ts);
    if (or_ts_check_pow(ts, pow_bits_for(sec, n))) { rc = 3; goto out; }
    int nz = n - skips + 1;
    zc = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    for (int i = 0; i < nz; i++) zc[i] = or_ts_sample_ef(ts);
    ef_t scsum, oval;
    opt = (ef_t *)malloc((size_t)nz * sizeof(ef_t));
    if (or_sc_verify_uskip(ts, a->degree + 1, n, skips, sec, &scsum, opt, &oval)) { rc = 4; goto out; }
    if (!ef_is_zero(scsum)) { rc = 6; goto out; }
    wud = (ef_t *)malloc(2 * (size_t)nw * sizeof(ef_t));
    if (or_ts_next_ef(ts, wud, 2 * (size_t)nw)) { rc = 3; goto out; }
    ef_t osel[64], zsel[64];
    or_selectors_at_ef(skips, opt[0], osel);
    /* global point [pre_up, wit_up, pre_down, wit_down] */
    gp = (ef_t *)malloc(2 * (size_t)C * sizeof(ef_t));
    for (int c = 0; c < np; c++) {
        const f_t *col = a->preprocessed + (size_t)c * ((size_t)1 << n);
        gp[c] = pre_eval(col, n, skips, 0, osel, opt + 1);
        gp[C + c] = pre_eval(col, n, skips, 1, osel, opt + 1);
    }
    for (int c = 0; c < nw; c++) {
        gp[np + c] = wud[c];
        gp[C + np + c] = wud[nw + c];
    }
    ap = (ef_t *)malloc((size_t)(ncons ? ncons : 1) * sizeof(ef_t));
    if (ncons) ap[0] = ef_one();
    for (int i = 1; i < ncons; i++) ap[i] = ef_mul(ap[i - 1], alpha);
    ef_t gce = or_prog_eval_ef(&a->prog, gp, ap);
    or_selectors_at_ef(skips, zc[0], zsel);
    ef_t dot = ef_zero();
    for (int i = 0; i < K; i++) dot = ef_add(dot, ef_mul(zsel[i], osel[i]));
    ef_t lhs = ef_mul(ef_mul(dot, or_eq_poly_outside(zc + 1, opt + 1, n - skips)), gce);
    if (!ef_eq(lhs, oval)) { rc = 6; goto out; }
    if (or_ts_check_pow(ts, pow_bits_for(sec, lc))) { rc = 3; goto out; }
    ef_t cbs[32];
    for (int i = 0; i < lc; i++) cbs[i] = or_ts_sample_ef(ts);
    ef_t alpha2 = or_ts_sample_ef(ts);
    ef_t sub[64];
    if (or_ts_next_ef(ts, sub, (size_t)K)) { rc = 3; goto out; }
    eqc = (ef_t *)malloc(((size_t)1 << lc) * sizeof(ef_t));
    or_eval_eq(cbs, lc, eqc);
    ef_t l = ef_zero(), ru = ef_zero(), rd = ef_zero();
    for (int i = 0; i < K; i++) l = ef_add(l, ef_mul(sub[i], osel[i]));
    for (int c = 0; c < nw; c++) {
        ru = ef_add(ru, ef_mul(wud[c], eqc[c]));
        rd = ef_add(rd, ef_mul(wud[nw + c], eqc[c]));
    }
    if (!ef_eq(l, ef_add(ru, ef_mul(rd, alpha2)))) { rc = 6; goto out; }
    ef_t eps[16];
    for (int i = 0; i < skips; i++) eps[i] = or_ts_sample_ef(ts);
    ef_t isum, ival;
    ipt = (ef_t *)malloc((size_t)n * sizeof(ef_t));
    if (or_sc_verify(ts, n, 2, sec, &isum, ipt, &ival)) { rc = 4; goto out; }
    if (!ef_eq(isum, or_mle_eval_ef(sub, skips, eps))) { rc = 6; goto out; }
    lde = (ef_t *)malloc((size_t)(2 * n) * sizeof(ef_t));
    for (int i = 0; i < skips; i++) lde[i] = eps[i];
    for (int i = 0; i < n - skips; i++) lde[skips + i] = opt[1 + i];
    for (int i = 0; i < n; i++) lde[n + i] = ipt[i];
    ef_t up = matrix_up_lde(lde, 2 * n), dn2 = matrix_down_lde(lde, 2 * n);
    ef_t expected = ef_mul(ival, ef_inv(ef_add(up, ef_mul(alpha2, dn2))));
    fp = (ef_t *)malloc((size_t)(lc + n) * sizeof(ef_t));
    for (int i = 0; i < lc; i++) fp[i] = cbs[i];
    for (int i = 0; i < n; i++) fp[lc + i] = ipt[i];
    if (or_whir_verify(&cfg, p16, p24, ts, &cm, fp, expected)) { rc = 2; goto out; }
out:
    free(zc);
    free(opt);
    free(ipt);
    free(lde);
    free(fp);
    free(wud);
    free(gp);
    free(ap);
    free(eqc);
    return rc;
}
