/* whir.c -- WHIR polynomial commitment: config, RS/NTT, Merkle, prove, verify (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference delegates the commitment to whir-p3 @ 9caef91
 * (crates/air/src/prove.rs:82-91,216-225; verify.rs:70-76,239-246;
 * table.rs:72-98).  Its source is absent, so this restates the published
 * WHIR protocol in whir-p3's structure; PARITY UNPINNED for the exact byte
 * layout.  Conventions fixed here (and shared with the HIP implementation):
 *  - WhirConfig::new round schedule for FoldingFactor::ConstantFromSecondRound
 *    (f0, f), CapacityBound soundness, 248-bit extension field.
 *  - A polynomial with m variables and folding factor k is committed as an
 *    interleaved Reed-Solomon code: column b in {0,1}^k (the top k variables,
 *    big-endian) holds the univariate embedding of f(b, .), i.e. the monomial
 *    coefficients (Mobius transform of the evaluations) evaluated by a
 *    natural-order NTT over the 2^(log_domain - k)-th roots of unity.  Leaf r
 *    is row r (F columns for the initial F commitment, EF columns flattened
 *    coefficient-wise otherwise), hashed with PaddingFreeSponge<P24,24,16,8>;
 *    nodes use TruncatedPermutation<P16,2,8,16>.
 *  - Sumcheck binds the most-significant variable first; round polynomial
 *    sent as evaluations at 0, 1, 2.
 *  - expand_from_univariate(y, m) = (y^(2^(m-1)), ..., y^2, y).
 *  - Query r of a codeword with 2^L rows is the point w_L^r; its folded
 *    value is sum_b eq(b, r_fold) leaf_r[b].
 *  - Queries: num_queries x sample_bits(L), sorted, deduplicated; each opened
 *    as hint: row values, then L sibling digests bottom-up.
 * Pinned by: NTT vs direct evaluation, Mobius vs multilinear evaluation,
 * verifier acceptance and tamper rejection, and the invariant that folded
 * query values equal the folded polynomial at the expanded point.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ------------------------------------------------------------ config -- */
/* SecurityAssumption (whir-p3 @ 9caef91 parameters/errors.rs, restated from its
 * published form; parity unpinned): the error terms of the WHIR round-by-round
 * soundness analysis for the three decoding regimes.  eta is the proximity slack:
 * JohnsonBound eta = sqrt(rho)/20 (log2 = -(r/2 + log2(10) + 1)), CapacityBound
 * eta = rho/2 (log2 = -(r + 1)), none in unique decoding. */
static const int FIELD_BITS = 248;
static const int MAX_VARS_SEND_COEFFS = 6;
static const double LOG2_10 = 3.321928094887362;

static double sa_log_eta(int sa, int r) {
    if (sa == OR_SA_UNIQUE_DECODING) return 0.0;
    if (sa == OR_SA_JOHNSON_BOUND) return -(0.5 * r + LOG2_10 + 1.0);
    return -((double)r + 1.0);
}
static double sa_list_size_bits(int sa, int log_degree, int r) {
    double le = sa_log_eta(sa, r);
    if (sa == OR_SA_UNIQUE_DECODING) return 0.0;
    if (sa == OR_SA_JOHNSON_BOUND) return 0.5 * r - (1.0 + le);
    return (double)(log_degree + r) - le;
}
static double sa_prox_gaps_error(int sa, int log_degree, int r, int num_functions) {
    double le = sa_log_eta(sa, r), err;
    if (sa == OR_SA_UNIQUE_DECODING) err = (double)(log_degree + r);
    else if (sa == OR_SA_JOHNSON_BOUND) {
        double sqrt_rho_20 = 1.0 + LOG2_10 + 0.5 * r;
        err = 2.0 * log_degree + 7.0 * ((sqrt_rho_20 < -le ? sqrt_rho_20 : -le) - 1.0);
    } else err = (double)(log_degree + 2 * r) - le;
    return (double)FIELD_BITS - (err + log2((double)num_functions - 1.0));
}
static double sa_ood_error(int sa, int log_degree, int r, int samples) {
    double lsb = sa_list_size_bits(sa, log_degree, r);
    double err = 2.0 * lsb + (double)(log_degree * samples);
    return (double)(samples * FIELD_BITS) + 1.0 - err;
}
static int sa_ood_samples(int sa, int sec, int log_degree, int r) {
    if (sa == OR_SA_UNIQUE_DECODING) return 0;
    for (int s = 1; s < 64; s++)
        if (sa_ood_error(sa, log_degree, r, s) >= (double)sec) return s;
    return -1;
}
/* -log2 of the per-query survival probability (1 - delta) */
static double sa_query_bits(int sa, int r) {
    if (sa == OR_SA_UNIQUE_DECODING) return -log2(0.5 * (1.0 + 1.0 / (double)(1 << r)));
    if (sa == OR_SA_JOHNSON_BOUND) return 0.5 * r;
    return (double)r;
}
static int sa_queries(int sa, int protocol_sec, int r) {
    return (int)ceil((double)protocol_sec / sa_query_bits(sa, r));
}
static double sa_queries_error(int sa, int r, int q) { return (double)q * sa_query_bits(sa, r); }
static double fold_sumcheck_error(int sa, int nv, int r) {
    return (double)FIELD_BITS - (sa_list_size_bits(sa, nv, r) + 1.0);
}
static int folding_pow_bits(int sa, int sec, int nv, int r) {
    double e1 = sa_prox_gaps_error(sa, nv, r, 2);
    double e2 = fold_sumcheck_error(sa, nv, r);
    double e = e1 < e2 ? e1 : e2;
    double x = (double)sec - e;
    return x > 0 ? (int)ceil(x) : 0;
}
static double queries_combination_error(int sa, int nv, int r, int ood, int q) {
    double lsb = sa_list_size_bits(sa, nv, r);
    return (double)FIELD_BITS - (log2((double)(ood + q)) + lsb + 1.0);
}

int or_whir_folding_at(const or_whir_cfg_t *c, int round) { return round == 0 ? c->f0 : c->f; }

int or_whir_config(or_whir_cfg_t *c, int v, int sec, int max_pow, int f0, int f, int rate,
                   int rs0, int sa) {
    memset(c, 0, sizeof *c);
    if (f0 > v || f <= 0 || rs0 > f0 || sa < 0 || sa > 2) return -1;
    c->soundness = sa;
    c->num_variables = v;
    c->security_level = sec;
    c->max_pow_bits = max_pow;
    c->f0 = f0;
    c->f = f;
    c->starting_log_inv_rate = rate;
    c->rs_reduction0 = rs0;
    int nv1 = v - f0;
    if (nv1 < MAX_VARS_SEND_COEFFS) {
        c->n_rounds = 0;
        c->final_sumcheck_rounds = nv1;
    } else {
        c->n_rounds = (nv1 - MAX_VARS_SEND_COEFFS + f - 1) / f;
        c->final_sumcheck_rounds = nv1 - c->n_rounds * f;
    }
    if (c->n_rounds > OR_WHIR_MAX_ROUNDS) return -1;
    int psec = sec - max_pow;
    if (psec < 0) psec = 0;
    c->commit_ood_samples = sa_ood_samples(sa, sec, v, rate);
    c->starting_folding_pow_bits = folding_pow_bits(sa, sec, v, rate);
    if (c->commit_ood_samples < 0) return -1;
    int nv = v - f0, lir = rate, logd = v + rate;
    for (int r = 0; r < c->n_rounds; r++) {
        int rs = r == 0 ? rs0 : 1;
        int fold = or_whir_folding_at(c, r);
        int next_rate = lir + (fold - rs);
        int q = sa_queries(sa, psec, lir);
        int ood = sa_ood_samples(sa, sec, nv, next_rate);
        if (ood < 0) return -1;
        double qe = sa_queries_error(sa, lir, q);
        double ce = queries_combination_error(sa, nv, next_rate, ood, q);
        double pw = (double)sec - (qe < ce ? qe : ce);
        or_whir_round_t *rp = &c->rounds[r];
        rp->pow_bits = pw > 0 ? (int)ceil(pw) : 0;
        rp->folding_pow_bits = folding_pow_bits(sa, sec, nv, next_rate);
        rp->num_queries = q;
        rp->ood_samples = ood;
        rp->log_inv_rate = lir;
        rp->num_variables = nv;
        rp->folding_factor = fold;
        rp->log_domain_size = logd;
        nv -= or_whir_folding_at(c, r + 1);
        lir = next_rate;
        logd -= rs;
    }
    c->final_queries = sa_queries(sa, psec, lir);
    double fp = (double)sec - sa_queries_error(sa, lir, c->final_queries);
    c->final_pow_bits = fp > 0 ? (int)ceil(fp) : 0;
    double ffp = (double)sec - (double)(FIELD_BITS - 1);
    c->final_folding_pow_bits = ffp > 0 ? (int)ceil(ffp) : 0;
    c->final_log_inv_rate = lir;
    c->final_log_domain_size = logd;
    /* every committed codeword lives on a power-of-two subgroup of F*: at most 2^24 */
    if (v + rate - f0 > 24) return -1;
    for (int r = 0; r < c->n_rounds; r++) {
        int rs = r == 0 ? rs0 : 1;
        if (c->rounds[r].log_domain_size - rs - or_whir_folding_at(c, r + 1) > 24) return -1;
    }
    return 0;
}

/* Domain separator: (tag, count) pairs for every transcript operation of the
 * commitment and the opening, absorbed before the first message.
 * tags: 1 observe, 2 sample, 3 hint, 4 pow.  (whir-p3 DomainSeparator
 * commit_statement + add_whir_proof; encoding parity unpinned.) */
size_t or_whir_domsep(const or_whir_cfg_t *c, f_t *out, size_t cap) {
    size_t n = 0;
#define OP(t, k) do { if (n + 2 <= cap) { out[n] = (t); out[n + 1] = (f_t)(k); } n += 2; } while (0)
    OP(1, 8);
    if (c->commit_ood_samples) { OP(2, 8 * c->commit_ood_samples); OP(1, 8 * c->commit_ood_samples); }
    OP(2, 8);
    for (int k = 0; k < c->f0; k++) { OP(1, 24); if (c->starting_folding_pow_bits) OP(4, c->starting_folding_pow_bits); OP(2, 8); }
    for (int r = 0; r < c->n_rounds; r++) {
        const or_whir_round_t *rp = &c->rounds[r];
        OP(1, 8);
        if (rp->ood_samples) { OP(2, 8 * rp->ood_samples); OP(1, 8 * rp->ood_samples); }
        if (rp->pow_bits) OP(4, rp->pow_bits);
        OP(2, rp->num_queries);
        OP(3, 0);
        OP(2, 8);
        for (int k = 0; k < or_whir_folding_at(c, r + 1); k++) { OP(1, 24); if (rp->folding_pow_bits) OP(4, rp->folding_pow_bits); OP(2, 8); }
    }
    OP(1, 8 << c->final_sumcheck_rounds);
    if (c->final_pow_bits) OP(4, c->final_pow_bits);
    OP(2, c->final_queries);
    OP(3, 0);
    for (int k = 0; k < c->final_sumcheck_rounds; k++) { OP(1, 24); if (c->final_folding_pow_bits) OP(4, c->final_folding_pow_bits); OP(2, 8); }
#undef OP
    return n;
}

/* ----------------------------------------------------------- RS / NTT -- */
void or_ntt(f_t *a, int log_n) {
    size_t n = (size_t)1 << log_n;
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { f_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int s = 1; s <= log_n; s++) {
        size_t len = (size_t)1 << s, half = len >> 1;
        f_t wl = f_two_adic_generator(s);
        f_t *tw = (f_t *)malloc(half * sizeof(f_t));
        tw[0] = 1;
        for (size_t k = 1; k < half; k++) tw[k] = f_mul(tw[k - 1], wl);
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; k++) {
                f_t u = a[i + k], v = f_mul(a[i + k + half], tw[k]);
                a[i + k] = f_add(u, v);
                a[i + k + half] = f_sub(u, v);
            }
        free(tw);
    }
}

void or_mobius(f_t *a, int log_n) {
    size_t n = (size_t)1 << log_n;
    for (int b = 0; b < log_n; b++) {
        size_t bit = (size_t)1 << b;
        for (size_t i = 0; i < n; i++)
            if (i & bit) a[i] = f_sub(a[i], a[i ^ bit]);
    }
}

/* evals: n_cols contiguous blocks of 2^log_m F; cw: n_cols blocks of 2^log_rows */
void or_rs_encode_columns(const f_t *evals, int log_m, int n_cols, int log_rows, f_t *cw) {
    size_t m = (size_t)1 << log_m, rows = (size_t)1 << log_rows;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < n_cols; b++) {
        f_t *col = cw + (size_t)b * rows;
        memset(col, 0, rows * sizeof(f_t));
        memcpy(col, evals + (size_t)b * m, m * sizeof(f_t));
        or_mobius(col, log_m);
        or_ntt(col, log_rows);
    }
}

/* ---------------------------------------------------------- Merkle -- */
void or_tree_build(or_tree_t *t, const or_p2_t *p16, const or_p2_t *p24) {
    size_t rows = (size_t)1 << t->log_rows;
    t->nodes = (f_t *)malloc(2 * rows * 8 * sizeof(f_t));
#pragma omp parallel
    {
        f_t *row = (f_t *)malloc((size_t)t->width * sizeof(f_t));
#pragma omp for schedule(static)
        for (size_t r = 0; r < rows; r++) {
            for (int c = 0; c < t->width; c++) row[c] = t->cw[(size_t)c * rows + r];
            or_hash_leaf(p24, row, (size_t)t->width, t->nodes + (rows + r) * 8);
        }
        free(row);
    }
    for (size_t lvl = rows >> 1; lvl >= 1; lvl >>= 1) {
#pragma omp parallel for schedule(static)
        for (size_t i = lvl; i < 2 * lvl; i++)
            or_compress(p16, t->nodes + 2 * i * 8, t->nodes + (2 * i + 1) * 8, t->nodes + i * 8);
    }
}

void or_tree_free(or_tree_t *t) {
    free(t->cw);
    free(t->nodes);
    t->cw = t->nodes = NULL;
}

static void tree_open(or_ts_t *ts, const or_tree_t *t, size_t r) {
    size_t rows = (size_t)1 << t->log_rows;
    for (int c = 0; c < t->width; c++) or_ts_hint_base(ts, &t->cw[(size_t)c * rows + r], 1);
    size_t idx = rows + r;
    for (int l = 0; l < t->log_rows; l++) {
        or_ts_hint_base(ts, t->nodes + (idx ^ 1) * 8, 8);
        idx >>= 1;
    }
}

static int tree_verify(const or_p2_t *p16, const or_p2_t *p24, const f_t root[8], int log_rows,
                       size_t r, const f_t *row, int width, const f_t *path) {
    f_t cur[8], tmp[8];
    or_hash_leaf(p24, row, (size_t)width, cur);
    size_t idx = r;
    for (int l = 0; l < log_rows; l++) {
        if (idx & 1) or_compress(p16, path + 8 * l, cur, tmp);
        else or_compress(p16, cur, path + 8 * l, tmp);
        memcpy(cur, tmp, sizeof cur);
        idx >>= 1;
    }
    return memcmp(cur, root, sizeof cur) == 0 ? 0 : -1;
}

/* ------------------------------------------------------------ commit -- */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

static int sample_queries(or_ts_t *ts, int n, int log_rows, uint32_t *out) {
    for (int i = 0; i < n; i++) out[i] = or_chal_sample_bits(&ts->ch, log_rows);
    qsort(out, (size_t)n, sizeof(uint32_t), cmp_u32);
    int u = 0;
    for (int i = 0; i < n; i++)
        if (u == 0 || out[u - 1] != out[i]) out[u++] = out[i];
    return u;
}

/* CommitmentWriter::commit (crates/air/src/prove.rs:82-91) */
void or_whir_commit(or_whir_witness_t *w, const or_whir_cfg_t *cfg, const or_p2_t *p16,
                    const or_p2_t *p24, or_ts_t *ts, const f_t *evals) {
    memset(w, 0, sizeof *w);
    w->cfg = cfg;
    w->p16 = p16;
    w->p24 = p24;
    int v = cfg->num_variables;
    size_t N = (size_t)1 << v;
    w->evals = (f_t *)malloc(N * sizeof(f_t));
    memcpy(w->evals, evals, N * sizeof(f_t));
    int f0 = cfg->f0;
    w->tree.log_rows = v + cfg->starting_log_inv_rate - f0;
    w->tree.width = 1 << f0;
    w->tree.cw = (f_t *)malloc(((size_t)1 << (v + cfg->starting_log_inv_rate)) * sizeof(f_t));
    or_rs_encode_columns(w->evals, v - f0, 1 << f0, w->tree.log_rows, w->tree.cw);
    or_tree_build(&w->tree, p16, p24);
    or_ts_add_base(ts, w->tree.nodes + 8, 8);
    w->n_ood = cfg->commit_ood_samples;
    ef_t *pt = (ef_t *)malloc((size_t)v * sizeof(ef_t));
    for (int i = 0; i < w->n_ood; i++) w->ood_points[i] = or_ts_sample_ef(ts);
    for (int i = 0; i < w->n_ood; i++) {
        or_expand_from_univariate(w->ood_points[i], v, pt);
        w->ood_answers[i] = or_mle_eval_f(w->evals, v, pt);
    }
    or_ts_add_ef(ts, w->ood_answers, (size_t)w->n_ood);
    free(pt);
}

void or_whir_witness_free(or_whir_witness_t *w) {
    free(w->evals);
    w->evals = NULL;
    or_tree_free(&w->tree);
}

/* ----------------------------------------------------------- prover -- */
typedef struct {
    int n;          /* variables */
    int base;       /* evals still in F */
    f_t *ef_base;   /* F evals (first round) */
    ef_t *e;        /* EF evals */
    ef_t *wts;      /* EF weights */
    ef_t sum;
} wsc_t;

static ef_t interp3(const ef_t h[3], ef_t r) {
    /* h(r) for the degree-2 polynomial through (0,h0),(1,h1),(2,h2) */
    ef_t one = ef_one(), two = ef_from_f(2);
    ef_t rm1 = ef_sub(r, one), rm2 = ef_sub(r, two);
    f_t inv2 = f_inv(2);
    ef_t t0 = ef_mul_f(ef_mul(ef_mul(h[0], rm1), rm2), inv2);
    ef_t t1 = ef_neg(ef_mul(ef_mul(h[1], r), rm2));
    ef_t t2 = ef_mul_f(ef_mul(ef_mul(h[2], r), rm1), inv2);
    return ef_add(ef_add(t0, t1), t2);
}

static void wsc_round(wsc_t *s, or_ts_t *ts, int pow_bits, ef_t *r_out) {
    size_t h = (size_t)1 << (s->n - 1);
    ef_t H0 = ef_zero(), H1 = ef_zero(), H2 = ef_zero();
#pragma omp parallel
    {
        ef_t a0 = ef_zero(), a1 = ef_zero(), a2 = ef_zero();
#pragma omp for schedule(static)
        for (size_t i = 0; i < h; i++) {
            ef_t w0 = s->wts[i], w1 = s->wts[i + h];
            ef_t w2 = ef_sub(ef_add(w1, w1), w0);
            if (s->base) {
                f_t e0 = s->ef_base[i], e1 = s->ef_base[i + h];
                f_t e2 = f_sub(f_add(e1, e1), e0);
                a0 = ef_add(a0, ef_mul_f(w0, e0));
                a1 = ef_add(a1, ef_mul_f(w1, e1));
                a2 = ef_add(a2, ef_mul_f(w2, e2));
            } else {
                ef_t e0 = s->e[i], e1 = s->e[i + h];
                ef_t e2 = ef_sub(ef_add(e1, e1), e0);
                a0 = ef_add(a0, ef_mul(w0, e0));
                a1 = ef_add(a1, ef_mul(w1, e1));
                a2 = ef_add(a2, ef_mul(w2, e2));
            }
        }
#pragma omp critical
        { H0 = ef_add(H0, a0); H1 = ef_add(H1, a1); H2 = ef_add(H2, a2); }
    }
    ef_t hv[3] = {H0, H1, H2};
    or_ts_add_ef(ts, hv, 3);
    or_ts_pow(ts, pow_bits);
    ef_t r = or_ts_sample_ef(ts);
    *r_out = r;
    s->sum = interp3(hv, r);
    if (s->base) {
        ef_t *ne = (ef_t *)malloc(h * sizeof(ef_t));
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < h; i++)
            ne[i] = ef_add(ef_from_f(s->ef_base[i]), ef_mul_f(r, f_sub(s->ef_base[i + h], s->ef_base[i])));
        s->e = ne;
        s->base = 0;
        s->ef_base = NULL; /* not owned */
    } else {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < h; i++) s->e[i] = ef_add(s->e[i], ef_mul(r, ef_sub(s->e[i + h], s->e[i])));
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < h; i++) s->wts[i] = ef_add(s->wts[i], ef_mul(r, ef_sub(s->wts[i + h], s->wts[i])));
    s->n--;
}

/* weights += scale * eq(., point) */
static void add_eq_ef(ef_t *wts, const ef_t *pt, int n, ef_t scale) {
    size_t N = (size_t)1 << n;
    ef_t *t = (ef_t *)malloc(N * sizeof(ef_t));
    or_eval_eq_scaled(pt, n, scale, t);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) wts[i] = ef_add(wts[i], t[i]);
    free(t);
}

/* weights += scale * eq(., expand(y)) for y in F: the eq table is over F */
static void add_eq_univ_f(ef_t *wts, f_t y, int n, ef_t scale) {
    size_t N = (size_t)1 << n;
    f_t *pt = (f_t *)malloc((size_t)(n ? n : 1) * sizeof(f_t));
    f_t cur = y;
    for (int i = 0; i < n; i++) { pt[n - 1 - i] = cur; cur = f_mul(cur, cur); }
    f_t *t = (f_t *)malloc(N * sizeof(f_t));
    or_eval_eq_f(pt, n, t);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) wts[i] = ef_add(wts[i], ef_mul_f(scale, t[i]));
    free(t);
    free(pt);
}

static ef_t eval_leaf_fold(const f_t *row, int width_f, int base, int fold, const ef_t *r) {
    int K = 1 << fold;
    ef_t *eq = (ef_t *)malloc((size_t)K * sizeof(ef_t));
    or_eval_eq(r, fold, eq);
    ef_t acc = ef_zero();
    for (int b = 0; b < K; b++) {
        if (base) acc = ef_add(acc, ef_mul_f(eq[b], row[b]));
        else { ef_t v; memcpy(v.c, row + 8 * b, sizeof v.c); acc = ef_add(acc, ef_mul(eq[b], v)); }
    }
    (void)width_f;
    free(eq);
    return acc;
}

/* Commit an EF polynomial (2^m evals) with folding factor k, domain 2^log_domain. */
static void commit_ef(or_tree_t *t, const ef_t *e, int m, int k, int log_domain,
                      const or_p2_t *p16, const or_p2_t *p24) {
    int log_rows = log_domain - k;
    int K = 1 << k;
    size_t cm = (size_t)1 << (m - k);
    /* split into 8K F columns: column (b, coord) */
    f_t *cols = (f_t *)malloc((size_t)K * 8 * cm * sizeof(f_t));
    for (int b = 0; b < K; b++)
        for (int j = 0; j < 8; j++)
            for (size_t i = 0; i < cm; i++) cols[((size_t)b * 8 + j) * cm + i] = e[(size_t)b * cm + i].c[j];
    t->log_rows = log_rows;
    t->width = 8 * K;
    t->cw = (f_t *)malloc((size_t)8 * K * ((size_t)1 << log_rows) * sizeof(f_t));
    or_rs_encode_columns(cols, m - k, 8 * K, log_rows, t->cw);
    free(cols);
    or_tree_build(t, p16, p24);
}

/* Prover::prove (crates/air/src/prove.rs:216-225) */
void or_whir_prove(or_whir_witness_t *w, or_ts_t *ts, const ef_t *point, ef_t value) {
    const or_whir_cfg_t *c = w->cfg;
    int v = c->num_variables;
    size_t N = (size_t)1 << v;
    ef_t *pt = (ef_t *)malloc((size_t)(v + 1) * sizeof(ef_t));
    wsc_t s;
    memset(&s, 0, sizeof s);
    s.n = v;
    s.base = 1;
    s.ef_base = w->evals;
    s.wts = (ef_t *)calloc(N, sizeof(ef_t));
    /* constraints: OOD first, then the statement */
    ef_t g = or_ts_sample_ef(ts), gp = ef_one();
    s.sum = ef_zero();
    for (int i = 0; i < w->n_ood; i++) {
        or_expand_from_univariate(w->ood_points[i], v, pt);
        add_eq_ef(s.wts, pt, v, gp);
        s.sum = ef_add(s.sum, ef_mul(gp, w->ood_answers[i]));
        gp = ef_mul(gp, g);
    }
    add_eq_ef(s.wts, point, v, gp);
    s.sum = ef_add(s.sum, ef_mul(gp, value));

    ef_t rfold[32];
    for (int k = 0; k < c->f0; k++) wsc_round(&s, ts, c->starting_folding_pow_bits, &rfold[k]);
    int prev_fold = c->f0;
    int prev_base = 1;
    or_tree_t prev = w->tree;  /* borrowed */
    or_tree_t own = {0, 0, NULL, NULL};
    int have_own = 0;
    uint32_t *qs = (uint32_t *)malloc(4096 * sizeof(uint32_t));
    for (int r = 0; r < c->n_rounds; r++) {
        const or_whir_round_t *rp = &c->rounds[r];
        int m = s.n;
        int rs = r == 0 ? c->rs_reduction0 : 1;
        int new_logd = rp->log_domain_size - rs;
        int k_next = or_whir_folding_at(c, r + 1);
        or_tree_t nt;
        commit_ef(&nt, s.e, m, k_next, new_logd, w->p16, w->p24);
        or_ts_add_base(ts, nt.nodes + 8, 8);
        ef_t oodp[8], ooda[8];
        for (int i = 0; i < rp->ood_samples; i++) oodp[i] = or_ts_sample_ef(ts);
        for (int i = 0; i < rp->ood_samples; i++) {
            or_expand_from_univariate(oodp[i], m, pt);
            ooda[i] = or_mle_eval_ef(s.e, m, pt);
        }
        or_ts_add_ef(ts, ooda, (size_t)rp->ood_samples);
        or_ts_pow(ts, rp->pow_bits);
        int log_rows = rp->log_domain_size - prev_fold;
        int nq = sample_queries(ts, rp->num_queries, log_rows, qs);
        ef_t *vals = (ef_t *)malloc((size_t)nq * sizeof(ef_t));
        size_t rowsN = (size_t)1 << prev.log_rows;
        f_t *row = (f_t *)malloc((size_t)prev.width * sizeof(f_t));
        for (int q = 0; q < nq; q++) {
            tree_open(ts, &prev, qs[q]);
            for (int cc = 0; cc < prev.width; cc++) row[cc] = prev.cw[(size_t)cc * rowsN + qs[q]];
            vals[q] = eval_leaf_fold(row, prev.width, prev_base, prev_fold, rfold);
        }
        free(row);
        ef_t gg = or_ts_sample_ef(ts), gpp = ef_one();
        for (int i = 0; i < rp->ood_samples; i++) {
            or_expand_from_univariate(oodp[i], m, pt);
            add_eq_ef(s.wts, pt, m, gpp);
            s.sum = ef_add(s.sum, ef_mul(gpp, ooda[i]));
            gpp = ef_mul(gpp, gg);
        }
        f_t gen = f_two_adic_generator(log_rows);
        for (int q = 0; q < nq; q++) {
            add_eq_univ_f(s.wts, f_pow(gen, qs[q]), m, gpp);
            s.sum = ef_add(s.sum, ef_mul(gpp, vals[q]));
            gpp = ef_mul(gpp, gg);
        }
        free(vals);
        for (int k = 0; k < k_next; k++) wsc_round(&s, ts, rp->folding_pow_bits, &rfold[k]);
        if (have_own) or_tree_free(&own);
        own = nt;
        have_own = 1;
        prev = own;
        prev_fold = k_next;
        prev_base = 0;
    }
    /* final round */
    if (s.base) {
        /* only possible when f0 == v; not reachable for our configs */
        ef_t *ne = (ef_t *)malloc(((size_t)1 << s.n) * sizeof(ef_t));
        for (size_t i = 0; i < ((size_t)1 << s.n); i++) ne[i] = ef_from_f(s.ef_base[i]);
        s.e = ne;
        s.base = 0;
    }
    or_ts_add_ef(ts, s.e, (size_t)1 << s.n);
    or_ts_pow(ts, c->final_pow_bits);
    {
        int log_rows = c->final_log_domain_size - prev_fold;
        int nq = sample_queries(ts, c->final_queries, log_rows, qs);
        for (int q = 0; q < nq; q++) tree_open(ts, &prev, qs[q]);
    }
    for (int k = 0; k < c->final_sumcheck_rounds; k++) wsc_round(&s, ts, c->final_folding_pow_bits, &rfold[k]);
    if (have_own) or_tree_free(&own);
    free(qs);
    free(pt);
    free(s.e);
    free(s.wts);
}

/* ---------------------------------------------------------- verifier -- */
typedef struct {
    int offset;     /* index of this set's first variable in the global point */
    int n;          /* variables */
    int count;
    ef_t gamma;
    ef_t *pts;      /* EF points (count * n) */
} cset_t;

static int vsc_rounds(or_ts_t *ts, int k, int pow_bits, ef_t *sum, ef_t *rs) {
    for (int i = 0; i < k; i++) {
        ef_t h[3];
        if (or_ts_next_ef(ts, h, 3)) return -1;
        if (!ef_eq(ef_add(h[0], h[1]), *sum)) return -2;
        if (or_ts_check_pow(ts, pow_bits)) return -3;
        rs[i] = or_ts_sample_ef(ts);
        *sum = interp3(h, rs[i]);
    }
    return 0;
}

/* CommitmentReader::parse_commitment (crates/air/src/verify.rs:72-76) */
int or_whir_parse_commitment(const or_whir_cfg_t *c, or_ts_t *ts, or_whir_commitment_t *out) {
    memset(out, 0, sizeof *out);
    if (or_ts_next_base(ts, out->root, 8)) return -1;
    out->n_ood = c->commit_ood_samples;
    for (int i = 0; i < out->n_ood; i++) out->ood_points[i] = or_ts_sample_ef(ts);
    if (or_ts_next_ef(ts, out->ood_answers, (size_t)out->n_ood)) return -1;
    return 0;
}

int or_whir_verify(const or_whir_cfg_t *c, const or_p2_t *p16, const or_p2_t *p24, or_ts_t *ts,
                   const or_whir_commitment_t *cm, const ef_t *point, ef_t value) {
    int v = c->num_variables;
    int rc = -1;
    f_t root[8];
    ef_t rglob[64];
    int nglob = 0;
    cset_t sets[OR_WHIR_MAX_ROUNDS + 2];
    int nsets = 0;
    ef_t *pt = (ef_t *)malloc((size_t)(v + 1) * sizeof(ef_t));
    f_t *row = (f_t *)malloc(4096 * sizeof(f_t));
    f_t *path = (f_t *)malloc(64 * 8 * sizeof(f_t));
    uint32_t *qs = (uint32_t *)malloc(4096 * sizeof(uint32_t));
    memset(sets, 0, sizeof sets);
    memcpy(root, cm->root, sizeof root);
    ef_t oodp[8], ooda[8];
    int n_ood = cm->n_ood;
    memcpy(oodp, cm->ood_points, sizeof oodp);
    memcpy(ooda, cm->ood_answers, sizeof ooda);
    ef_t g = or_ts_sample_ef(ts), gp = ef_one(), sum = ef_zero();
    sets[0].offset = 0;
    sets[0].n = v;
    sets[0].count = n_ood + 1;
    sets[0].gamma = g;
    sets[0].pts = (ef_t *)malloc((size_t)(n_ood + 1) * v * sizeof(ef_t));
    for (int i = 0; i < n_ood; i++) {
        or_expand_from_univariate(oodp[i], v, sets[0].pts + (size_t)i * v);
        sum = ef_add(sum, ef_mul(gp, ooda[i]));
        gp = ef_mul(gp, g);
    }
    memcpy(sets[0].pts + (size_t)n_ood * v, point, (size_t)v * sizeof(ef_t));
    sum = ef_add(sum, ef_mul(gp, value));
    nsets = 1;
    if (vsc_rounds(ts, c->f0, c->starting_folding_pow_bits, &sum, rglob)) goto out;
    nglob = c->f0;
    int prev_fold = c->f0, prev_base = 1, prev_log_rows = v + c->starting_log_inv_rate - c->f0;
    f_t prev_root[8];
    memcpy(prev_root, root, sizeof root);
    int m = v - c->f0;
    for (int r = 0; r < c->n_rounds; r++) {
        const or_whir_round_t *rp = &c->rounds[r];
        int rs = r == 0 ? c->rs_reduction0 : 1;
        int k_next = or_whir_folding_at(c, r + 1);
        f_t nroot[8];
        if (or_ts_next_base(ts, nroot, 8)) goto out;
        for (int i = 0; i < rp->ood_samples; i++) oodp[i] = or_ts_sample_ef(ts);
        if (or_ts_next_ef(ts, ooda, (size_t)rp->ood_samples)) goto out;
        if (or_ts_check_pow(ts, rp->pow_bits)) goto out;
        int log_rows = rp->log_domain_size - prev_fold;
        if (log_rows != prev_log_rows) goto out;
        int nq = sample_queries(ts, rp->num_queries, log_rows, qs);
        int width = prev_base ? (1 << prev_fold) : 8 << prev_fold;
        ef_t *vals = (ef_t *)malloc((size_t)nq * sizeof(ef_t));
        for (int q = 0; q < nq; q++) {
            if (or_ts_next_hint_base(ts, row, (size_t)width)) { free(vals); goto out; }
            if (or_ts_next_hint_base(ts, path, (size_t)log_rows * 8)) { free(vals); goto out; }
            if (tree_verify(p16, p24, prev_root, log_rows, qs[q], row, width, path)) { free(vals); rc = -10; goto out; }
            vals[q] = eval_leaf_fold(row, width, prev_base, prev_fold, rglob + nglob - prev_fold);
        }
        ef_t gg = or_ts_sample_ef(ts), gpp = ef_one();
        cset_t *cs = &sets[nsets++];
        cs->offset = nglob;
        cs->n = m;
        cs->count = rp->ood_samples + nq;
        cs->gamma = gg;
        cs->pts = (ef_t *)malloc((size_t)cs->count * (m ? m : 1) * sizeof(ef_t));
        for (int i = 0; i < rp->ood_samples; i++) {
            or_expand_from_univariate(oodp[i], m, cs->pts + (size_t)i * m);
            sum = ef_add(sum, ef_mul(gpp, ooda[i]));
            gpp = ef_mul(gpp, gg);
        }
        f_t gen = f_two_adic_generator(log_rows);
        for (int q = 0; q < nq; q++) {
            or_expand_from_univariate(ef_from_f(f_pow(gen, qs[q])), m,
                                      cs->pts + (size_t)(rp->ood_samples + q) * m);
            sum = ef_add(sum, ef_mul(gpp, vals[q]));
            gpp = ef_mul(gpp, gg);
        }
        free(vals);
        if (vsc_rounds(ts, k_next, rp->folding_pow_bits, &sum, rglob + nglob)) goto out;
        nglob += k_next;
        m -= k_next;
        memcpy(prev_root, nroot, sizeof nroot);
        prev_log_rows = rp->log_domain_size - rs - k_next;
        prev_fold = k_next;
        prev_base = 0;
    }
    {
        int fv = c->final_sumcheck_rounds;
        if (fv != m) goto out;
        size_t FN = (size_t)1 << fv;
        ef_t *fin = (ef_t *)malloc(FN * sizeof(ef_t));
        if (or_ts_next_ef(ts, fin, FN)) { free(fin); goto out; }
        if (or_ts_check_pow(ts, c->final_pow_bits)) { free(fin); goto out; }
        int log_rows = c->final_log_domain_size - prev_fold;
        if (log_rows != prev_log_rows) { free(fin); goto out; }
        int nq = sample_queries(ts, c->final_queries, log_rows, qs);
        int width = prev_base ? (1 << prev_fold) : 8 << prev_fold;
        f_t gen = f_two_adic_generator(log_rows);
        for (int q = 0; q < nq; q++) {
            if (or_ts_next_hint_base(ts, row, (size_t)width) ||
                or_ts_next_hint_base(ts, path, (size_t)log_rows * 8)) { free(fin); goto out; }
            if (tree_verify(p16, p24, prev_root, log_rows, qs[q], row, width, path)) { free(fin); rc = -10; goto out; }
            ef_t fv_q = eval_leaf_fold(row, width, prev_base, prev_fold, rglob + nglob - prev_fold);
            or_expand_from_univariate(ef_from_f(f_pow(gen, qs[q])), fv, pt);
            if (!ef_eq(fv_q, or_mle_eval_ef(fin, fv, pt))) { free(fin); rc = -11; goto out; }
        }
        if (vsc_rounds(ts, fv, c->final_folding_pow_bits, &sum, rglob + nglob)) { free(fin); goto out; }
        nglob += fv;
        ef_t fval = or_mle_eval_ef(fin, fv, rglob + nglob - fv);
        free(fin);
        /* combined weight at the global folding point */
        ef_t W = ef_zero();
        for (int si = 0; si < nsets; si++) {
            cset_t *cs = &sets[si];
            ef_t gpow = ef_one();
            for (int k = 0; k < cs->count; k++) {
                ef_t e = or_eq_poly_outside(cs->pts + (size_t)k * cs->n, rglob + cs->offset, cs->n);
                W = ef_add(W, ef_mul(gpow, e));
                gpow = ef_mul(gpow, cs->gamma);
            }
        }
        if (!ef_eq(sum, ef_mul(fval, W))) { rc = -12; goto out; }
        if (nglob != v) goto out;
    }
    rc = 0;
out:
    for (int i = 0; i < nsets; i++) free(sets[i].pts);
    free(pt);
    free(row);
    free(path);
    free(qs);
    return rc;
}
