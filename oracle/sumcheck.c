/* sumcheck.c -- generic sumcheck with univariate skip (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates crates/sumcheck/src/prove.rs:22-335 and verify.rs:23-113 of the
 * reference line for line (first-party code, so the round structure, the
 * evaluation points, the selector-eq product, the missing_mul_factor and the
 * fold order are the reference's own).  Folds bind the most-significant
 * variables (crates/utils/src/multilinear.rs:10-123).  Pinned by the verifier
 * round checks, brute-force round polynomials for small n and sum invariants.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

/* SumcheckGrinding::Auto (crates/sumcheck/src/lib.rs:21-31), EF::bits() = 8 * 31 */
static int sc_pow_bits(int sec_bits, int degree) {
    int ef_bits = 248;
    int t = ef_bits - (int)or_log2_up((size_t)degree + 1);
    if (t < 0) t = 0;
    int r = sec_bits - t;
    return r < 0 ? 0 : r;
}
/* SumcheckGrinding::pow_bits for each mode */
static int g_grind_kind = OR_GRIND_AUTO, g_grind_value = 128;
static int grind_bits(int degree) {
    if (g_grind_kind == OR_GRIND_NONE) return 0;
    if (g_grind_kind == OR_GRIND_CUSTOM) return g_grind_value;
    return sc_pow_bits(g_grind_value, degree);
}

/* ---------- constraint programs (straight-line SSA) ---------- */
int or_prog_check(const or_prog_t *p) {
    if (p->n_inputs < 0 || p->n_instr < 0) return -1;
    for (int j = 0; j < p->n_instr; j++) {
        const uint32_t *w = p->code + 4 * (size_t)j;
        uint32_t lim = (uint32_t)(p->n_inputs + j);
        int nops = (w[0] == OR_OP_ADD || w[0] == OR_OP_SUB || w[0] == OR_OP_MUL) ? 2 : w[0] == OR_OP_CONST ? 0 : 1;
        if (w[0] > OR_OP_ASSERT) return -1;
        if (w[0] == OR_OP_CONST && w[1] >= OR_P) return -1;
        for (int o = 0; o < nops; o++) {
            uint32_t v = w[1 + o];
            if (v >= lim) return -1;
            /* an ASSERT defines no value */
            if (v >= (uint32_t)p->n_inputs && p->code[4 * (size_t)(v - p->n_inputs)] == OR_OP_ASSERT) return -1;
        }
    }
    return 0;
}
int or_prog_n_asserts(const or_prog_t *p) {
    int k = 0;
    for (int j = 0; j < p->n_instr; j++) k += p->code[4 * (size_t)j] == OR_OP_ASSERT;
    return k;
}
ef_t or_prog_eval_f(const or_prog_t *p, const f_t *pt, const ef_t *batching) {
    f_t *v = (f_t *)malloc((size_t)(p->n_inputs + p->n_instr) * sizeof(f_t));
    memcpy(v, pt, (size_t)p->n_inputs * sizeof(f_t));
    ef_t acc = ef_zero();
    int k = 0;
    for (int j = 0; j < p->n_instr; j++) {
        const uint32_t *w = p->code + 4 * (size_t)j;
        f_t r = 0;
        switch (w[0]) {
        case OR_OP_CONST: r = w[1]; break;
        case OR_OP_ADD: r = f_add(v[w[1]], v[w[2]]); break;
        case OR_OP_SUB: r = f_sub(v[w[1]], v[w[2]]); break;
        case OR_OP_MUL: r = f_mul(v[w[1]], v[w[2]]); break;
        case OR_OP_NEG: r = f_neg(v[w[1]]); break;
        default: r = v[w[1]]; acc = ef_add(acc, ef_mul_f(batching[k++], r)); break;
        }
        v[p->n_inputs + j] = r;
    }
    free(v);
    return acc;
}
ef_t or_prog_eval_ef(const or_prog_t *p, const ef_t *pt, const ef_t *batching) {
    ef_t *v = (ef_t *)malloc((size_t)(p->n_inputs + p->n_instr) * sizeof(ef_t));
    memcpy(v, pt, (size_t)p->n_inputs * sizeof(ef_t));
    ef_t acc = ef_zero();
    int k = 0;
    for (int j = 0; j < p->n_instr; j++) {
        const uint32_t *w = p->code + 4 * (size_t)j;
        ef_t r;
        switch (w[0]) {
        case OR_OP_CONST: r = ef_from_f(w[1]); break;
        case OR_OP_ADD: r = ef_add(v[w[1]], v[w[2]]); break;
        case OR_OP_SUB: r = ef_sub(v[w[1]], v[w[2]]); break;
        case OR_OP_MUL: r = ef_mul(v[w[1]], v[w[2]]); break;
        case OR_OP_NEG: r = ef_neg(v[w[1]]); break;
        default: r = v[w[1]]; acc = ef_add(acc, ef_mul(batching[k++], r)); break;
        }
        v[p->n_inputs + j] = r;
    }
    free(v);
    return acc;
}

/* ---------- point evaluation (SumcheckComputation::eval) ---------- */
static ef_t comp_eval_f(const or_sc_comp_t *c, const f_t *pt) {
    switch (c->kind) {
    case OR_SC_AIR: return or_air_eval_f(c->air, pt, c->alpha_pows);
    case OR_SC_PROGRAM: return or_prog_eval_f(c->prog, pt, c->alpha_pows);
    case OR_SC_PRODUCT: return ef_from_f(f_mul(pt[0], pt[1]));
    default: return ef_from_f(f_mul(f_mul(pt[0], pt[0]), pt[1]));
    }
}
static ef_t comp_eval_ef(const or_sc_comp_t *c, const ef_t *pt) {
    switch (c->kind) {
    case OR_SC_AIR: return or_air_eval_ef(c->air, pt, c->alpha_pows);
    case OR_SC_PROGRAM: return or_prog_eval_ef(c->prog, pt, c->alpha_pows);
    case OR_SC_PRODUCT: return ef_mul(pt[0], pt[1]);
    default: return ef_mul(ef_square(pt[0]), pt[1]);
    }
}

/* compute_over_hypercube (crates/sumcheck/src/prove.rs:256-319) */
static ef_t over_hypercube(const or_sc_comp_t *c, int n_mles, int base, void **pols, size_t N,
                           const ef_t *eq) {
    ef_t total = ef_zero();
#pragma omp parallel
    {
        ef_t part = ef_zero();
        f_t *pf = (f_t *)malloc((size_t)n_mles * sizeof(f_t));
        ef_t *pe = (ef_t *)malloc((size_t)n_mles * sizeof(ef_t));
#pragma omp for schedule(static)
        for (size_t x = 0; x < N; x++) {
            ef_t r;
            if (base) {
                for (int m = 0; m < n_mles; m++) pf[m] = ((f_t *)pols[m])[x];
                r = comp_eval_f(c, pf);
            } else {
                for (int m = 0; m < n_mles; m++) pe[m] = ((ef_t *)pols[m])[x];
                r = comp_eval_ef(c, pe);
            }
            if (eq) r = ef_mul(r, eq[x]);
            part = ef_add(part, r);
        }
        free(pf);
        free(pe);
#pragma omp critical
        total = ef_add(total, part);
    }
    return total;
}

/* fold_multilinear_in_small_field: out[i] = sum_j s_j m[i + j*N/2^k] with F scalars */
static void *fold_small(int base, const void *m, size_t N, const f_t *s, int ns) {
    size_t M = N / (size_t)ns;
    if (base) {
        f_t *out = (f_t *)malloc(M * sizeof(f_t));
        const f_t *in = (const f_t *)m;
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < M; i++) {
            f_t acc = 0;
            for (int j = 0; j < ns; j++) acc = f_add(acc, f_mul(in[i + (size_t)j * M], s[j]));
            out[i] = acc;
        }
        return out;
    }
    ef_t *out = (ef_t *)malloc(M * sizeof(ef_t));
    const ef_t *in = (const ef_t *)m;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < M; i++) {
        ef_t acc = ef_zero();
        for (int j = 0; j < ns; j++) acc = ef_add(acc, ef_mul_f(in[i + (size_t)j * M], s[j]));
        out[i] = acc;
    }
    return out;
}

/* fold_multilinear_in_large_field: EF scalars; input F or EF, output EF */
static ef_t *fold_large(int base, const void *m, size_t N, const ef_t *s, int ns) {
    size_t M = N / (size_t)ns;
    ef_t *out = (ef_t *)malloc(M * sizeof(ef_t));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < M; i++) {
        ef_t acc = ef_zero();
        for (int j = 0; j < ns; j++) {
            ef_t t = base ? ef_mul_f(s[j], ((const f_t *)m)[i + (size_t)j * M])
                          : ef_mul(s[j], ((const ef_t *)m)[i + (size_t)j * M]);
            acc = ef_add(acc, t);
        }
        out[i] = acc;
    }
    return out;
}

/* sc_round (crates/sumcheck/src/prove.rs:97-254).  Returns folded EF mles. */
static ef_t **sc_round(int skips, int n_mles, int base, void **mles, int *n_vars,
                       const or_sc_comp_t *comp, const ef_t *eq_factor, int is_zerocheck,
                       or_ts_t *ts, int degree, ef_t *sum, int sec_bits, ef_t *challenges,
                       int round, ef_t *missing, int *has_missing) {
    size_t N = (size_t)1 << *n_vars;
    int K = 1 << skips;
    ef_t *eq_mle = NULL;
    if (eq_factor) {
        int ne = *n_vars - skips; /* eq_factor[1 + round ..] has this many entries */
        eq_mle = (ef_t *)malloc(((size_t)1 << ne) * sizeof(ef_t));
        or_eval_eq(eq_factor + 1 + round, ne, eq_mle);
    }
    int max_z = degree * (K - 1);
    int npts = max_z + 1;
    f_t *xs = (f_t *)malloc((size_t)npts * sizeof(f_t));
    ef_t *ys = (ef_t *)malloc((size_t)npts * sizeof(ef_t));
    int np = 0;
    int start = 0;
    if (is_zerocheck) {
        for (int i = 0; i < K; i++) { xs[np] = (f_t)i; ys[np] = ef_zero(); np++; }
        start = K;
    }
    f_t *sel = (f_t *)malloc((size_t)K * sizeof(f_t));
    for (int z = start; z <= max_z; z++) {
        ef_t sz;
        if (z == K - 1) {
            if (eq_factor) {
                ef_t r = eq_factor[round];
                if (skips == 1) {
                    sz = ef_mul(ef_sub(*sum, ef_mul(ys[0], ef_sub(ef_one(), r))), ef_inv(r));
                } else {
                    ef_t *se = (ef_t *)malloc((size_t)K * sizeof(ef_t));
                    or_selectors_at_ef(skips, r, se);
                    ef_t acc = ef_zero();
                    for (int i = 0; i < K - 1; i++) acc = ef_add(acc, ef_mul(ys[i], se[i]));
                    sz = ef_mul(ef_sub(*sum, acc), ef_inv(se[K - 1]));
                    free(se);
                }
            } else {
                ef_t acc = ef_zero();
                for (int i = 0; i < np; i++) acc = ef_add(acc, ys[i]);
                sz = ef_sub(*sum, acc);
            }
        } else {
            void **folded = (void **)malloc((size_t)n_mles * sizeof(void *));
            int owned = 1;
            if (skips == 1 && z == 0) {
                for (int m = 0; m < n_mles; m++) folded[m] = mles[m]; /* first half */
                owned = 0;
            } else {
                if (skips == 1) { sel[0] = f_sub(1, (f_t)z); sel[1] = (f_t)z; }
                else or_selectors_at_f(skips, (f_t)z, sel);
                for (int m = 0; m < n_mles; m++) folded[m] = fold_small(base, mles[m], N, sel, K);
            }
            sz = over_hypercube(comp, n_mles, base, folded, N / (size_t)K, eq_mle);
            if (*has_missing) sz = ef_mul(sz, *missing);
            if (owned) for (int m = 0; m < n_mles; m++) free(folded[m]);
            free(folded);
        }
        xs[np] = (f_t)z;
        ys[np] = sz;
        np++;
    }
    ef_t *p = (ef_t *)malloc((size_t)(np + K) * sizeof(ef_t));
    or_lagrange_interp(xs, ys, np, p);
    int plen = np;
    if (eq_factor) {
        ef_t r = eq_factor[round];
        ef_t *q;
        int qlen;
        if (skips == 1) {
            q = (ef_t *)malloc(2 * sizeof(ef_t));
            q[0] = ef_sub(ef_one(), r);
            q[1] = ef_sub(ef_add(r, r), ef_one());
            qlen = 2;
        } else {
            f_t *qx = (f_t *)malloc((size_t)K * sizeof(f_t));
            ef_t *qy = (ef_t *)malloc((size_t)K * sizeof(ef_t));
            or_selectors_at_ef(skips, r, qy);
            for (int i = 0; i < K; i++) qx[i] = (f_t)i;
            q = (ef_t *)malloc((size_t)K * sizeof(ef_t));
            or_lagrange_interp(qx, qy, K, q);
            qlen = K;
            free(qx);
            free(qy);
        }
        ef_t *pq = (ef_t *)malloc((size_t)(plen + qlen - 1) * sizeof(ef_t));
        or_poly_mul(p, plen, q, qlen, pq);
        free(p);
        free(q);
        p = pq;
        plen = plen + qlen - 1;
    }
    or_ts_add_ef(ts, p, (size_t)plen);
    ef_t c = or_ts_sample_ef(ts);
    challenges[round] = c;
    *sum = or_poly_eval(p, plen, c);
    *n_vars -= skips;
    (void)sec_bits;
    or_ts_pow(ts, grind_bits((degree + (eq_factor ? 1 : 0)) * (K - 1)));

    ef_t *fs = (ef_t *)malloc((size_t)K * sizeof(ef_t));
    if (skips == 1) { fs[0] = ef_sub(ef_one(), c); fs[1] = c; }
    else or_selectors_at_ef(skips, c, fs);
    if (eq_factor) {
        ef_t r = eq_factor[round], f;
        if (skips == 1) {
            f = ef_add(ef_mul(ef_sub(ef_one(), r), ef_sub(ef_one(), c)), ef_mul(r, c));
        } else {
            ef_t *sr = (ef_t *)malloc((size_t)K * sizeof(ef_t));
            or_selectors_at_ef(skips, r, sr);
            f = ef_zero();
            for (int i = 0; i < K; i++) f = ef_add(f, ef_mul(sr[i], fs[i]));
            free(sr);
        }
        *missing = *has_missing ? ef_mul(f, *missing) : f;
        *has_missing = 1;
    }
    ef_t **out = (ef_t **)malloc((size_t)n_mles * sizeof(ef_t *));
    for (int m = 0; m < n_mles; m++) out[m] = fold_large(base, mles[m], N, fs, K);
    free(fs);
    free(p);
    free(sel);
    free(xs);
    free(ys);
    free(eq_mle);
    return out;
}

void or_sc_prove_ex(int skips, int n_mles, int base_field, const void *const *mles, int n_vars,
                    const or_sc_comp_t *comp, int degree, const ef_t *eq_factor, int is_zerofier,
                    or_ts_t *ts, ef_t sum, int n_rounds, int grind_kind, int grind_value,
                    const ef_t *missing_mul_factor, ef_t *challenges, ef_t *folded, ef_t *final_sum) {
    if (n_rounds <= 0) n_rounds = n_vars - skips + 1;
    g_grind_kind = grind_kind;
    g_grind_value = grind_value;
    ef_t missing = missing_mul_factor ? *missing_mul_factor : ef_zero();
    int has_missing = missing_mul_factor != NULL;
    int nv = n_vars;
    ef_t **cur = sc_round(skips, n_mles, base_field, (void **)mles, &nv, comp, eq_factor,
                          is_zerofier, ts, degree, &sum, 0, challenges, 0, &missing,
                          &has_missing);
    for (int i = 1; i < n_rounds; i++) {
        ef_t **nxt = sc_round(1, n_mles, 0, (void **)cur, &nv, comp, eq_factor, 0, ts, degree,
                              &sum, 0, challenges, i, &missing, &has_missing);
        for (int m = 0; m < n_mles; m++) free(cur[m]);
        free(cur);
        cur = nxt;
    }
    size_t len = (size_t)1 << nv;
    for (int m = 0; m < n_mles; m++) {
        if (folded) memcpy(folded + (size_t)m * len, cur[m], len * sizeof(ef_t));
        free(cur[m]);
    }
    free(cur);
    *final_sum = sum;
    g_grind_kind = OR_GRIND_AUTO;
}

void or_sc_prove(int skips, int n_mles, int base_field, const void *const *mles, int n_vars,
                 const or_sc_comp_t *comp, int degree, const ef_t *eq_factor, int is_zerofier,
                 or_ts_t *ts, ef_t sum, int sec_bits, ef_t *challenges, ef_t *final_evals,
                 ef_t *final_sum) {
    or_sc_prove_ex(skips, n_mles, base_field, mles, n_vars, comp, degree, eq_factor, is_zerofier, ts,
                   sum, 0, OR_GRIND_AUTO, sec_bits, NULL, challenges, final_evals, final_sum);
}

/* ---------------- verifier (crates/sumcheck/src/verify.rs) ---------------- */
static int verify_core(or_ts_t *ts, int n, const int *degs, const int *set_sizes, int sec_bits,
                       ef_t *sum, ef_t *point, ef_t *value) {
    ef_t target = ef_zero();
    for (int r = 0; r < n; r++) {
        int d = degs[r];
        ef_t *co = (ef_t *)malloc((size_t)(d + 1) * sizeof(ef_t));
        if (or_ts_next_ef(ts, co, (size_t)d + 1)) { free(co); return -1; }
        ef_t cs = ef_zero();
        for (int s = 0; s < set_sizes[r]; s++) cs = ef_add(cs, or_poly_eval(co, d + 1, ef_from_f((f_t)s)));
        if (r == 0) *sum = cs;
        else if (!ef_eq(target, cs)) { free(co); return -2; }
        ef_t c = or_ts_sample_ef(ts);
        if (or_ts_check_pow(ts, sc_pow_bits(sec_bits, d))) { free(co); return -3; }
        target = or_poly_eval(co, d + 1, c);
        point[r] = c;
        free(co);
    }
    *value = target;
    return 0;
}

int or_sc_verify(or_ts_t *ts, int n_vars, int degree, int sec_bits, ef_t *sum, ef_t *point,
                 ef_t *value) {
    int *degs = (int *)malloc((size_t)(n_vars + 1) * sizeof(int));
    int *sets = (int *)malloc((size_t)(n_vars + 1) * sizeof(int));
    for (int i = 0; i < n_vars; i++) { degs[i] = degree; sets[i] = 2; }
    int r = verify_core(ts, n_vars, degs, sets, sec_bits, sum, point, value);
    free(degs);
    free(sets);
    return r;
}

int or_sc_verify_uskip(or_ts_t *ts, int degree, int n_vars, int skips, int sec_bits, ef_t *sum,
                       ef_t *point, ef_t *value) {
    int n = n_vars - skips + 1;
    int *degs = (int *)malloc((size_t)n * sizeof(int));
    int *sets = (int *)malloc((size_t)n * sizeof(int));
    degs[0] = degree * ((1 << skips) - 1);
    sets[0] = 1 << skips;
    for (int i = 1; i < n; i++) { degs[i] = degree; sets[i] = 2; }
    int r = verify_core(ts, n, degs, sets, sec_bits, sum, point, value);
    free(degs);
    free(sets);
    return r;
}
