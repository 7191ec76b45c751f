/* poly.c -- multilinear and univariate helpers (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Multilinear tables are big-endian: entry i, variable k <-> bit (n-1-k) of i
 * (docs/main.tex:81-82; whir-p3 EvaluationsList::eval_eq).  Pinned by brute
 * force over the hypercube in the tests.  Univariate polynomials are dense
 * coefficient vectors, low degree first (whir-p3 WhirDensePolynomial);
 * interpolation results are fixed by the mathematics (unique polynomial of
 * degree < #points), pinned by evaluation round trips.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* crates/utils/src/misc.rs:1-8 */
size_t or_log2_up(size_t x) {
    if (x <= 1) return 0;
    size_t r = 0, v = x - 1;
    while (v) { r++; v >>= 1; }
    return r;
}

/* Levels of 2^12 entries and more read the previous level from a copy so that
 * their entries can be computed in parallel; smaller ones double in place, from
 * the top entry down.  Both give the same table. */
#define OR_PAR_MIN 4096
void or_eval_eq_scaled(const ef_t *p, int n, ef_t scale, ef_t *out) {
    out[0] = scale;
    ef_t *prev = NULL;
    for (int k = 0; k < n; k++) {
        size_t half = (size_t)1 << k;
        if (half < OR_PAR_MIN) {
            for (size_t j = half; j-- > 0;) {
                ef_t hi = ef_mul(out[j], p[k]);
                out[2 * j] = ef_sub(out[j], hi);
                out[2 * j + 1] = hi;
            }
            continue;
        }
        prev = (ef_t *)realloc(prev, half * sizeof(ef_t));
        memcpy(prev, out, half * sizeof(ef_t));
#pragma omp parallel for schedule(static)
        for (size_t j = 0; j < half; j++) {
            ef_t hi = ef_mul(prev[j], p[k]);
            out[2 * j] = ef_sub(prev[j], hi);
            out[2 * j + 1] = hi;
        }
    }
    free(prev);
}

void or_eval_eq(const ef_t *p, int n, ef_t *out) { or_eval_eq_scaled(p, n, ef_one(), out); }

void or_eval_eq_f(const f_t *p, int n, f_t *out) {
    out[0] = 1;
    f_t *prev = NULL;
    for (int k = 0; k < n; k++) {
        size_t half = (size_t)1 << k;
        if (half < OR_PAR_MIN) {
            for (size_t j = half; j-- > 0;) {
                f_t hi = f_mul(out[j], p[k]);
                out[2 * j] = f_sub(out[j], hi);
                out[2 * j + 1] = hi;
            }
            continue;
        }
        prev = (f_t *)realloc(prev, half * sizeof(f_t));
        memcpy(prev, out, half * sizeof(f_t));
#pragma omp parallel for schedule(static)
        for (size_t j = 0; j < half; j++) {
            f_t hi = f_mul(prev[j], p[k]);
            out[2 * j] = f_sub(prev[j], hi);
            out[2 * j + 1] = hi;
        }
    }
    free(prev);
}

/* MultilinearPoint::eq_poly_outside: prod_i (a_i b_i + (1-a_i)(1-b_i)) */
ef_t or_eq_poly_outside(const ef_t *a, const ef_t *b, int n) {
    ef_t acc = ef_one(), one = ef_one();
    for (int i = 0; i < n; i++) {
        ef_t t = ef_add(ef_mul(a[i], b[i]), ef_mul(ef_sub(one, a[i]), ef_sub(one, b[i])));
        acc = ef_mul(acc, t);
    }
    return acc;
}

/* EvaluationsList::evaluate: sequential MSB-first binding. */
ef_t or_mle_eval_ef(const ef_t *evals, int n, const ef_t *point) {
    size_t N = (size_t)1 << n;
    ef_t *t = (ef_t *)malloc(N * sizeof(ef_t));
    memcpy(t, evals, N * sizeof(ef_t));
    for (int k = 0; k < n; k++) {
        size_t h = N >> (k + 1);
#pragma omp parallel for schedule(static) if (h >= OR_PAR_MIN)
        for (size_t i = 0; i < h; i++) t[i] = ef_add(t[i], ef_mul(point[k], ef_sub(t[i + h], t[i])));
    }
    ef_t r = t[0];
    free(t);
    return r;
}

ef_t or_mle_eval_f(const f_t *evals, int n, const ef_t *point) {
    size_t N = (size_t)1 << n;
    if (n == 0) return ef_from_f(evals[0]);
    size_t h = N >> 1;
    ef_t *t = (ef_t *)malloc(h * sizeof(ef_t));
#pragma omp parallel for schedule(static) if (h >= OR_PAR_MIN)
    for (size_t i = 0; i < h; i++)
        t[i] = ef_add(ef_from_f(evals[i]), ef_mul_f(point[0], f_sub(evals[i + h], evals[i])));
    for (int k = 1; k < n; k++) {
        size_t hh = N >> (k + 1);
#pragma omp parallel for schedule(static) if (hh >= OR_PAR_MIN)
        for (size_t i = 0; i < hh; i++) t[i] = ef_add(t[i], ef_mul(point[k], ef_sub(t[i + hh], t[i])));
    }
    ef_t r = t[0];
    free(t);
    return r;
}

/* MultilinearPoint::expand_from_univariate: (x^(2^(n-1)), ..., x^2, x) */
void or_expand_from_univariate(ef_t x, int n, ef_t *out) {
    ef_t cur = x;
    for (int i = 0; i < n; i++) {
        out[n - 1 - i] = cur;
        cur = ef_square(cur);
    }
}

void or_lagrange_interp(const f_t *xs, const ef_t *ys, int n, ef_t *coeffs) {
    /* master polynomial M = prod (X - x_i), over F */
    f_t *m = (f_t *)calloc((size_t)n + 1, sizeof(f_t));
    f_t *q = (f_t *)calloc((size_t)n, sizeof(f_t));
    m[0] = 1;
    for (int i = 0; i < n; i++) {
        for (int d = i + 1; d > 0; d--) m[d] = f_sub(m[d - 1], f_mul(xs[i], m[d]));
        m[0] = f_neg(f_mul(xs[i], m[0]));
    }
    for (int k = 0; k < n; k++) coeffs[k] = ef_zero();
    for (int i = 0; i < n; i++) {
        /* q = M / (X - x_i) by synthetic division */
        f_t carry = m[n];
        for (int d = n - 1; d >= 0; d--) {
            q[d] = carry;
            carry = f_add(m[d], f_mul(carry, xs[i]));
        }
        f_t den = 1;
        for (int j = 0; j < n; j++)
            if (j != i) den = f_mul(den, f_sub(xs[i], xs[j]));
        ef_t s = ef_mul_f(ys[i], f_inv(den));
        for (int d = 0; d < n; d++) coeffs[d] = ef_add(coeffs[d], ef_mul_f(s, q[d]));
    }
    free(m);
    free(q);
}

ef_t or_poly_eval(const ef_t *c, int n, ef_t x) {
    ef_t acc = ef_zero();
    for (int i = n - 1; i >= 0; i--) acc = ef_add(ef_mul(acc, x), c[i]);
    return acc;
}

void or_poly_mul(const ef_t *a, int na, const ef_t *b, int nb, ef_t *out) {
    for (int i = 0; i < na + nb - 1; i++) out[i] = ef_zero();
    for (int i = 0; i < na; i++)
        for (int j = 0; j < nb; j++) out[i + j] = ef_add(out[i + j], ef_mul(a[i], b[j]));
}

/* crates/utils/src/univariate.rs:5-15: Lagrange basis over {0..2^k-1}, evaluated at x. */
void or_selectors_at_f(int k, f_t x, f_t *out) {
    int N = 1 << k;
    if (x < (f_t)N) {
        for (int i = 0; i < N; i++) out[i] = (f_t)(i == (int)x);
        return;
    }
    for (int i = 0; i < N; i++) {
        f_t num = 1, den = 1;
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            num = f_mul(num, f_sub(x, (f_t)j));
            den = f_mul(den, f_sub((f_t)i, (f_t)j));
        }
        out[i] = f_mul(num, f_inv(den));
    }
}

void or_selectors_at_ef(int k, ef_t x, ef_t *out) {
    int N = 1 << k;
    for (int i = 0; i < N; i++) {
        ef_t num = ef_one();
        f_t den = 1;
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            num = ef_mul(num, ef_sub(x, ef_from_f((f_t)j)));
            den = f_mul(den, f_sub((f_t)i, (f_t)j));
        }
        out[i] = ef_mul_f(num, f_inv(den));
    }
}
