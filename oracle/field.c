/* field.c -- KoalaBear F and its degree-8 binomial extension (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * F is Plonky3's KoalaBear (p3-koala-bear @ d0c4a36, MontyField31 with
 * p = 2^31 - 2^24 + 1); the reference names it at src/examples/poseidon2.rs:30.
 * EF = BinomialExtensionField<F, 8> (src/examples/poseidon2.rs:31) with W = 3.
 * The oracle keeps canonical values; Montgomery form only matters at the RNG
 * boundary (f_from_monty).  Pinned by: Python big-int arithmetic, Fermat
 * inverse, irreducibility of X^8 - 3 (sympy), the two-adic generator table
 * of order 2^24.
 */
#include <stdlib.h>
#include "oracle.h"

f_t f_pow(f_t a, uint64_t e) {
    f_t r = 1;
    while (e) {
        if (e & 1) r = f_mul(r, a);
        a = f_mul(a, a);
        e >>= 1;
    }
    return r;
}

f_t f_inv(f_t a) { return f_pow(a, OR_P - 2); }

/* MontyField31::new_monty(m) holds m as Montgomery form: canonical = m * R^-1. */
f_t f_from_monty(uint32_t m) {
    static f_t rinv = 0;
    if (!rinv) rinv = f_inv((f_t)(((uint64_t)1 << 32) % OR_P));
    return f_mul(m % OR_P, rinv);
}

/* Plonky3 KoalaBear: multiplicative generator 3, TWO_ADICITY 24; the 2^bits
 * generator is 3^((p-1)/2^bits). */
f_t f_two_adic_generator(int bits) {
    /* Plonky3 asserts bits <= TWO_ADICITY: no larger power-of-two subgroup exists */
    if (bits < 0 || bits > 24) abort();
    f_t g = f_pow(3, (OR_P - 1) >> 24);
    for (int i = bits; i < 24; i++) g = f_mul(g, g);
    return g;
}

ef_t ef_zero(void) { ef_t r = {{0}}; return r; }
ef_t ef_one(void) { ef_t r = {{1}}; return r; }
ef_t ef_from_f(f_t a) { ef_t r = {{0}}; r.c[0] = a; return r; }

ef_t ef_add(ef_t a, ef_t b) {
    for (int i = 0; i < OR_EF_D; i++) a.c[i] = f_add(a.c[i], b.c[i]);
    return a;
}
ef_t ef_sub(ef_t a, ef_t b) {
    for (int i = 0; i < OR_EF_D; i++) a.c[i] = f_sub(a.c[i], b.c[i]);
    return a;
}
ef_t ef_neg(ef_t a) {
    for (int i = 0; i < OR_EF_D; i++) a.c[i] = f_neg(a.c[i]);
    return a;
}

/* Schoolbook product reduced by X^8 = W. */
/* Schoolbook product, X^8 = W.  Each product of canonical values is below
 * p^2 < 2^62, so a 64-bit sum holds four of them: the coefficient sums are kept in
 * two halves by the first factor's index (i < 4, i >= 4), each reduced once. */
ef_t ef_mul(ef_t a, ef_t b) {
    uint64_t lo[2][8] = {{0}}, hi[2][8] = {{0}};
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            uint64_t p = (uint64_t)a.c[i] * b.c[j];
            if (i + j < 8) lo[i >> 2][i + j] += p; else hi[i >> 2][i + j - 8] += p;
        }
    ef_t r;
    for (int k = 0; k < 8; k++) {
        f_t l = f_add((f_t)(lo[0][k] % OR_P), (f_t)(lo[1][k] % OR_P));
        f_t h = f_add((f_t)(hi[0][k] % OR_P), (f_t)(hi[1][k] % OR_P));
        r.c[k] = f_add(l, f_mul(h, OR_EF_W));
    }
    return r;
}

ef_t ef_mul_f(ef_t a, f_t b) {
    for (int i = 0; i < OR_EF_D; i++) a.c[i] = f_mul(a.c[i], b);
    return a;
}

ef_t ef_square(ef_t a) { return ef_mul(a, a); }

ef_t ef_pow(ef_t a, uint64_t e) {
    ef_t r = ef_one();
    while (e) {
        if (e & 1) r = ef_mul(r, a);
        a = ef_mul(a, a);
        e >>= 1;
    }
    return r;
}

/* Frobenius x -> x^p acts on X^k as multiplication by (W^((p-1)/8))^k since
 * X^p = X * (X^8)^((p-1)/8).  Inverse via the norm: a^-1 = a^(r-1) / N(a),
 * r = 1 + p + ... + p^7, a^(r-1) = prod_{i=1..7} frob^i(a). */
static ef_t ef_frob(ef_t a, int times) {
    f_t z = f_pow(OR_EF_W, (OR_P - 1) / 8);
    f_t zt = f_pow(z, (uint64_t)times);
    f_t m = 1;
    for (int k = 0; k < 8; k++) {
        a.c[k] = f_mul(a.c[k], m);
        m = f_mul(m, zt);
    }
    return a;
}

ef_t ef_inv(ef_t a) {
    ef_t acc = ef_one();
    for (int i = 1; i < 8; i++) acc = ef_mul(acc, ef_frob(a, i));
    ef_t n = ef_mul(acc, a); /* in F */
    return ef_mul_f(acc, f_inv(n.c[0]));
}

int ef_eq(ef_t a, ef_t b) {
    for (int i = 0; i < 8; i++) if (a.c[i] != b.c[i]) return 0;
    return 1;
}
int ef_is_zero(ef_t a) {
    for (int i = 0; i < 8; i++) if (a.c[i]) return 0;
    return 1;
}
