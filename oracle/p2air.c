/* p2air.c -- Poseidon2-16 KoalaBear AIR: constants, trace, constraints (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference proves Poseidon2Air<F, GenericPoseidon2LinearLayersKoalaBear,
 * 16, 3, 0, 4, 20> (src/examples/poseidon2.rs:33-36,108-146) from
 * p3-poseidon2-air @ d0c4a36 (parity unpinned, source absent).  Column layout
 * (Poseidon2Cols, SBOX_REGISTERS = 0 so S-boxes take no columns):
 *   [0]        export
 *   [1..17)    inputs[16]
 *   [17..81)   beginning_full_rounds[4].post[16]
 *   [81..101)  partial_rounds[20].post_sbox
 *   [101..165) ending_full_rounds[4].post[16]
 * Constraint order (fixes the alpha powers, crates/utils/src/constraints_folder.rs:51-56):
 *   export*(export-1); then per beginning full round 16 x (state_i - post_i);
 *   per partial round (s0 - post_sbox); per ending full round 16 x (state_i - post_i).
 * Only the local row is read; the "down" half of the point is unused.
 * Pinned by: every constraint vanishing on generated traces and the trace
 * rows reproducing the Poseidon2 permutation of the same constants.
 */
#include <string.h>
#include "oracle.h"

/* RoundConstants::from_rng: beginning[4][16], partial[20], ending[4][16] */
void or_air_consts_from_rng(or_air_consts_t *c, or_rng_t *rng) {
    for (int r = 0; r < OR_AIR_HALF_FULL; r++)
        for (int i = 0; i < 16; i++) c->beginning[r][i] = or_rng_sample_f(rng);
    for (int r = 0; r < OR_AIR_PARTIAL; r++) c->partial[r] = or_rng_sample_f(rng);
    for (int r = 0; r < OR_AIR_HALF_FULL; r++)
        for (int i = 0; i < 16; i++) c->ending[r][i] = or_rng_sample_f(rng);
}

static inline f_t cube(f_t x) { return f_mul(f_mul(x, x), x); }

/* generate_trace_rows (src/examples/poseidon2.rs:122-131), column-major output. */
void or_air_generate_trace(const or_air_consts_t *c, const f_t *inputs, int log_n, f_t *cols) {
    size_t N = (size_t)1 << log_n;
#pragma omp parallel for schedule(static)
    for (size_t row = 0; row < N; row++) {
        f_t s[16];
        int col = 0;
#define PUT(v) cols[(size_t)(col++) * N + row] = (v)
        PUT(1);
        for (int i = 0; i < 16; i++) { s[i] = inputs[row * 16 + i]; PUT(s[i]); }
        or_p2_mds_light(s, 16);
        for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
            for (int i = 0; i < 16; i++) s[i] = cube(f_add(s[i], c->beginning[r][i]));
            or_p2_mds_light(s, 16);
            for (int i = 0; i < 16; i++) PUT(s[i]);
        }
        for (int r = 0; r < OR_AIR_PARTIAL; r++) {
            s[0] = cube(f_add(s[0], c->partial[r]));
            PUT(s[0]);
            or_p2_internal_layer(s, 16);
        }
        for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
            for (int i = 0; i < 16; i++) s[i] = cube(f_add(s[i], c->ending[r][i]));
            or_p2_mds_light(s, 16);
            for (int i = 0; i < 16; i++) PUT(s[i]);
        }
#undef PUT
    }
}

/* ---- constraint evaluation over F (the packed base-field path,
 *      crates/sumcheck/src/prove.rs:273-306) ---- */
ef_t or_air_eval_f(const or_air_consts_t *c, const f_t *x, const ef_t *ap) {
    ef_t acc = ef_zero();
    int k = 0;
#define EMIT(v) acc = ef_add(acc, ef_mul_f(ap[k++], (v)))
    EMIT(f_mul(x[0], f_sub(x[0], 1)));
    f_t s[16];
    for (int i = 0; i < 16; i++) s[i] = x[1 + i];
    or_p2_mds_light(s, 16);
    int col = 17;
    for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
        for (int i = 0; i < 16; i++) s[i] = cube(f_add(s[i], c->beginning[r][i]));
        or_p2_mds_light(s, 16);
        for (int i = 0; i < 16; i++) { EMIT(f_sub(s[i], x[col + i])); s[i] = x[col + i]; }
        col += 16;
    }
    for (int r = 0; r < OR_AIR_PARTIAL; r++) {
        f_t t = cube(f_add(s[0], c->partial[r]));
        EMIT(f_sub(t, x[col]));
        s[0] = x[col++];
        or_p2_internal_layer(s, 16);
    }
    for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
        for (int i = 0; i < 16; i++) s[i] = cube(f_add(s[i], c->ending[r][i]));
        or_p2_mds_light(s, 16);
        for (int i = 0; i < 16; i++) { EMIT(f_sub(s[i], x[col + i])); s[i] = x[col + i]; }
        col += 16;
    }
#undef EMIT
    return acc;
}

/* ---- the same over EF (crates/sumcheck/src/prove.rs:307-318,
 *      crates/utils/src/constraints_folder.rs) ---- */
static void mds_light_ef(ef_t *s) {
    for (int c = 0; c < 16; c += 4) {
        ef_t *x = s + c;
        ef_t t01 = ef_add(x[0], x[1]), t23 = ef_add(x[2], x[3]);
        ef_t t0123 = ef_add(t01, t23);
        ef_t t01123 = ef_add(t0123, x[1]);
        ef_t t01233 = ef_add(t0123, x[3]);
        ef_t y3 = ef_add(t01233, ef_add(x[0], x[0]));
        ef_t y1 = ef_add(t01123, ef_add(x[2], x[2]));
        ef_t y0 = ef_add(t01123, t01);
        ef_t y2 = ef_add(t01233, t23);
        x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
    }
    ef_t sums[4] = {ef_zero(), ef_zero(), ef_zero(), ef_zero()};
    for (int i = 0; i < 16; i++) sums[i & 3] = ef_add(sums[i & 3], s[i]);
    for (int i = 0; i < 16; i++) s[i] = ef_add(s[i], sums[i & 3]);
}

static void internal_ef(ef_t *s) {
    const f_t *d = or_p2_internal_diag(16);
    ef_t sum = ef_zero();
    for (int i = 0; i < 16; i++) sum = ef_add(sum, s[i]);
    for (int i = 0; i < 16; i++) s[i] = ef_add(sum, ef_mul_f(s[i], d[i]));
}

static inline ef_t cube_ef(ef_t x) { return ef_mul(ef_square(x), x); }

ef_t or_air_eval_ef(const or_air_consts_t *c, const ef_t *x, const ef_t *ap) {
    ef_t acc = ef_zero();
    int k = 0;
#define EMIT(v) acc = ef_add(acc, ef_mul(ap[k++], (v)))
    EMIT(ef_mul(x[0], ef_sub(x[0], ef_one())));
    ef_t s[16];
    for (int i = 0; i < 16; i++) s[i] = x[1 + i];
    mds_light_ef(s);
    int col = 17;
    for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
        for (int i = 0; i < 16; i++) s[i] = cube_ef(ef_add(s[i], ef_from_f(c->beginning[r][i])));
        mds_light_ef(s);
        for (int i = 0; i < 16; i++) { EMIT(ef_sub(s[i], x[col + i])); s[i] = x[col + i]; }
        col += 16;
    }
    for (int r = 0; r < OR_AIR_PARTIAL; r++) {
        ef_t t = cube_ef(ef_add(s[0], ef_from_f(c->partial[r])));
        EMIT(ef_sub(t, x[col]));
        s[0] = x[col++];
        internal_ef(s);
    }
    for (int r = 0; r < OR_AIR_HALF_FULL; r++) {
        for (int i = 0; i < 16; i++) s[i] = cube_ef(ef_add(s[i], ef_from_f(c->ending[r][i])));
        mds_light_ef(s);
        for (int i = 0; i < 16; i++) { EMIT(ef_sub(s[i], x[col + i])); s[i] = x[col + i]; }
        col += 16;
    }
#undef EMIT
    return acc;
}
