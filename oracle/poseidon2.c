/* poseidon2.c -- Poseidon2 over KoalaBear and the Merkle hash/compress (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference instantiates Poseidon2KoalaBear<16> and <24> from the seeded
 * rng (src/examples/poseidon2.rs:22-26,148-151).  Structure restated from
 * p3-poseidon2 / p3-koala-bear @ d0c4a36 (parity unpinned, sources absent):
 *   initial:  mds_light; 4 x (add rc, x^3 on all lanes, mds_light)
 *   internal: rounds_p x (s0 += rc, s0 = s0^3, s = sum(s) + diag * s)
 *   terminal: 4 x (add rc, x^3 on all lanes, mds_light)
 * mds_light = M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] per 4-chunk, then
 * each lane adds the sum of its residue class mod 4.  Constants drawn as
 * new_from_rng_128: 4 x [F;W] initial, 4 x [F;W] terminal, rounds_p F.
 * Pinned by: invertibility and dense-matrix restatements of both linear
 * layers, x^3 being a permutation (gcd(3, p-1) = 1), forward/inverse round
 * trip in the tests.
 */
#include <string.h>
#include "oracle.h"

static f_t INTERNAL16[16], INTERNAL24[24];
static int diag_ready = 0;

/* diag vectors (canonical):
 * 16: [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24]
 * 24: [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/4, 1/8, 1/16, 1/32, 1/64, 1/2^24,
 *      -1/2^8, -1/8, -1/16, -1/32, -1/64, -1/2^7, -1/2^9, -1/2^24] */
static f_t inv2k(int k) { return f_inv(f_pow(2, (uint64_t)k)); }
static void init_diag(void) {
    if (diag_ready) return;
    f_t d16[16] = {f_neg(2), 1, 2, inv2k(1), 3, 4, f_neg(inv2k(1)), f_neg(3), f_neg(4),
                   inv2k(8), inv2k(3), inv2k(24), f_neg(inv2k(8)), f_neg(inv2k(3)),
                   f_neg(inv2k(4)), f_neg(inv2k(24))};
    f_t d24[24] = {f_neg(2), 1, 2, inv2k(1), 3, 4, f_neg(inv2k(1)), f_neg(3), f_neg(4),
                   inv2k(8), inv2k(2), inv2k(3), inv2k(4), inv2k(5), inv2k(6), inv2k(24),
                   f_neg(inv2k(8)), f_neg(inv2k(3)), f_neg(inv2k(4)), f_neg(inv2k(5)),
                   f_neg(inv2k(6)), f_neg(inv2k(7)), f_neg(inv2k(9)), f_neg(inv2k(24))};
    memcpy(INTERNAL16, d16, sizeof d16);
    memcpy(INTERNAL24, d24, sizeof d24);
    diag_ready = 1;
}

const f_t *or_p2_internal_diag(int width) {
    init_diag();
    return width == 16 ? INTERNAL16 : INTERNAL24;
}

void or_p2_init_from_rng(or_p2_t *p, int width, or_rng_t *rng) {
    memset(p, 0, sizeof *p);
    p->width = width;
    p->rounds_f = 8;
    p->rounds_p = width == 16 ? 20 : 23;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < width; i++) p->ext_init[r][i] = or_rng_sample_f(rng);
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < width; i++) p->ext_term[r][i] = or_rng_sample_f(rng);
    for (int r = 0; r < p->rounds_p; r++) p->internal[r] = or_rng_sample_f(rng);
}

static void mat4(f_t *x) {
    f_t t01 = f_add(x[0], x[1]), t23 = f_add(x[2], x[3]);
    f_t t0123 = f_add(t01, t23);
    f_t t01123 = f_add(t0123, x[1]);
    f_t t01233 = f_add(t0123, x[3]);
    f_t y3 = f_add(t01233, f_add(x[0], x[0]));
    f_t y1 = f_add(t01123, f_add(x[2], x[2]));
    f_t y0 = f_add(t01123, t01);
    f_t y2 = f_add(t01233, t23);
    x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
}

void or_p2_mds_light(f_t *s, int width) {
    for (int c = 0; c < width; c += 4) mat4(s + c);
    f_t sums[4] = {0, 0, 0, 0};
    for (int i = 0; i < width; i++) sums[i & 3] = f_add(sums[i & 3], s[i]);
    for (int i = 0; i < width; i++) s[i] = f_add(s[i], sums[i & 3]);
}

void or_p2_internal_layer(f_t *s, int width) {
    const f_t *d = or_p2_internal_diag(width);
    f_t sum = 0;
    for (int i = 0; i < width; i++) sum = f_add(sum, s[i]);
    for (int i = 0; i < width; i++) s[i] = f_add(sum, f_mul(d[i], s[i]));
}

static inline f_t cube(f_t x) { return f_mul(f_mul(x, x), x); }

void or_p2_permute(const or_p2_t *p, f_t *s) {
    int w = p->width;
    or_p2_mds_light(s, w);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < w; i++) s[i] = cube(f_add(s[i], p->ext_init[r][i]));
        or_p2_mds_light(s, w);
    }
    for (int r = 0; r < p->rounds_p; r++) {
        s[0] = cube(f_add(s[0], p->internal[r]));
        or_p2_internal_layer(s, w);
    }
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < w; i++) s[i] = cube(f_add(s[i], p->ext_term[r][i]));
        or_p2_mds_light(s, w);
    }
}

/* PaddingFreeSponge<P24, 24, 16, 8>: overwrite-mode absorption of 16-element
 * chunks (a short last chunk overwrites only its prefix), squeeze 8. */
void or_hash_leaf(const or_p2_t *p24, const f_t *row, size_t len, f_t out[8]) {
    f_t s[24];
    memset(s, 0, sizeof s);
    for (size_t off = 0; off < len; off += 16) {
        size_t n = len - off < 16 ? len - off : 16;
        memcpy(s, row + off, n * sizeof(f_t));
        or_p2_permute(p24, s);
    }
    memcpy(out, s, 8 * sizeof(f_t));
}

/* TruncatedPermutation<P16, 2, 8, 16>: permute(l || r)[..8]. */
void or_compress(const or_p2_t *p16, const f_t l[8], const f_t r[8], f_t out[8]) {
    f_t s[16];
    memcpy(s, l, 8 * sizeof(f_t));
    memcpy(s + 8, r, 8 * sizeof(f_t));
    or_p2_permute(p16, s);
    memcpy(out, s, 8 * sizeof(f_t));
}
