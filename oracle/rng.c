/* rng.c -- StdRng (ChaCha12) restatement for synthetic inputs (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference draws every synthetic value from StdRng::seed_from_u64(0)
 * (src/examples/poseidon2.rs:104-120,148-149).  rand 0.9.1 StdRng is
 * rand_chacha 0.9.0 ChaCha12Rng (64-bit block counter in words 12-13, stream
 * id in words 14-15, 4 blocks buffered), seeded via rand_core 0.9.3's
 * seed_from_u64 which expands the u64 with PCG32.  F is sampled as in
 * p3-monty-31: next_u32() >> 1, rejected when >= p, read as Montgomery form.
 * Parity unpinned for the seeding and sampling (dependency sources absent);
 * the ChaCha block function is pinned by RFC 7539 vectors in the tests.
 */
#include <string.h>
#include "oracle.h"

#define ROTL(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define QR(a, b, c, d)                                        \
    a += b; d ^= a; d = ROTL(d, 16);                          \
    c += d; b ^= c; b = ROTL(b, 12);                          \
    a += b; d ^= a; d = ROTL(d, 8);                           \
    c += d; b ^= c; b = ROTL(b, 7);

static void chacha_core(const uint32_t in[16], int rounds, uint32_t out[16]) {
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int i = 0; i < rounds; i += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

static const uint32_t SIGMA[4] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                     uint32_t out[16]) {
    uint32_t in[16];
    memcpy(in, SIGMA, 16);
    memcpy(in + 4, key, 32);
    in[12] = (uint32_t)counter;
    in[13] = (uint32_t)(counter >> 32);
    in[14] = (uint32_t)stream;
    in[15] = (uint32_t)(stream >> 32);
    chacha_core(in, rounds, out);
}

void or_chacha_block_ietf(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3],
                          int rounds, uint32_t out[16]) {
    uint32_t in[16];
    memcpy(in, SIGMA, 16);
    memcpy(in + 4, key, 32);
    in[12] = counter;
    in[13] = nonce[0];
    in[14] = nonce[1];
    in[15] = nonce[2];
    chacha_core(in, rounds, out);
}

static uint32_t pcg32(uint64_t *state) {
    const uint64_t MUL = 6364136223846793005ull;
    const uint64_t INC = 11634580027462260723ull;
    *state = *state * MUL + INC;
    uint64_t s = *state;
    uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t rot = (uint32_t)(s >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

void or_rng_seed_u64(or_rng_t *r, uint64_t seed) {
    memset(r, 0, sizeof *r);
    for (int i = 0; i < 8; i++) r->key[i] = pcg32(&seed); /* little-endian bytes -> word */
    r->counter = 0;
    r->idx = 64;
    r->rounds = 12;
}

uint32_t or_rng_next_u32(or_rng_t *r) {
    if (r->idx >= 64) {
        for (int b = 0; b < 4; b++) or_chacha_block(r->key, r->counter + b, 0, r->rounds, r->buf + 16 * b);
        r->counter += 4;
        r->idx = 0;
    }
    return r->buf[r->idx++];
}

f_t or_rng_sample_f(or_rng_t *r) {
    for (;;) {
        uint32_t u = or_rng_next_u32(r) >> 1;
        if (u < OR_P) return f_from_monty(u);
    }
}
