// host_p16.h -- Poseidon2-16 for the host side of the transcript, on AVX2.
//
// The duplex challenger runs one Poseidon2-16 permutation per 8 observed or
// sampled elements on the host, between GPU round trips, so its latency adds
// directly to the proof time.  The state is held as two 8-lane vectors; a
// Montgomery product of 8 lanes takes six 32x32->64 multiplies (even and odd
// lanes), the external layer's 4x4 blocks are 128-bit-lane shuffles, and the
// internal layer multiplies by the Montgomery images of the diagonal.  Same
// results as kb::p2_permute<16, 20> (the proofs are compared bit for bit).
// The partial rounds' S-box is a scalar chain, so its cube takes one Montgomery
// reduction by 2^64 instead of two by 2^32 (cube64).
#pragma once
#include <stdint.h>
#include "common.h"
#if !defined(__HIP_DEVICE_COMPILE__)
#include <immintrin.h>

namespace wh {

struct HostP16 {
    static __m256i vp() { return _mm256_set1_epi32((int)kb::P); }
    static __m256i vadd(__m256i a, __m256i b) {
        __m256i s = _mm256_add_epi32(a, b);
        return _mm256_min_epu32(s, _mm256_sub_epi32(s, vp()));
    }
    // Montgomery product of 8 lanes (inputs < P): hi(a b) - hi(m P), m = lo(a b) MU
    static __m256i vmul(__m256i a, __m256i b) {
        const __m256i P = vp(), MU = _mm256_set1_epi32((int)kb::MU);
        __m256i pe = _mm256_mul_epu32(a, b);
        __m256i po = _mm256_mul_epu32(_mm256_srli_epi64(a, 32), _mm256_srli_epi64(b, 32));
        __m256i qe = _mm256_mul_epu32(_mm256_mul_epu32(pe, MU), P);
        __m256i qo = _mm256_mul_epu32(_mm256_mul_epu32(po, MU), P);
        __m256i ph = _mm256_blend_epi32(_mm256_srli_epi64(pe, 32), po, 0xAA);
        __m256i qh = _mm256_blend_epi32(_mm256_srli_epi64(qe, 32), qo, 0xAA);
        __m256i d = _mm256_sub_epi32(ph, qh);
        return _mm256_min_epu32(d, _mm256_add_epi32(d, P));
    }
    static __m256i vcube(__m256i x) { return vmul(vmul(x, x), x); }
    // M4 on each 4-element block: row r = block sum + x_r + 2 x_(r+1 mod 4)
    static __m256i mat4(__m256i x) {
        __m256i s = vadd(x, _mm256_shuffle_epi32(x, 0xB1));
        s = vadd(s, _mm256_shuffle_epi32(s, 0x4E));
        __m256i nx = _mm256_shuffle_epi32(x, 0x39);
        return vadd(vadd(s, x), vadd(nx, nx));
    }
    static void mds_light(__m256i &lo, __m256i &hi) {
        lo = mat4(lo);
        hi = mat4(hi);
        __m256i u = vadd(lo, hi);
        u = vadd(u, _mm256_permute2x128_si256(u, u, 0x01));  // sum of the 4 blocks per residue
        lo = vadd(lo, u);
        hi = vadd(hi, u);
    }
    static __m256i ld(const kb::F *p) { return _mm256_loadu_si256((const __m256i *)p); }

    static constexpr uint64_t pinv64() {  // P^-1 mod 2^64 by Newton's iteration (3 -> 96 bits)
        uint64_t x = kb::P;
        for (int i = 0; i < 5; i++) x *= 2 - (uint64_t)kb::P * x;
        return x;
    }
    // Montgomery cube of t < P: t^3 < 2^93 < 2^64 P, so one reduction by 2^64 gives
    // t^3 / 2^64 mod P, which for t = a 2^32 is (a^3) 2^32, the Montgomery image of
    // a^3.  With m = lo64(T) P^-1, T - m P is a multiple of 2^64 and
    // hi(T) - hi(m P) lies in (-P, 2^29): one conditional add of P.
    static uint32_t cube64(uint32_t t) {
        const uint64_t t2 = (uint64_t)t * t;
        const unsigned __int128 T = (unsigned __int128)t2 * t;
        const uint64_t m = (uint64_t)T * pinv64();
        const uint64_t mh = (uint64_t)(((unsigned __int128)m * kb::P) >> 64);
        const uint32_t d = (uint32_t)((uint64_t)(T >> 64) - mh);
        return kb::umin(d, d + kb::P);
    }

    static void permute(kb::F *s, const kb::P2Consts16 &k) {
        __m256i lo = ld(s), hi = ld(s + 8);
        const __m256i dlo = _mm256_loadu_si256((const __m256i *)kb::DIAG.d16);
        const __m256i dhi = _mm256_loadu_si256((const __m256i *)(kb::DIAG.d16 + 8));
        mds_light(lo, hi);
        for (int r = 0; r < 4; r++) {
            lo = vcube(vadd(lo, ld(k.ext_init[r])));
            hi = vcube(vadd(hi, ld(k.ext_init[r] + 8)));
            mds_light(lo, hi);
        }
        // Partial rounds with lane 0 kept as a scalar: with x the state entering a
        // round, y = (x0 + rc)^3, S = sum_{i>=1} x_i and s = S + y, the next state is
        // x0' = s + d0 y = S - y (d0 = -2) and x_i' = s + d_i x_i (i >= 1), and
        // S' = 15 s + sum_{i>=1} d_i x_i.  The vector products d_i x_i and their sum
        // depend only on x, so they overlap the scalar S-box chain, whose step is
        // one subtraction (from S + rc, formed while the cube runs) and the cube.
        const __m256i lane0 = _mm256_setr_epi32(-1, 0, 0, 0, 0, 0, 0, 0);
        auto hsum_rest = [&](__m256i a, __m256i b) {  // sum of lanes 1..15 of (a, b)
            __m256i u = vadd(_mm256_andnot_si256(lane0, a), b);
            u = vadd(u, _mm256_permute2x128_si256(u, u, 0x01));
            u = vadd(u, _mm256_shuffle_epi32(u, 0x4E));
            u = vadd(u, _mm256_shuffle_epi32(u, 0xB1));
            return kb::F{(uint32_t)_mm256_cvtsi256_si32(u)};
        };
        const kb::F fifteen = kb::to_mont(15);
        kb::F x0{(uint32_t)_mm256_cvtsi256_si32(lo)};
        kb::F rest = hsum_rest(lo, hi);
        __m256i vlo = vmul(lo, dlo), vhi = vmul(hi, dhi);
        kb::F w = hsum_rest(vlo, vhi);
        kb::F t = kb::fadd(x0, k.internal[0]);  // the S-box input x0 + rc
        for (int r = 0; r < 20; r++) {
            const kb::F y{cube64(t.v)};
            const kb::F sr = kb::fadd(rest, y);
            if (r + 1 < 20) t = kb::fsub(kb::fadd(rest, k.internal[r + 1]), y);
            else x0 = kb::fsub(rest, y);
            rest = kb::fadd(kb::fmul(fifteen, sr), w);
            const __m256i sv = _mm256_set1_epi32((int)sr.v);
            lo = vadd(sv, vlo);
            hi = vadd(sv, vhi);
            if (r + 1 < 20) {
                vlo = vmul(lo, dlo);
                vhi = vmul(hi, dhi);
                w = hsum_rest(vlo, vhi);
            }
        }
        lo = _mm256_blend_epi32(lo, _mm256_set1_epi32((int)x0.v), 0x01);
        for (int r = 0; r < 4; r++) {
            lo = vcube(vadd(lo, ld(k.ext_term[r])));
            hi = vcube(vadd(hi, ld(k.ext_term[r] + 8)));
            mds_light(lo, hi);
        }
        _mm256_storeu_si256((__m256i *)s, lo);
        _mm256_storeu_si256((__m256i *)(s + 8), hi);
    }
};
static_assert((uint64_t)kb::P * HostP16::pinv64() == 1, "P^-1 mod 2^64");

}  // namespace wh
#endif
