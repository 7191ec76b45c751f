// common.h -- device helpers shared by the prover kernels: compile-time
// Poseidon2 diagonals, wave/block reductions of extension-field values and
// the host-side launch/check macros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include "kb.h"
#include "reduce.h"

namespace kb {

constexpr uint32_t cmont(uint64_t c) { return (uint32_t)((c << 32) % P); }
constexpr uint64_t cinv2k(int k) {
    uint64_t x = 1;
    for (int i = 0; i < k; i++) x = (x & 1) ? (x + P) / 2 : x / 2;
    return x;
}
constexpr uint64_t cneg(uint64_t x) { return x ? P - x : 0; }

struct DiagTables {
    uint32_t d16[16];
    uint32_t d24[24];
};

constexpr DiagTables make_diag() {
    DiagTables t{};
    const uint64_t c16[16] = {cneg(2), 1, 2, cinv2k(1), 3, 4, cneg(cinv2k(1)), cneg(3), cneg(4),
                              cinv2k(8), cinv2k(3), cinv2k(24), cneg(cinv2k(8)), cneg(cinv2k(3)),
                              cneg(cinv2k(4)), cneg(cinv2k(24))};
    const uint64_t c24[24] = {cneg(2), 1, 2, cinv2k(1), 3, 4, cneg(cinv2k(1)), cneg(3), cneg(4),
                              cinv2k(8), cinv2k(2), cinv2k(3), cinv2k(4), cinv2k(5), cinv2k(6),
                              cinv2k(24), cneg(cinv2k(8)), cneg(cinv2k(3)), cneg(cinv2k(4)),
                              cneg(cinv2k(5)), cneg(cinv2k(6)), cneg(cinv2k(7)), cneg(cinv2k(9)),
                              cneg(cinv2k(24))};
    for (int i = 0; i < 16; i++) t.d16[i] = cmont(c16[i]);
    for (int i = 0; i < 24; i++) t.d24[i] = cmont(c24[i]);
    return t;
}

constexpr DiagTables DIAG = make_diag();
__device__ __constant__ static const DiagTables DDIAG = make_diag();

// Tables of the partial rounds by convolution (PartialConv): powers of the
// internal diagonal's entries 1..W-1 and their column sums, from the canonical
// diagonal (d = Montgomery image / 2^32, recovered as mont * R^-1).
constexpr uint64_t cmulmod(uint64_t a, uint64_t b) { return a * b % P; }
constexpr uint64_t from_cmont(uint32_t m) {  // m R^-1 mod P, R^-1 = (2^32 mod P)^(P-2)
    uint64_t b = (1ull << 32) % P, e = P - 2, inv = 1;
    while (e) { if (e & 1) inv = cmulmod(inv, b); b = cmulmod(b, b); e >>= 1; }
    return cmulmod(m, inv);
}
template <int W, int RP>
constexpr PartialConv<W, RP> make_pconv(const uint32_t *diag_mont) {
    PartialConv<W, RP> t{};
    uint64_t pw[W];
    for (int i = 1; i < W; i++) pw[i] = 1;
    for (int k = 1; k <= RP; k++) {
        uint64_t cs = 0;
        for (int i = 1; i < W; i++) {
            pw[i] = cmulmod(pw[i], from_cmont(diag_mont[i]));
            t.dpow[k - 1][i - 1] = cmont(pw[i]);
            cs = (cs + pw[i]) % P;
        }
        t.csum[k - 1] = cmont(cs);
    }
    // h = d^RP / 2 of each pair (d = the entry DiagIdx::plus(j))
    for (int j = 0; j < DiagIdx<W>::NP; j++) {
        const uint64_t d = from_cmont(diag_mont[DiagIdx<W>::plus(j)]);
        uint64_t h = cinv2k(1);
        for (int k = 0; k < RP; k++) h = cmulmod(h, d);
        t.hpow[j] = cmont(h);
    }
    t.wm1 = cmont(W - 1);
    return t;
}
// DiagIdx lists every entry 1..W-1 once, each pair as d and -d
template <int W>
constexpr bool diag_idx_ok(const uint32_t *diag_mont) {
    using D = DiagIdx<W>;
    int seen[W] = {};
    for (int j = 0; j < D::NP; j++) {
        const int a = D::plus(j), b = D::minus(j);
        if (a < 1 || a >= W || b < 1 || b >= W) return false;
        seen[a]++, seen[b]++;
        if ((from_cmont(diag_mont[a]) + from_cmont(diag_mont[b])) % P != 0) return false;
    }
    for (int j = 0; j < D::NS; j++) {
        const int a = D::single(j);
        if (a < 1 || a >= W) return false;
        seen[a]++;
    }
    for (int i = 1; i < W; i++)
        if (seen[i] != 1) return false;
    return 2 * D::NP + D::NS == W - 1;
}
static_assert(diag_idx_ok<16>(DIAG.d16), "width-16 diagonal pairs");
static_assert(diag_idx_ok<24>(DIAG.d24), "width-24 diagonal pairs");
constexpr PartialConv<16, 20> PCONV16 = make_pconv<16, 20>(DIAG.d16);
constexpr PartialConv<24, 23> PCONV24 = make_pconv<24, 23>(DIAG.d24);
__device__ __constant__ static const PartialConv<16, 20> DPCONV16 = make_pconv<16, 20>(DIAG.d16);
__device__ __constant__ static const PartialConv<24, 23> DPCONV24 = make_pconv<24, 23>(DIAG.d24);

// Poseidon2 with the fixed KoalaBear diagonals.
template <typename C>
__device__ __forceinline__ void p16_permute(F *s, const C &k) { p2_permute_conv<16, 20>(s, k, DPCONV16); }
template <typename C>
__device__ __forceinline__ void p24_permute(F *s, const C &k) { p2_permute_conv<24, 23>(s, k, DPCONV24); }

}  // namespace kb

#define HIP_CHECK(x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "HIP error %s at %s:%d: %s\n", #x, __FILE__, __LINE__,        \
                    hipGetErrorString(e_));                                               \
            throw std::runtime_error(hipGetErrorString(e_));                              \
        }                                                                                 \
    } while (0)
