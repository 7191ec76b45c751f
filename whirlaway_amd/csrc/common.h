// common.h -- device helpers shared by the prover kernels: compile-time
// Poseidon2 diagonals, wave/block reductions of extension-field values and
// the host-side launch/check macros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include "kb.h"

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
    t.wm1 = cmont(W - 1);
    return t;
}
constexpr PartialConv<16, 20> PCONV16 = make_pconv<16, 20>(DIAG.d16);
constexpr PartialConv<24, 23> PCONV24 = make_pconv<24, 23>(DIAG.d24);
__device__ __constant__ static const PartialConv<16, 20> DPCONV16 = make_pconv<16, 20>(DIAG.d16);
__device__ __constant__ static const PartialConv<24, 23> DPCONV24 = make_pconv<24, 23>(DIAG.d24);

// Poseidon2 with the fixed KoalaBear diagonals.
template <typename C>
__device__ __forceinline__ void p16_permute(F *s, const C &k) { p2_permute_conv<16, 20>(s, k, DPCONV16); }
template <typename C>
__device__ __forceinline__ void p24_permute(F *s, const C &k) { p2_permute_conv<24, 23>(s, k, DPCONV24); }

// Lazy accumulator of EF x F products: 64-bit per coefficient (4 products of
// values < P stay below 2^64 after a fold64), reduced once in get().
struct AccF {
    uint64_t s[8];
    EF r;
    int n;
    __device__ __forceinline__ AccF() : n(0) {
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = 0;
        r = ezero();
    }
    __device__ __forceinline__ void flush() {
#pragma unroll
        for (int i = 0; i < 8; i++) { r.c[i] = fadd(r.c[i], F{redc64(s[i])}); s[i] = 0; }
        n = 0;
    }
    // sums stay below 2^64: shrink them (fold64, one multiply-add each) after every
    // fourth product and reduce only in get()
    __device__ __forceinline__ void add(const EF &a, F b) {
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += (uint64_t)a.c[i].v * b.v;
        if (++n == 4) {
#pragma unroll
            for (int i = 0; i < 8; i++) s[i] = fold64(s[i]);
            n = 0;
        }
    }
    __device__ __forceinline__ EF get() { flush(); return r; }
};

// ---------------------------------------------------------- reductions --
__device__ __forceinline__ EF wave_sum(EF v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 8; c++) v.c[c] = fadd(v.c[c], F{(uint32_t)__shfl_xor((int)v.c[c].v, off, 64)});
    }
    return v;
}

// Sum over the block (blockDim.x multiple of 64, at most 1024); result valid in thread 0.
__device__ __forceinline__ EF block_sum(EF v, EF *lds /* >= 16 entries */) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    EF r = ezero();
    if (threadIdx.x == 0) {
        int nw = blockDim.x >> 6;
        for (int i = 0; i < nw; i++) r = eadd(r, lds[i]);
    }
    return r;
}

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
