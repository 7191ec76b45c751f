// reduce.h -- lazy EF x F accumulation and wave/block sums of extension-field
// values, shared by the prover kernels.
#pragma once
#include "kb.h"

namespace kb {

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
