// sc_generic.hip -- sumcheck::prove (crates/sumcheck/src/prove.rs:22-254) for
// any computation given as a constraint program, and its C-ABI entry point.
//
// Round r binds 2^k values of the most significant variables (k = skips in
// round 0, else 1).  For each evaluation point z of the round a block row
// (blockIdx.y) walks the hypercube of the remaining variables: every input the
// program loads is that multilinear folded on the fly by z's scalars
// (selectors of the skipped variables, or (1 - z, z)), the program's asserted
// values are combined with the batching scalars, weighted by eq(eq_factor[1 +
// r..], x) when an eq factor is given, and summed; the host interpolates the
// round polynomial, runs the transcript, and every table is folded by the
// challenge's scalars in one launch.  All arithmetic is exact in F_p, so the
// block/tree order of the sums does not change the result.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <queue>
#include <type_traits>
#include "../../include/whirlaway_hip.h"
#include "sc_generic.h"
#include "trace.h"

using namespace wl;

namespace wg {

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }
constexpr int BUNDLE_WORDS = 10;  // a bundle: 2 header words and up to 4 entries of 2

__device__ __forceinline__ F d_add(F a, F b) { return fadd(a, b); }
__device__ __forceinline__ EF d_add(const EF &a, const EF &b) { return eadd(a, b); }
__device__ __forceinline__ F d_sub(F a, F b) { return fsub(a, b); }
__device__ __forceinline__ EF d_sub(const EF &a, const EF &b) { return esub(a, b); }
__device__ __forceinline__ F d_mul(F a, F b) { return fmul(a, b); }
__device__ __forceinline__ EF d_mul(const EF &a, const EF &b) { return emul(a, b); }
// x times a constant held in a slot (an F, or an EF with only coefficient 0): EF x F
__device__ __forceinline__ F d_mulc(F a, F c) { return fmul(a, c); }
__device__ __forceinline__ EF d_mulc(const EF &a, const EF &c) { return emulf(a, c.c[0]); }
__device__ __forceinline__ F d_neg(F a) { return fneg(a); }
__device__ __forceinline__ EF d_neg(const EF &a) { return eneg(a); }
template <class T> __device__ __forceinline__ T d_const(F c);
template <> __device__ __forceinline__ F d_const<F>(F c) { return c; }
template <> __device__ __forceinline__ EF d_const<EF>(F c) { return efrom(c); }

// input `in` at hypercube point x, folded over its top k variables by the
// point's K scalars: sum_j s[j] table_in[j half + x]
__device__ __forceinline__ F load_folded(const F *tab, size_t half, size_t x, const F *s, int K) {
    uint64_t acc = 0;
    for (int j = 0; j < K; j++) {
        acc += (uint64_t)s[j].v * tab[(size_t)j * half + x].v;
        if ((j & 3) == 3) acc = fold64(acc);
    }
    return F{redc64(acc)};
}
__device__ __forceinline__ EF load_folded(const EF *tab, size_t half, size_t x, const F *s, int K) {
    AccF acc;
    for (int j = 0; j < K; j++) acc.add(tab[(size_t)j * half + x], s[j]);
    return acc.get();
}

// Block partials of sum_x eq[x] * sum_k batching[k] c_k(point_z(x)) for the round's
// evaluation point blockIdx.y (scalars zsc[blockIdx.y][0..K)) over the constraints of
// group blockIdx.z of a partition (part: G + 1 word offsets, then the groups' slot
// code); tables: n inputs of `len` entries each, contiguous.  An ASSERT names its
// batching scalar in its destination field.  parts[(z G + g) gridDim.x + block].
template <class T, int MAXR>
__global__ void __launch_bounds__(128) k_eval(const T *tables, size_t len, int k, const F *zsc, const uint32_t *part,
                                              int G, const EF *batching, const EF *eq, EF *parts) {
    __shared__ EF lds[16];
    const int K = 1 << k;
    const size_t half = len >> k;
    const F *s = zsc + (size_t)blockIdx.y * K;
    const uint32_t *prog = part + (G + 1) + part[blockIdx.z];
    const int n_prog = (int)(part[blockIdx.z + 1] - part[blockIdx.z]) / 2;
    T regs[MAXR];
    EF total = ezero();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < half; x += stride) {
        EF acc = ezero();
        // the next instruction's words are fetched while this one executes (the
        // program buffer has two words of padding after the last group)
        uint32_t nw0 = prog[0], nw1 = prog[1];
        for (int pc = 0; pc < n_prog; pc++) {
            const uint32_t w0 = nw0, w1 = nw1;
            nw0 = prog[2 * pc + 2];
            nw1 = prog[2 * pc + 3];
            const uint32_t op = w0 & 0xff, dst = w0 >> 8, a = w1 & 0xffff, b = w1 >> 16;
            switch (op) {
            case OP_LOAD: regs[dst] = load_folded(tables + (size_t)a * len, half, x, s, K); break;
            case OP_CONST: regs[dst] = d_const<T>(F{w1}); break;
            case OP_ADD: regs[dst] = d_add(regs[a], regs[b]); break;
            case OP_SUB: regs[dst] = d_sub(regs[a], regs[b]); break;
            case OP_MUL: regs[dst] = d_mul(regs[a], regs[b]); break;
            case OP_MULC: regs[dst] = d_mulc(regs[a], regs[b]); break;
            case OP_NEG: regs[dst] = d_neg(regs[a]); break;
            default: {  // OP_ASSERT
                if constexpr (std::is_same<T, F>::value) acc = eadd(acc, emulf(batching[dst], regs[a]));
                else acc = eadd(acc, emul(batching[dst], regs[a]));
            }
            }
        }
        if (eq) acc = emul(acc, eq[x]);
        total = eadd(total, acc);
    }
    const EF r = block_sum(total, lds);
    if (threadIdx.x == 0) parts[((size_t)blockIdx.y * G + blockIdx.z) * gridDim.x + blockIdx.x] = r;
}

// k_eval with the register file in LDS instead of private (scratch) memory, for
// launches with few blocks: 64 threads per block, slot s of lane l at words
// [(s W + c) 64 + l] (W words per value), so a wave's accesses hit 64 distinct
// banks.  An interpreted instruction then waits on LDS, not on scratch.
template <class T>
__global__ void __launch_bounds__(64) k_eval_lds(const T *tables, size_t len, int k, const F *zsc, const uint32_t *part,
                                                 int G, const EF *batching, const EF *eq, EF *parts) {
    extern __shared__ uint32_t lregs[];
    __shared__ EF lds[16];
    constexpr int WD = (int)(sizeof(T) / 4);
    const int lane = threadIdx.x;
    auto rd = [&](uint32_t slot) {
        T v;
        uint32_t *pv = (uint32_t *)&v;
#pragma unroll
        for (int c = 0; c < WD; c++) pv[c] = lregs[(slot * WD + c) * 64 + lane];
        return v;
    };
    auto wr = [&](uint32_t slot, const T &v) {
        const uint32_t *pv = (const uint32_t *)&v;
#pragma unroll
        for (int c = 0; c < WD; c++) lregs[(slot * WD + c) * 64 + lane] = pv[c];
    };
    const int K = 1 << k;
    const size_t half = len >> k;
    const F *s = zsc + (size_t)blockIdx.y * K;
    const uint32_t *prog = part + (G + 1) + part[blockIdx.z];
    const int n_prog = (int)(part[blockIdx.z + 1] - part[blockIdx.z]) / 2;
    EF total = ezero();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < half; x += stride) {
        EF acc = ezero();
        // the next instruction's words are fetched while this one executes (the
        // program buffer has two words of padding after the last group)
        uint32_t nw0 = prog[0], nw1 = prog[1];
        for (int pc = 0; pc < n_prog; pc++) {
            const uint32_t w0 = nw0, w1 = nw1;
            nw0 = prog[2 * pc + 2];
            nw1 = prog[2 * pc + 3];
            const uint32_t op = w0 & 0xff, dst = w0 >> 8, a = w1 & 0xffff, b = w1 >> 16;
            switch (op) {
            case OP_LOAD: wr(dst, load_folded(tables + (size_t)a * len, half, x, s, K)); break;
            case OP_CONST: wr(dst, d_const<T>(F{w1})); break;
            case OP_ADD: wr(dst, d_add(rd(a), rd(b))); break;
            case OP_SUB: wr(dst, d_sub(rd(a), rd(b))); break;
            case OP_MUL: wr(dst, d_mul(rd(a), rd(b))); break;
            case OP_MULC: wr(dst, d_mulc(rd(a), rd(b))); break;
            case OP_NEG: wr(dst, d_neg(rd(a))); break;
            default: {  // OP_ASSERT
                if constexpr (std::is_same<T, F>::value) acc = eadd(acc, emulf(batching[dst], rd(a)));
                else acc = eadd(acc, emul(batching[dst], rd(a)));
            }
            }
        }
        if (eq) acc = emul(acc, eq[x]);
        total = eadd(total, acc);
    }
    const EF r = block_sum(total, lds);
    if (threadIdx.x == 0) parts[((size_t)blockIdx.y * G + blockIdx.z) * gridDim.x + blockIdx.x] = r;
}

// ---- extension-field rounds with 16 lanes per point ----
// Lane l of a 16-lane DPP row holds coefficient l mod 8 of every value of its
// point (the row's two halves hold the same coefficients, so a rotation within the
// row is a rotation of the 8 coefficients).  Additions are lane-local; a product
// c_k = sum_r a_{k - r mod 8} b_r (3 when r > k, from X^8 = 3) takes the rotations
// of a (row_ror) and the broadcasts of b and 3b (row_newbcast) into one 64-bit sum
// per lane.  An interpreted instruction costs a few lane instructions instead of a
// whole extension product per lane, which shortens the dependent chain of a point.
template <int R>
__device__ __forceinline__ uint32_t dpp_ror(uint32_t v) {
    if constexpr (R == 0) return v;
    else return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x120 + R, 0xF, 0xF, false);
}
template <int R>
__device__ __forceinline__ uint32_t dpp_bcast(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + R, 0xF, 0xF, false);
}
// 64-bit sum of a_{k - r} b_r (times 3 when r > k) over r, for coefficient k of this lane
template <int R = 0>
__device__ __forceinline__ void lane_mul_acc(uint64_t &acc, uint32_t a, uint32_t b, uint32_t b3, int k) {
    if constexpr (R < 8) {
        if (R == 4) acc = fold64(acc);
        const uint32_t ar = dpp_ror<R>(a), br = dpp_bcast<R>(b), b3r = dpp_bcast<R>(b3);
        acc += (uint64_t)ar * (R > k ? b3r : br);
        lane_mul_acc<R + 1>(acc, a, b, b3, k);
    }
}
__device__ __forceinline__ uint32_t lane_mul(uint32_t a, uint32_t b, int k) {
    uint64_t acc = 0;
    lane_mul_acc(acc, a, b, ftimes3(F{b}).v, k);
    return redc64(acc);
}
// sum_r a_{k - r} B_r (3 B_r when r > k) for a uniform extension constant B (B3 = 3B)
template <int R = 0>
__device__ __forceinline__ void lane_mul_const_acc(uint64_t &acc, uint32_t a, const EF &B, const uint32_t *B3, int k) {
    if constexpr (R < 8) {
        if (R == 4) acc = fold64(acc);
        acc += (uint64_t)dpp_ror<R>(a) * (R > k ? B3[R] : B.c[R].v);
        lane_mul_const_acc<R + 1>(acc, a, B, B3, k);
    }
}

// k_eval for extension-field tables: 16 points per 256-thread block, register file
// in LDS ([slot][thread], one coefficient per thread); partials as k_eval's.
__global__ void __launch_bounds__(256) k_eval_lanes(const EF *tables, size_t len, int k, const F *zsc,
                                                    const uint32_t *part, int G, const EF *batching, const EF *eq,
                                                    EF *parts) {
    extern __shared__ uint32_t lregs[];
    __shared__ uint32_t red[4][8];
    const int t = threadIdx.x, co = t & 7, pt = t >> 4;
    auto rd = [&](uint32_t slot) { return lregs[slot * 256 + t]; };
    auto wr = [&](uint32_t slot, uint32_t v) { lregs[slot * 256 + t] = v; };
    const int K = 1 << k;
    const size_t half = len >> k;
    const F *s = zsc + (size_t)blockIdx.y * K;
    const uint32_t *prog = part + (G + 1) + part[blockIdx.z];
    const int n_prog = (int)(part[blockIdx.z + 1] - part[blockIdx.z]) / 2;
    uint32_t total = 0;  // coefficient co of this lane's points' sum
    const size_t stride = (size_t)gridDim.x * 16;
    // every lane of the block runs the same number of iterations (DPP needs the row)
    for (size_t x0 = (size_t)blockIdx.x * 16; x0 < half; x0 += stride) {
        const size_t x = x0 + pt;
        const bool live = x < half;
        const size_t xl = live ? x : x0;
        uint64_t acc = 0;
        int n_acc = 0;
        uint32_t nw0 = prog[0], nw1 = prog[1];
        for (int pc = 0; pc < n_prog; pc++) {
            const uint32_t w0 = nw0, w1 = nw1;
            nw0 = prog[2 * pc + 2];
            nw1 = prog[2 * pc + 3];
            const uint32_t op = w0 & 0xff, dst = w0 >> 8, a = w1 & 0xffff, b = w1 >> 16;
            switch (op) {
            case OP_LOAD: {
                const EF *tab = tables + (size_t)a * len;
                uint64_t f = 0;
                for (int j = 0; j < K; j++) {
                    if (j && (j & 3) == 0) f = fold64(f);
                    f += (uint64_t)s[j].v * tab[(size_t)j * half + xl].c[co].v;
                }
                wr(dst, redc64(f));
                break;
            }
            case OP_CONST: wr(dst, co == 0 ? w1 : 0u); break;
            case OP_ADD: wr(dst, fadd(F{rd(a)}, F{rd(b)}).v); break;
            case OP_SUB: wr(dst, fsub(F{rd(a)}, F{rd(b)}).v); break;
            case OP_MUL: wr(dst, lane_mul(rd(a), rd(b), co)); break;
            case OP_MULC: wr(dst, fmul(F{rd(a)}, F{dpp_bcast<0>(rd(b))}).v); break;
            case OP_NEG: wr(dst, fneg(F{rd(a)}).v); break;
            default: {  // OP_ASSERT: batching[dst] * value into the lane's lazy sum
                const EF B = batching[dst];
                uint32_t B3[8];
#pragma unroll
                for (int i = 0; i < 8; i++) B3[i] = ftimes3(B.c[i]).v;
                if (n_acc == 64) { acc = fold64(acc); n_acc = 0; }
                uint64_t prod = 0;
                lane_mul_const_acc(prod, rd(a), B, B3, co);
                acc += fold64(prod);
                n_acc++;
            }
            }
        }
        uint32_t v = redc64(acc);
        if (eq) v = lane_mul(v, eq[xl].c[co].v, co);
        if (live) total = fadd(F{total}, F{v}).v;
    }
    // sum over the block's 16 points: the four rows of a wave, then the four waves
    total = fadd(F{total}, F{(uint32_t)__shfl_xor((int)total, 16, 64)}).v;
    total = fadd(F{total}, F{(uint32_t)__shfl_xor((int)total, 32, 64)}).v;
    if ((t & 63) < 8) red[t >> 6][t & 7] = total;
    __syncthreads();
    if (t < 8) {
        uint32_t r = red[0][t];
        for (int w = 1; w < 4; w++) r = fadd(F{r}, F{red[w][t]}).v;
        EF *o = parts + ((size_t)blockIdx.y * G + blockIdx.z) * gridDim.x + blockIdx.x;
        o->c[t] = F{r};
    }
}

// k_eval_lanes over bundles (Partition::bcode): each step runs up to four
// independent instructions of one op, their operand reads first, then the
// arithmetic (the products' DPP chains interleave), then the writes.  The next
// bundle's words are fetched while the current one runs.
template <int C>
using BundleN = std::integral_constant<int, C>;
__global__ void __launch_bounds__(256) k_eval_bundles(const EF *tables, size_t len, int k, const F *zsc,
                                                      const uint32_t *bpart, int G, const EF *batching,
                                                      const EF *eq, EF *parts) {
    extern __shared__ uint32_t lregs[];
    __shared__ uint32_t red[4][8];
    const int t = threadIdx.x, co = t & 7, pt = t >> 4;
    auto rd = [&](uint32_t slot) { return lregs[slot * 256 + t]; };
    auto wr = [&](uint32_t slot, uint32_t v) { lregs[slot * 256 + t] = v; };
    const int K = 1 << k;
    const size_t half = len >> k;
    const F *s = zsc + (size_t)blockIdx.y * K;
    const uint32_t *prog = bpart + (G + 1) + bpart[blockIdx.z];
    const int n_words = (int)(bpart[blockIdx.z + 1] - bpart[blockIdx.z]);
    uint32_t total = 0;  // coefficient co of this lane's points' sum
    const size_t stride = (size_t)gridDim.x * 16;
    // every lane of the block runs the same number of iterations (DPP needs the row)
    for (size_t x0 = (size_t)blockIdx.x * 16; x0 < half; x0 += stride) {
        const size_t x = x0 + pt;
        const bool live = x < half;
        const size_t xl = live ? x : x0;
        uint64_t acc = 0;
        int n_acc = 0;
        uint32_t nb[BUNDLE_WORDS];
#pragma unroll
        for (int i = 0; i < BUNDLE_WORDS; i++) nb[i] = prog[i];
        for (int pc = 0; pc < n_words;) {
            uint32_t e[BUNDLE_WORDS];
#pragma unroll
            for (int i = 0; i < BUNDLE_WORDS; i++) e[i] = nb[i];
            const uint32_t op = e[0] & 0xff, cnt = e[0] >> 8;
            const int next = pc + 2 + 2 * (int)cnt;
#pragma unroll
            for (int i = 0; i < BUNDLE_WORDS; i++) nb[i] = prog[next + i];
            auto run = [&](auto fn) {
                switch (cnt) {
                case 1: fn(BundleN<1>{}); break;
                case 2: fn(BundleN<2>{}); break;
                case 3: fn(BundleN<3>{}); break;
                default: fn(BundleN<4>{}); break;
                }
            };
            switch (op) {
            case OP_LOAD:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint64_t f[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) f[i] = 0;
                    for (int j = 0; j < K; j++) {
#pragma unroll
                        for (int i = 0; i < NC; i++) {
                            const EF *tab = tables + (size_t)e[3 + 2 * i] * len;
                            if (j && (j & 3) == 0) f[i] = fold64(f[i]);
                            f[i] += (uint64_t)s[j].v * tab[(size_t)j * half + xl].c[co].v;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < NC; i++) wr(e[2 + 2 * i], redc64(f[i]));
                });
                break;
            case OP_CONST:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
#pragma unroll
                    for (int i = 0; i < NC; i++) wr(e[2 + 2 * i], co == 0 ? e[3 + 2 * i] : 0u);
                });
                break;
            case OP_ADD:
            case OP_SUB:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint32_t a[NC], b[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) a[i] = rd(e[3 + 2 * i] & 0xffff), b[i] = rd(e[3 + 2 * i] >> 16);
#pragma unroll
                    for (int i = 0; i < NC; i++)
                        wr(e[2 + 2 * i], op == OP_ADD ? fadd(F{a[i]}, F{b[i]}).v : fsub(F{a[i]}, F{b[i]}).v);
                });
                break;
            case OP_NEG:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint32_t a[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) a[i] = rd(e[3 + 2 * i] & 0xffff);
#pragma unroll
                    for (int i = 0; i < NC; i++) wr(e[2 + 2 * i], fneg(F{a[i]}).v);
                });
                break;
            case OP_MUL:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint32_t a[NC], b[NC], r[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) a[i] = rd(e[3 + 2 * i] & 0xffff), b[i] = rd(e[3 + 2 * i] >> 16);
#pragma unroll
                    for (int i = 0; i < NC; i++) r[i] = lane_mul(a[i], b[i], co);
#pragma unroll
                    for (int i = 0; i < NC; i++) wr(e[2 + 2 * i], r[i]);
                });
                break;
            case OP_MULC:
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint32_t a[NC], c[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) a[i] = rd(e[3 + 2 * i] & 0xffff), c[i] = rd(e[3 + 2 * i] >> 16);
#pragma unroll
                    for (int i = 0; i < NC; i++) wr(e[2 + 2 * i], fmul(F{a[i]}, F{dpp_bcast<0>(c[i])}).v);
                });
                break;
            default:  // OP_ASSERT: batching[rank] * value into the lane's lazy sum
                run([&](auto CN) {
                    constexpr int NC = decltype(CN)::value;
                    uint32_t a[NC];
                    EF B[NC];
#pragma unroll
                    for (int i = 0; i < NC; i++) a[i] = rd(e[3 + 2 * i]), B[i] = batching[e[2 + 2 * i]];
#pragma unroll
                    for (int i = 0; i < NC; i++) {
                        uint32_t B3[8];
#pragma unroll
                        for (int q = 0; q < 8; q++) B3[q] = ftimes3(B[i].c[q]).v;
                        if (n_acc == 64) { acc = fold64(acc); n_acc = 0; }
                        uint64_t prod = 0;
                        lane_mul_const_acc(prod, a[i], B[i], B3, co);
                        acc += fold64(prod);
                        n_acc++;
                    }
                });
            }
            pc = next;
        }
        uint32_t v = redc64(acc);
        if (eq) v = lane_mul(v, eq[xl].c[co].v, co);
        if (live) total = fadd(F{total}, F{v}).v;
    }
    // sum over the block's 16 points: the four rows of a wave, then the four waves
    total = fadd(F{total}, F{(uint32_t)__shfl_xor((int)total, 16, 64)}).v;
    total = fadd(F{total}, F{(uint32_t)__shfl_xor((int)total, 32, 64)}).v;
    if ((t & 63) < 8) red[t >> 6][t & 7] = total;
    __syncthreads();
    if (t < 8) {
        uint32_t r = red[0][t];
        for (int w = 1; w < 4; w++) r = fadd(F{r}, F{red[w][t]}).v;
        EF *o = parts + ((size_t)blockIdx.y * G + blockIdx.z) * gridDim.x + blockIdx.x;
        o->c[t] = F{r};
    }
}

// out[m][i] = sum_j sc[j] in[m][j M + i], M = len / K, for every table m
template <class T>
__global__ void k_fold_tables(const T *in, size_t len, int K, const EF *sc, EF *out, size_t total) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const size_t M = len / K, m = t / M, i = t - m * M;
    const T *src = in + m * len + i;
    if constexpr (std::is_same<T, F>::value) {
        AccF acc;
        for (int j = 0; j < K; j++) acc.add(sc[j], src[(size_t)j * M]);
        out[t] = acc.get();
    } else {
        EF acc = ezero();
        for (int j = 0; j < K; j++) acc = eadd(acc, emul(sc[j], src[(size_t)j * M]));
        out[t] = acc;
    }
}

__global__ void k_up_down(const F *pre, int n_pre, const F *wit, int n_cols, size_t N, F *out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_cols * N) return;
    const size_t c = t / N, x = t - c * N;
    const F *col = (int)c < n_pre ? pre + c * N : wit + (c - n_pre) * N;
    out[t] = col[x == N - 1 ? N - 2 : x];                    // up
    out[(size_t)n_cols * N + t] = col[x == N - 1 ? N - 1 : x + 1];  // down
}
void columns_up_down(const F *pre, int n_pre, const F *wit, int n_cols, int log_n, F *out, hipStream_t s) {
    const size_t N = (size_t)1 << log_n, total = (size_t)n_cols * N;
    k_up_down<<<nblk(total, 256), 256, 0, s>>>(pre, n_pre, wit, n_cols, N, out);
}

// --------------------------------------------------------------- compile --
namespace {

int n_operands(uint32_t op) {
    return (op == OP_ADD || op == OP_SUB || op == OP_MUL) ? 2 : (op == OP_NEG || op == OP_ASSERT) ? 1 : 0;
}

// Slot code of the instructions j with keep[j] set (their operands must be kept or
// inputs): a linear scan gives each live value a slot, operands are freed before the
// result is placed, inputs are loaded where first used.  An ASSERT carries its
// batching index (its rank among all ASSERTs of the program) in the destination
// field.  Returns the register-file size.
int emit_slots(const uint32_t *code, size_t n_instr, int n_inputs, const std::vector<char> &keep,
               const std::vector<int> &assert_rank, std::vector<uint32_t> &dev) {
    const size_t nv = (size_t)n_inputs + n_instr;
    std::vector<long> last(nv, -1);
    for (size_t j = 0; j < n_instr; j++) {
        if (!keep[j]) continue;
        const uint32_t *w = code + 4 * j;
        for (int o = 0; o < n_operands(w[0]); o++) last[w[1 + o]] = (long)j;
    }
    std::vector<int> slot(nv, -1), free_slots;
    int n_slots = 0;
    auto take = [&]() {
        if (!free_slots.empty()) { int s = free_slots.back(); free_slots.pop_back(); return s; }
        return n_slots++;
    };
    auto emit = [&](uint32_t op, uint32_t dst, uint32_t w1) { dev.push_back(op | (dst << 8)); dev.push_back(w1); };
    auto is_const = [&](uint32_t v) { return v >= (uint32_t)n_inputs && code[4 * (v - n_inputs)] == OP_CONST; };
    for (size_t j = 0; j < n_instr; j++) {
        if (!keep[j]) continue;
        const uint32_t *w = code + 4 * j;
        const int nops = n_operands(w[0]);
        for (int o = 0; o < nops; o++) {
            const uint32_t v = w[1 + o];
            if (v < (uint32_t)n_inputs && slot[v] < 0) {
                slot[v] = take();
                emit(OP_LOAD, (uint32_t)slot[v], v);
            }
        }
        const uint32_t a = nops >= 1 ? (uint32_t)slot[w[1]] : 0, b = nops >= 2 ? (uint32_t)slot[w[2]] : 0;
        if (w[0] == OP_ASSERT) emit(OP_ASSERT, (uint32_t)assert_rank[j], a);
        // operands whose last use is here (twice-used operands are freed once)
        for (int o = 0; o < nops; o++) {
            const uint32_t v = w[1 + o];
            if (last[v] == (long)j && slot[v] >= 0 && !(o == 1 && w[1] == w[2])) {
                free_slots.push_back(slot[v]);
                slot[v] = -2;
            }
        }
        if (w[0] == OP_ASSERT) continue;
        const size_t vid = (size_t)n_inputs + j;
        if (last[vid] < 0) continue;  // dead value: not computed
        slot[vid] = take();
        // a product with a constant operand: extension times base field (the constant's
        // slot holds it in coefficient 0)
        if (w[0] == OP_CONST) emit(OP_CONST, (uint32_t)slot[vid], to_mont(w[1]).v);
        else if (w[0] == OP_MUL && is_const(w[2])) emit(OP_MULC, (uint32_t)slot[vid], a | (b << 16));
        else if (w[0] == OP_MUL && is_const(w[1])) emit(OP_MULC, (uint32_t)slot[vid], b | (a << 16));
        else emit(w[0], (uint32_t)slot[vid], a | (b << 16));
    }
    return std::max(1, n_slots);
}

// The instructions j with keep[j] set as bundles for k_eval_bundles.  A list
// schedule over the dependence graph (input loads are nodes too, placed at their
// first user): each step takes the ready node first in program order and up to
// three more ready nodes of the same op, earliest first, so that independent work
// (the products of an internal layer, the additions of a balanced sum, the loads
// of a round's inputs) runs in one dispatch with its memory and LDS latencies
// overlapped, while values stay live about as long as in program order.  Slots are assigned over the bundle order: a
// bundle reads all its operands before it writes, so a result may take the slot an
// operand of the same bundle frees.  Returns the register-file size.
int emit_bundles(const uint32_t *code, size_t n_instr, int n_inputs, const std::vector<char> &keep,
                 const std::vector<int> &assert_rank, std::vector<uint32_t> &dev) {
    const size_t NN = n_instr + (size_t)n_inputs;  // node n_instr + v loads input v
    auto node_of = [&](uint32_t v) { return v < (uint32_t)n_inputs ? n_instr + v : (size_t)(v - n_inputs); };
    auto is_const = [&](uint32_t v) { return v >= (uint32_t)n_inputs && code[4 * (v - n_inputs)] == OP_CONST; };
    // operands as nodes (a repeated operand once)
    auto operands = [&](size_t j, size_t *o) {
        const uint32_t *w = code + 4 * j;
        const int nops = n_operands(w[0]);
        int n = 0;
        for (int q = 0; q < nops; q++) {
            const size_t u = node_of(w[1 + q]);
            if (n == 0 || o[0] != u) o[n++] = u;
        }
        return n;
    };
    // live nodes: the kept instructions on a path to an ASSERT, and the inputs they read
    std::vector<char> live(NN, 0);
    for (size_t j = n_instr; j-- > 0;) {
        if (!keep[j]) continue;
        if (code[4 * j] == OP_ASSERT) live[j] = 1;
        if (!live[j]) continue;
        size_t o[2];
        const int n = operands(j, o);
        for (int q = 0; q < n; q++) live[o[q]] = 1;
    }
    std::vector<std::vector<uint32_t>> users(NN);
    std::vector<int> ndeps(NN, 0);
    for (size_t j = 0; j < n_instr; j++) {
        if (!live[j]) continue;
        size_t o[2];
        const int n = operands(j, o);
        for (int q = 0; q < n; q++) users[o[q]].push_back((uint32_t)j);
        ndeps[j] = n;
    }
    auto kind = [&](size_t u) -> uint32_t {
        if (u >= n_instr) return OP_LOAD;
        const uint32_t *w = code + 4 * u;
        if (w[0] == OP_MUL && (is_const(w[1]) || is_const(w[2]))) return OP_MULC;
        return w[0];
    };
    // program position: an instruction's index, an input load's first user
    std::vector<long> pos(NN, 0);
    for (size_t u = 0; u < NN; u++) {
        if (u < n_instr) pos[u] = (long)u;
        else pos[u] = users[u].empty() ? 0 : (long)users[u].front();
    }
    // list schedule: a ready queue per op, earliest program position first
    using Key = std::pair<long, uint32_t>;
    std::vector<std::priority_queue<Key, std::vector<Key>, std::greater<Key>>> ready(8);
    for (size_t u = 0; u < NN; u++)
        if (live[u] && ndeps[u] == 0) ready[kind(u)].push({pos[u], (uint32_t)u});
    std::vector<std::vector<uint32_t>> bundles;
    std::vector<uint32_t> bundle_op;
    for (;;) {
        int best_k = -1;
        for (int k = 0; k < 8; k++)
            if (!ready[k].empty() && (best_k < 0 || ready[k].top() < ready[best_k].top())) best_k = k;
        if (best_k < 0) break;
        std::vector<uint32_t> b;
        while (b.size() < 4 && !ready[best_k].empty()) {
            b.push_back(ready[best_k].top().second);
            ready[best_k].pop();
        }
        for (uint32_t u : b)
            for (uint32_t v : users[u])
                if (--ndeps[v] == 0) ready[kind(v)].push({pos[v], v});
        bundles.push_back(std::move(b));
        bundle_op.push_back((uint32_t)best_k);
    }
    // slots over the bundle order
    std::vector<long> last(NN, -1);
    for (size_t bi = 0; bi < bundles.size(); bi++)
        for (uint32_t u : bundles[bi]) {
            size_t o[2];
            const int n = u < n_instr ? operands(u, o) : 0;
            for (int q = 0; q < n; q++) last[o[q]] = (long)bi;
        }
    std::vector<int> slot(NN, -1), free_slots;
    int n_slots = 0;
    auto take = [&]() {
        if (!free_slots.empty()) { int s2 = free_slots.back(); free_slots.pop_back(); return s2; }
        return n_slots++;
    };
    for (size_t bi = 0; bi < bundles.size(); bi++) {
        const std::vector<uint32_t> &b = bundles[bi];
        const uint32_t op = bundle_op[bi];
        // operand slots, read before the bundle frees or writes any
        std::vector<uint32_t> w1(b.size(), 0);
        for (size_t i = 0; i < b.size(); i++) {
            const size_t u = b[i];
            if (u >= n_instr) { w1[i] = (uint32_t)(u - n_instr); continue; }
            const uint32_t *w = code + 4 * u;
            if (op == OP_CONST) w1[i] = to_mont(w[1]).v;
            else if (op == OP_MULC) {
                const bool c2 = is_const(w[2]);
                const uint32_t xs = (uint32_t)slot[node_of(c2 ? w[1] : w[2])];
                const uint32_t cs = (uint32_t)slot[node_of(c2 ? w[2] : w[1])];
                w1[i] = xs | (cs << 16);
            } else {
                const int nops = n_operands(w[0]);
                const uint32_t a = nops >= 1 ? (uint32_t)slot[node_of(w[1])] : 0;
                const uint32_t bb = nops >= 2 ? (uint32_t)slot[node_of(w[2])] : 0;
                w1[i] = a | (bb << 16);
            }
        }
        for (size_t i = 0; i < b.size(); i++) {
            size_t o[2];
            const int n = b[i] < n_instr ? operands(b[i], o) : 0;
            for (int q = 0; q < n; q++)
                if (last[o[q]] == (long)bi && slot[o[q]] >= 0) {
                    free_slots.push_back(slot[o[q]]);
                    slot[o[q]] = -2;
                }
        }
        dev.push_back(op | ((uint32_t)b.size() << 8));
        dev.push_back(0);
        for (size_t i = 0; i < b.size(); i++) {
            const size_t u = b[i];
            uint32_t w0;
            if (op == OP_ASSERT) w0 = (uint32_t)assert_rank[u];
            else w0 = (uint32_t)(slot[u] = take());
            dev.push_back(w0);
            dev.push_back(w1[i]);
        }
    }
    return std::max(1, n_slots);
}

// One partition (see compile): false when the program has too few constraints for G groups.
// bundles: emit only the bundled code (Partition::bcode; left empty, and the
// partition's bundles unused, when they would keep more than 256 values live),
// else only the one-instruction slot code.
bool build_partition(const uint32_t *code, size_t n_instr, int n_inputs, int G, bool bundles, Program &out) {
    std::vector<char> is_assert(n_instr, 0);
    std::vector<int> assert_rank(n_instr, -1);
    int n_asserts = 0;
    for (size_t j = 0; j < n_instr; j++)
        if (code[4 * j] == OP_ASSERT) { is_assert[j] = 1; assert_rank[j] = n_asserts++; }
    if (n_asserts > (1 << 24)) throw std::runtime_error("too many asserted constraints");
    out.n_asserts = n_asserts;
    if (G > 1 && n_asserts < 2 * G) return false;
    std::vector<int> weight(n_instr, 1);
    for (size_t j = 0; j < n_instr; j++)
        if (code[4 * j] == OP_MUL) weight[j] = 8;
    long total = 0;
    {
        std::vector<char> used(n_instr, 0);
        for (size_t j = n_instr; j-- > 0;) {
            if (!is_assert[j] && !used[j]) continue;
            used[j] = 1;
            total += weight[j];
            const uint32_t *w = code + 4 * j;
            for (int o = 0; o < n_operands(w[0]); o++)
                if (w[1 + o] >= (uint32_t)n_inputs) used[w[1 + o] - n_inputs] = 1;
        }
    }
    std::vector<uint32_t> stack;
    Partition part;
    std::vector<char> keep(n_instr, 0);
    long group_w = 0;
    auto close_group = [&]() {
        if (bundles) {
            part.boff.push_back((uint32_t)part.bcode.size());
            part.bn_slots = std::max(part.bn_slots, emit_bundles(code, n_instr, n_inputs, keep, assert_rank, part.bcode));
        } else {
            part.off.push_back((uint32_t)part.code.size());
            part.n_slots = std::max(part.n_slots, emit_slots(code, n_instr, n_inputs, keep, assert_rank, part.code));
        }
        part.groups++;
        std::fill(keep.begin(), keep.end(), 0);
        group_w = 0;
    };
    for (size_t j = 0; j < n_instr; j++) {
        if (!is_assert[j]) continue;
        // the cone of this ASSERT
        stack.assign(1, (uint32_t)j);
        while (!stack.empty()) {
            const uint32_t i = stack.back();
            stack.pop_back();
            if (keep[i]) continue;
            keep[i] = 1;
            group_w += weight[i];
            const uint32_t *w = code + 4 * i;
            for (int o = 0; o < n_operands(w[0]); o++)
                if (w[1 + o] >= (uint32_t)n_inputs) stack.push_back(w[1 + o] - n_inputs);
        }
        if (G > 1 && part.groups < G - 1 && group_w * G >= total) close_group();
    }
    if (part.groups == 0 || std::count(keep.begin(), keep.end(), 1) > 0) close_group();
    if (bundles) {
        part.boff.push_back((uint32_t)part.bcode.size());
        if (part.bn_slots > 256) {  // beyond the operand encoding: no bundles for this partition
            part.bcode.clear();
            part.boff.clear();
        }
    } else {
        part.off.push_back((uint32_t)part.code.size());
        if (part.n_slots > 256) throw std::runtime_error("program keeps more than 256 values live");
    }
    out.parts.push_back(std::move(part));
    return true;
}

// Sums written as chains of additions (s = s + x_i) rebalanced into trees: an ADD
// whose only use is another ADD is folded into that ADD's sum, and each sum of
// L >= 3 terms is emitted as L - 1 additions combining equal-size partial sums.
// Field addition is exact, so every value the program defines is unchanged; the
// dependent depth of an L-term sum falls from L - 1 to about log2 L, which lets
// the bundled interpreter run its additions side by side.
std::vector<uint32_t> rebalance_sums(const uint32_t *code, size_t n, int n_inputs) {
    std::vector<int> uses(n, 0);
    std::vector<long> user(n, -1);
    for (size_t j = 0; j < n; j++) {
        const uint32_t *w = code + 4 * j;
        for (int o = 0; o < n_operands(w[0]); o++)
            if (w[1 + o] >= (uint32_t)n_inputs) {
                const size_t d = w[1 + o] - n_inputs;
                uses[d]++;
                user[d] = (long)j;
            }
    }
    std::vector<char> inner(n, 0);
    for (size_t j = 0; j < n; j++)
        inner[j] = code[4 * j] == OP_ADD && uses[j] == 1 && code[4 * user[j]] == OP_ADD;
    std::vector<uint32_t> out, map(n_inputs + n, 0);
    for (int v = 0; v < n_inputs; v++) map[v] = (uint32_t)v;
    auto emit = [&](uint32_t op, uint32_t a, uint32_t b) {
        out.push_back(op), out.push_back(a), out.push_back(b), out.push_back(0);
        return (uint32_t)(n_inputs + out.size() / 4 - 1);
    };
    std::vector<uint32_t> stack, leaves, level, next;
    for (size_t j = 0; j < n; j++) {
        const uint32_t *w = code + 4 * j;
        if (inner[j]) continue;  // part of the sum of its (later) user
        if (w[0] != OP_ADD) {
            const int nops = n_operands(w[0]);
            const uint32_t a = nops >= 1 ? map[w[1]] : w[1], b = nops >= 2 ? map[w[2]] : w[2];
            map[n_inputs + j] = emit(w[0], a, b);
            continue;
        }
        // the terms of this sum, left to right
        leaves.clear();
        stack.assign({w[2], w[1]});
        while (!stack.empty()) {
            const uint32_t v = stack.back();
            stack.pop_back();
            if (v >= (uint32_t)n_inputs && inner[v - n_inputs]) {
                const uint32_t *u = code + 4 * (v - n_inputs);
                stack.push_back(u[2]);
                stack.push_back(u[1]);
            } else {
                leaves.push_back(map[v]);
            }
        }
        // pairs of partial sums of equal size combine as soon as both exist (a
        // binary counter over the terms): depth ~log2 L, and only ~log2 L partial
        // sums are live at a time
        level.clear();  // the partial sums
        next.clear();   // their sizes (powers of two, decreasing)
        for (uint32_t v : leaves) {
            level.push_back(v);
            next.push_back(1);
            while (next.size() >= 2 && next[next.size() - 1] == next[next.size() - 2]) {
                const uint32_t r = emit(OP_ADD, level[level.size() - 2], level.back());
                level.pop_back(), next.pop_back();
                level.back() = r;
                next.back() *= 2;
            }
        }
        while (level.size() > 1) {
            const uint32_t r = emit(OP_ADD, level[level.size() - 2], level.back());
            level.pop_back(), next.pop_back();
            level.back() = r;
        }
        map[n_inputs + j] = level[0];
    }
    return out;
}

}  // namespace

void compile(const uint32_t *code_in, size_t n_instr_in, int n_inputs, Program &out) {
    if (n_inputs < 1 || n_inputs > 65535) throw std::runtime_error("a program reads 1 to 65535 inputs");
    {
        std::vector<char> is_assert(n_instr_in, 0);
        for (size_t j = 0; j < n_instr_in; j++) {
            const uint32_t *w = code_in + 4 * j;
            const uint64_t lim = (uint64_t)n_inputs + j;
            auto use = [&](uint32_t v) {
                if (v >= lim) throw std::runtime_error("program operand refers to a later value");
                if (v >= (uint32_t)n_inputs && is_assert[v - n_inputs])
                    throw std::runtime_error("program operand refers to an ASSERT, which defines no value");
            };
            switch (w[0]) {
            case OP_CONST:
                if (w[1] >= P) throw std::runtime_error("program constant is not a canonical field element");
                break;
            case OP_ADD: case OP_SUB: case OP_MUL: use(w[1]); use(w[2]); break;
            case OP_NEG: use(w[1]); break;
            case OP_ASSERT: use(w[1]); is_assert[j] = 1; break;
            default: throw std::runtime_error("unknown program op");
            }
        }
    }
    out.n_inputs = n_inputs;
    out.parts.clear();
    // the bundled code (extension-field rounds) from the program with its sums
    // rebalanced; the one-instruction code from the program as written, whose
    // long-lived balanced-sum terms measured slower in the base-field round
    const std::vector<uint32_t> bal = rebalance_sums(code_in, n_instr_in, n_inputs);
    Program bprog;
    // Partitions of the constraints into G groups (G = 1, 4, 16): consecutive
    // ASSERTs, each group the union of its constraints' dependency cones (shared
    // subexpressions are computed in every group that needs them), closed once it
    // holds its share of the program's weighted instructions (a product weighs 8).
    // Rounds with few points evaluate a point's constraints in G threads at once.
    for (int G : {1, 4, 16}) {
        if (!build_partition(code_in, n_instr_in, n_inputs, G, false, out)) break;
        Partition &q = out.parts.back();
        // bundles only where the rebalanced program splits into as many groups and
        // they fit the operand encoding; otherwise k_eval_lanes runs this partition
        if (build_partition(bal.data(), bal.size() / 4, n_inputs, G, true, bprog) &&
            bprog.parts.back().groups == q.groups && !bprog.parts.back().bcode.empty()) {
            Partition &b = bprog.parts.back();
            q.bn_slots = b.bn_slots;
            q.boff = std::move(b.boff);
            q.bcode = std::move(b.bcode);
        }
    }
    out.n_slots = out.parts[0].n_slots;
    out.dev = out.parts[0].code;
}

void upload(wl_ctx *c, Program &p) {
    std::vector<uint32_t> all;
    p.part_base.clear();
    for (const Partition &q : p.parts) {
        p.part_base.push_back(all.size());
        all.insert(all.end(), q.off.begin(), q.off.end());
        all.insert(all.end(), q.code.begin(), q.code.end());
    }
    all.push_back(0);  // padding read by the last instruction's prefetch
    all.push_back(0);
    p.bpart_base.clear();
    for (const Partition &q : p.parts) {
        p.bpart_base.push_back(all.size());
        all.insert(all.end(), q.boff.begin(), q.boff.end());
        all.insert(all.end(), q.bcode.begin(), q.bcode.end());
        all.insert(all.end(), BUNDLE_WORDS, 0u);  // read by the last bundle's prefetch
    }
    p.dev_buf = DBuf<uint32_t>(c->pool, all.size());
    c->h2d(p.dev_buf.p, all.data(), all.size());
}

// The launch of one evaluation.  Extension-field tables: 16 lanes per point
// (k_eval_lanes) with the partition whose groups bring the round's lanes to 2^17,
// or the finest one.  Base-field tables: the partition whose groups bring the
// round's threads (points x evaluation points) closest to 2^17 from below, or the
// finest one; the register file in LDS when the grid of 64-thread blocks is at most
// one block per CU and the slots fit, else in scratch with 128-thread blocks.
enum EvalMode { EVAL_SCRATCH, EVAL_LDS, EVAL_LANES, EVAL_BUNDLES };
struct EvalLaunch {
    size_t pi;
    EvalMode mode;
    int gx, G;
    size_t lds_bytes;
};
constexpr size_t EVAL_LDS_MAX = 160 * 1024 - 1024;  // LDS per CU, less the reductions'
template <class T>
EvalLaunch plan_eval(const Program &p, size_t len, int k, int nz) {
    const size_t half = len >> k;
    EvalLaunch e;
    if (std::is_same<T, EF>::value) {
        size_t pi = 0;
        while (pi + 1 < p.parts.size() && half * (size_t)nz * (size_t)p.parts[pi].groups * 16 < ((size_t)1 << 17))
            pi++;
        const size_t bundle_bytes = p.parts[pi].bcode.empty() ? SIZE_MAX : (size_t)p.parts[pi].bn_slots * 256 * 4;
        e.lds_bytes = (size_t)p.parts[pi].n_slots * 256 * 4;
        if (bundle_bytes <= EVAL_LDS_MAX || e.lds_bytes <= EVAL_LDS_MAX) {
            e.pi = pi;
            e.mode = bundle_bytes <= EVAL_LDS_MAX ? EVAL_BUNDLES : EVAL_LANES;
            if (e.mode == EVAL_BUNDLES) e.lds_bytes = bundle_bytes;
            e.G = p.parts[pi].groups;
            e.gx = (int)std::min<size_t>(nblk(half, 16), 1024);
            return e;
        }
    }
    size_t pi = 0;
    while (pi + 1 < p.parts.size() && half * (size_t)nz * (size_t)p.parts[pi].groups < ((size_t)1 << 17)) pi++;
    e.pi = pi;
    e.G = p.parts[pi].groups;
    const int gx64 = (int)std::min<size_t>(nblk(half, 64), 1024);
    e.lds_bytes = (size_t)p.parts[pi].n_slots * sizeof(T) * 64;
    // base-field slots are one word: with at most 32 KB per block several blocks share a
    // CU, so large launches take the LDS register file too
    const bool lds_small = (size_t)gx64 * nz * e.G <= 256 && e.lds_bytes <= EVAL_LDS_MAX;
    const bool lds_any = std::is_same<T, F>::value && e.lds_bytes <= (32u << 10);
    e.mode = lds_small || lds_any ? EVAL_LDS : EVAL_SCRATCH;
    e.gx = e.mode == EVAL_LDS ? gx64 : (int)std::min<size_t>(nblk(half, 128), 1024);
    return e;
}
template <class T>
int launch_eval(const Program &p, const T *tab, size_t len, int k, const F *zsc, int nz, const EF *bat, const EF *eq,
                EF *parts, hipStream_t s) {
    const EvalLaunch e = plan_eval<T>(p, len, k, nz);
    const Partition &q = p.parts[e.pi];
    const dim3 grid(e.gx, nz, e.G);
    const uint32_t *part = p.dev_buf.p + p.part_base[e.pi];
    if (e.mode == EVAL_BUNDLES) {
        if constexpr (std::is_same<T, EF>::value) {
            static std::once_flag once;
            std::call_once(once, [] {
                (void)hipFuncSetAttribute((const void *)k_eval_bundles, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)EVAL_LDS_MAX);
            });
            const uint32_t *bpart = p.dev_buf.p + p.bpart_base[e.pi];
            k_eval_bundles<<<grid, 256, e.lds_bytes, s>>>(tab, len, k, zsc, bpart, e.G, bat, eq, parts);
        }
    } else if (e.mode == EVAL_LANES) {
        if constexpr (std::is_same<T, EF>::value) {
            static std::once_flag once;
            std::call_once(once, [] {
                (void)hipFuncSetAttribute((const void *)k_eval_lanes, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)EVAL_LDS_MAX);
            });
            k_eval_lanes<<<grid, 256, e.lds_bytes, s>>>(tab, len, k, zsc, part, e.G, bat, eq, parts);
        }
    } else if (e.mode == EVAL_LDS) {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)k_eval_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)EVAL_LDS_MAX);
        });
        k_eval_lds<T><<<grid, 64, e.lds_bytes, s>>>(tab, len, k, zsc, part, e.G, bat, eq, parts);
    } else if (q.n_slots <= 16) {
        k_eval<T, 16><<<grid, 128, 0, s>>>(tab, len, k, zsc, part, e.G, bat, eq, parts);
    } else if (q.n_slots <= 64) {
        k_eval<T, 64><<<grid, 128, 0, s>>>(tab, len, k, zsc, part, e.G, bat, eq, parts);
    } else {
        k_eval<T, 256><<<grid, 128, 0, s>>>(tab, len, k, zsc, part, e.G, bat, eq, parts);
    }
    return e.gx * e.G;
}
// partials per evaluation point launch_eval will write (for sizing the buffer)
template <class T>
int eval_parts(const Program &p, size_t len, int k, int nz) {
    const EvalLaunch e = plan_eval<T>(p, len, k, nz);
    return e.gx * e.G;
}

static int grind_bits(const Args &a, int degree) {
    if (a.grind_kind == GRIND_NONE) return 0;
    if (a.grind_kind == GRIND_CUSTOM) return a.grind_value;
    return std::max(0, a.grind_value - std::max(0, 248 - (int)wh::log2_up((size_t)degree + 1)));
}

// ----------------------------------------------------------------- prove --
std::vector<EF> prove(wl_ctx *c, wh::Transcript &ts, const P2Dev *pdev, const void *tables, bool base, int n_mles,
                      int n_vars, const Program &prog, const Args &a, DBuf<EF> &folded, EF &sum) {
    const int skips = a.skips;
    if (skips < 1 || skips > 6 || n_vars < skips || n_mles != prog.n_inputs || a.degree < 1)
        throw std::runtime_error("sumcheck: invalid skips, degree or number of multilinears");
    const int nr = a.n_rounds > 0 ? a.n_rounds : n_vars - skips + 1;
    if (nr > n_vars - skips + 1) throw std::runtime_error("sumcheck: more rounds than variables");
    if (a.eq_factor && (int)a.eq_factor->size() != n_vars - skips + 1)
        throw std::runtime_error("sumcheck: eq_factor must hold n_vars - skips + 1 values");
    if ((int)(a.batching ? a.batching->size() : 0) < prog.n_asserts)
        throw std::runtime_error("sumcheck: fewer batching scalars than asserted constraints");
    if ((size_t)a.degree * ((1u << skips) - 1) + 1 > 64) throw std::runtime_error("sumcheck: round degree above 63");
    DBuf<EF> dbat = upload_ef(c, a.batching ? *a.batching : std::vector<EF>());
    EF missing = a.missing_mul_factor ? *a.missing_mul_factor : ezero();
    bool has_missing = a.missing_mul_factor != nullptr;
    sum = a.sum;
    std::vector<EF> challenges;
    const void *cur = tables;
    bool cur_base = base;
    DBuf<EF> owned;
    int nv = n_vars;
    for (int r = 0; r < nr; r++) {
        wl::Span span_round("sumcheck_round");  // crates/sumcheck/src/prove.rs:95
        const int k = r == 0 ? skips : 1, K = 1 << k;
        const size_t len = (size_t)1 << nv;
        const bool zerocheck = r == 0 && a.is_zerofier;
        DBuf<EF> eqm;
        if (a.eq_factor) {
            std::vector<EF> pt(a.eq_factor->begin() + 1 + r, a.eq_factor->end());
            if ((int)pt.size() != nv - k) throw std::runtime_error("sumcheck: eq_factor length");
            if (pt.empty()) eqm = upload_ef(c, std::vector<EF>(1, eone()));
            else eqm = eq_dev(c, pt, eone());
        }
        const int maxz = a.degree * (K - 1), start = zerocheck ? K : 0;
        std::vector<int> zs;
        std::vector<F> zsc;
        for (int z = start; z <= maxz; z++) {
            if (z == K - 1) continue;  // derived from the claim
            zs.push_back(z);
            if (k == 1) {
                zsc.push_back(fsub(fone(), fsmall((uint32_t)z)));
                zsc.push_back(fsmall((uint32_t)z));
            } else {
                std::vector<F> sel = wh::selectors_f(k, fsmall((uint32_t)z));
                zsc.insert(zsc.end(), sel.begin(), sel.end());
            }
        }
        std::vector<EF> vals;
        if (!zs.empty()) {
            DBuf<F> dz(c->pool, zsc.size());
            c->h2d_async(dz.p, zsc.data(), zsc.size());
            const int np = cur_base ? eval_parts<F>(prog, len, k, (int)zs.size())
                                    : eval_parts<EF>(prog, len, k, (int)zs.size());
            DBuf<EF> parts(c->pool, (size_t)np * zs.size());
            if (cur_base) launch_eval(prog, (const F *)cur, len, k, dz, (int)zs.size(), dbat, eqm.p, parts, c->stream);
            else launch_eval(prog, (const EF *)cur, len, k, dz, (int)zs.size(), dbat, eqm.p, parts, c->stream);
            HIP_CHECK(hipGetLastError());
            vals = reduce_many(c, parts, (int)zs.size(), np);
        }
        // the round's points (z, p(z)), z = 0 .. maxz
        std::vector<F> xs;
        std::vector<EF> ys;
        if (zerocheck)
            for (int i = 0; i < K; i++) { xs.push_back(fsmall((uint32_t)i)); ys.push_back(ezero()); }
        size_t vi = 0;
        std::vector<EF> sel_e;
        if (a.eq_factor && k > 1) sel_e = wh::selectors_ef(k, (*a.eq_factor)[r]);
        for (int z = start; z <= maxz; z++) {
            EF v;
            if (z == K - 1) {
                if (a.eq_factor) {
                    const EF e = (*a.eq_factor)[r];
                    if (k == 1) {
                        v = emul(esub(sum, emul(ys[0], esub(eone(), e))), wh::einv(e));
                    } else {
                        EF acc = ezero();
                        for (int i = 0; i < K - 1; i++) acc = eadd(acc, emul(ys[i], sel_e[i]));
                        v = emul(esub(sum, acc), wh::einv(sel_e[K - 1]));
                    }
                } else {
                    EF acc = ezero();
                    for (const EF &y : ys) acc = eadd(acc, y);
                    v = esub(sum, acc);
                }
            } else {
                v = vals[vi++];
                if (has_missing) v = emul(v, missing);
            }
            xs.push_back(fsmall((uint32_t)z));
            ys.push_back(v);
        }
        std::vector<EF> p = wh::lagrange(xs, ys);
        if (a.eq_factor) {
            const EF e = (*a.eq_factor)[r];
            std::vector<EF> q;
            if (k == 1) {
                q = {esub(eone(), e), esub(edbl(e), eone())};
            } else {
                std::vector<F> qx;
                for (int i = 0; i < K; i++) qx.push_back(fsmall((uint32_t)i));
                q = wh::lagrange(qx, sel_e);
            }
            p = wh::pmul(p, q);
        }
        ts.add_ef(p.data(), p.size());
        const EF ch = ts.sample_ef();
        challenges.push_back(ch);
        sum = wh::peval(p, ch);
        grind(c, ts, pdev, grind_bits(a, (a.degree + (a.eq_factor ? 1 : 0)) * (K - 1)));
        std::vector<EF> fs;
        if (k == 1) fs = {esub(eone(), ch), ch};
        else fs = wh::selectors_ef(k, ch);
        if (a.eq_factor) {
            const EF e = (*a.eq_factor)[r];
            EF f;
            if (k == 1) {
                f = eadd(emul(esub(eone(), e), esub(eone(), ch)), emul(e, ch));
            } else {
                f = ezero();
                for (int i = 0; i < K; i++) f = eadd(f, emul(sel_e[i], fs[i]));
            }
            missing = has_missing ? emul(f, missing) : f;
            has_missing = true;
        }
        DBuf<EF> dfs = upload_ef(c, fs);
        const size_t total = (size_t)n_mles * (len / K);
        DBuf<EF> next(c->pool, total);
        if (cur_base) k_fold_tables<F><<<nblk(total, 256), 256, 0, c->stream>>>((const F *)cur, len, K, dfs, next, total);
        else k_fold_tables<EF><<<nblk(total, 256), 256, 0, c->stream>>>((const EF *)cur, len, K, dfs, next, total);
        HIP_CHECK(hipGetLastError());
        owned = std::move(next);
        cur = owned.p;
        cur_base = false;
        nv -= k;
    }
    folded = std::move(owned);
    return challenges;
}

}  // namespace wg

// ===================================================================== C ABI ==
extern "C" {

int wl_sumcheck_prove(wl_ctx *c, wl_transcript *t, uint32_t skips, const uint32_t *mles, uint32_t n_mles,
                      int mles_are_ef, uint32_t n_vars, const uint32_t *program, uint32_t n_instr, uint32_t degree,
                      const uint32_t *batching, uint32_t n_batching, const uint32_t *eq_factor, int is_zerofier,
                      const uint32_t *sum, uint32_t n_rounds, int grinding, uint32_t grinding_value,
                      const uint32_t *missing_mul_factor, uint32_t *challenges_out, uint32_t *folded_out,
                      uint32_t *sum_out) {
    if (!c || !t || t->verifier || !mles || !program || !sum || n_mles < 1 || n_vars > 30 || skips < 1 ||
        skips > n_vars || (n_batching && !batching) || grinding < 0 || grinding > 2) {
        wl_set_error("bad arguments");
        return WL_ERR_ARG;
    }
    try {
        HIP_CHECK(hipSetDevice(c->device));
        wg::Program prog;
        wg::compile(program, n_instr, (int)n_mles, prog);
        wg::upload(c, prog);
        const size_t N = (size_t)1 << n_vars, w = mles_are_ef ? 8 : 1;
        DBuf<uint32_t> tab(c->pool, N * n_mles * w);
        HIP_CHECK(hipMemcpyAsync(tab.p, mles, N * n_mles * w * 4, hipMemcpyHostToDevice, c->stream));
        wk::convert_to_mont(tab.p, N * n_mles * w, c->stream);
        std::vector<EF> bat(n_batching), eqf;
        for (uint32_t i = 0; i < n_batching; i++) bat[i] = wh::to_mont_ef(batching + 8 * (size_t)i);
        if (eq_factor)
            for (uint32_t i = 0; i + skips <= n_vars; i++) eqf.push_back(wh::to_mont_ef(eq_factor + 8 * (size_t)i));
        EF mmf{};
        if (missing_mul_factor) mmf = wh::to_mont_ef(missing_mul_factor);
        wg::Args a;
        a.skips = (int)skips;
        a.degree = (int)degree;
        a.batching = &bat;
        a.eq_factor = eq_factor ? &eqf : nullptr;
        a.is_zerofier = is_zerofier != 0;
        a.sum = wh::to_mont_ef(sum);
        a.n_rounds = (int)n_rounds;
        a.grind_kind = grinding;
        a.grind_value = (int)grinding_value;
        a.missing_mul_factor = missing_mul_factor ? &mmf : nullptr;
        P2Dev ph;
        memset(&ph, 0, sizeof ph);
        ph.p16 = t->p16;
        const P2Dev *pd = context_p2(c, ph);
        c->ring_reset();
        DBuf<EF> folded;
        EF fsum;
        std::vector<EF> ch;
        try {
            ch = wg::prove(c, t->ts, pd, tab.p, !mles_are_ef, (int)n_mles, (int)n_vars, prog, a, folded, fsum);
        } catch (...) {
            c->release_all();
            throw;
        }
        for (size_t i = 0; i < ch.size(); i++)
            if (challenges_out) wh::from_mont_ef(ch[i], challenges_out + 8 * i);
        if (sum_out) wh::from_mont_ef(fsum, sum_out);
        if (folded_out) {
            wk::convert_from_mont((uint32_t *)folded.p, folded.n * 8, c->stream);
            HIP_CHECK(hipMemcpyAsync(folded_out, folded.p, folded.n * 32, hipMemcpyDeviceToHost, c->stream));
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return WL_OK;
    } catch (const std::exception &e) {
        wl_set_error(e.what());
        return WL_ERR_DEVICE;
    }
}

}  // extern "C"
