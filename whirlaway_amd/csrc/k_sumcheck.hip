// k_sumcheck.hip -- HIP kernels for the zerocheck (univariate skip), the
// inner product sumcheck, multilinear folds and eq tables.
//
// Reference behaviour (file:line in the reference repository):
//   zerocheck round 0     crates/sumcheck/src/prove.rs:131-188 (skip LDE, z = 2^k .. 3(2^k - 1))
//   rounds >= 1           crates/sumcheck/src/prove.rs:138-188 (z in {0,2,3}, z = 1 derived)
//   compute_over_hypercube crates/sumcheck/src/prove.rs:256-319
//   folds                 crates/utils/src/multilinear.rs:10-123 (most-significant variables)
//   columns up/down       crates/air/src/utils.rs:140-160 (index views, never materialized)
//   linear combination    crates/utils/src/multilinear.rs:125-147
//   matrix_*_folded       crates/air/src/uni_skip_utils.rs:5-34
// All sums are exact in F_p, so block/tree reassociation is bit-identical to
// the sequential definition.
#include "common.h"
#include "kernels.h"
#include <stdexcept>

namespace wk {
using namespace kb;

// ----------------------------------------------------------------- basic --
__global__ void k_to_mont(uint32_t *d, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = kb::to_mont(d[i]).v;
}
__global__ void k_from_mont(uint32_t *d, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = kb::from_mont(F{d[i]});
}
static inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void convert_to_mont(uint32_t *d, size_t n, hipStream_t s) {
    if (n) k_to_mont<<<nblk(n), 256, 0, s>>>(d, n);
}
void convert_from_mont(uint32_t *d, size_t n, hipStream_t s) {
    if (n) k_from_mont<<<nblk(n), 256, 0, s>>>(d, n);
}

__global__ void k_zero(uint32_t *d, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) d[i] = 0;
}
void fill_zero(void *d, size_t bytes, hipStream_t s) {
    size_t n = bytes / 4;
    if (!n) return;
    unsigned g = nblk(n);
    if (g > 65536) g = 65536;
    k_zero<<<g, 256, 0, s>>>((uint32_t *)d, n);
}

__global__ void k_reduce_parts(const EF *in, int n_parts, EF *out) {
    __shared__ EF lds[16];
    const EF *p = in + (size_t)blockIdx.x * n_parts;
    EF acc = ezero();
    const int st = blockDim.x;
    int i = threadIdx.x;
    for (; i + 3 * st < n_parts; i += 4 * st) {
        const EF a = p[i], b = p[i + st], c = p[i + 2 * st], d = p[i + 3 * st];
        acc = eadd(acc, eadd(eadd(a, b), eadd(c, d)));
    }
    for (; i < n_parts; i += st) acc = eadd(acc, p[i]);
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}
void reduce_parts(const EF *in, int n_items, int n_parts, EF *out, hipStream_t s) {
    k_reduce_parts<<<n_items, 256, 0, s>>>(in, n_parts, out);
}

// blocks copy interleaved 16-byte pieces, taking words below n_canon out of
// Montgomery form on the way; with several blocks the last one to finish
// (counter, left at 0) stores the flag
__device__ __forceinline__ uint32_t pub_word(uint32_t v, size_t i, size_t n_canon) {
    return i < n_canon ? from_mont(F{v}) : v;
}
__global__ void k_publish(const uint32_t *src, size_t n, uint32_t *dst, volatile uint32_t *flag, uint32_t seq,
                          uint32_t *counter, size_t n_canon) {
    const size_t n4 = n / 4, stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool al = ((((uintptr_t)src) | ((uintptr_t)dst)) & 15) == 0;
    if (al) {
        for (size_t i = t; i < n4; i += stride) {
            uint4 v = ((const uint4 *)src)[i];
            if (4 * i < n_canon) {
                v.x = pub_word(v.x, 4 * i, n_canon);
                v.y = pub_word(v.y, 4 * i + 1, n_canon);
                v.z = pub_word(v.z, 4 * i + 2, n_canon);
                v.w = pub_word(v.w, 4 * i + 3, n_canon);
            }
            ((uint4 *)dst)[i] = v;
        }
        for (size_t i = 4 * n4 + t; i < n; i += stride) dst[i] = pub_word(src[i], i, n_canon);
    } else {
        for (size_t i = t; i < n; i += stride) dst[i] = pub_word(src[i], i, n_canon);
    }
    // Every wave's system fence waits for its own stores; after the barrier thread 0
    // releases them to the host with its own fence and raises the flag.  (Each fence
    // also writes L2 back, ~0.15 us per wave, so the launch gives a lone block only
    // as many waves as its data.)
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (gridDim.x > 1) {
        if (atomicAdd(counter, 1u) != gridDim.x - 1) return;
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence_system();
    *flag = seq;
}
void publish(const uint32_t *src, size_t n, uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq,
             uint32_t *counter, hipStream_t s, size_t n_canon) {
    // one block up to 16 KiB, then a block per 16 KiB, at most 64
    const int blocks = (int)std::min<size_t>(64, std::max<size_t>(1, (n + 4095) / 4096));
    // one thread per four 16-byte words: 256 threads for a block's 16 KiB, and a lone
    // block in whole waves
    const int threads = blocks > 1 ? 256 : (int)std::min<size_t>(1024, std::max<size_t>(64, ((n + 15) / 16 + 63) / 64 * 64));
    k_publish<<<blocks, threads, 0, s>>>(src, n, host_dst, flag, seq, counter, n_canon);
}
// rows of row_words words, stride_words apart in device memory, packed into host memory
__global__ void k_publish_rows(const uint32_t *src, int rows, int row_words, size_t stride_words, uint32_t *dst,
                               volatile uint32_t *flag, uint32_t seq) {
    const int n = rows * row_words;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        dst[i] = src[(size_t)(i / row_words) * stride_words + (size_t)(i % row_words)];
    __threadfence_system();  // as in k_publish
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        *flag = seq;
    }
}
void publish_rows(const uint32_t *src, int rows, int row_words, size_t stride_words, uint32_t *host_dst,
                  volatile uint32_t *flag, uint32_t seq, hipStream_t s) {
    const int n = rows * row_words;  // eight words per thread, in whole waves
    const int threads = std::min(1024, std::max(64, ((n + 7) / 8 + 63) / 64 * 64));
    k_publish_rows<<<1, threads, 0, s>>>(src, rows, row_words, stride_words, host_dst, flag, seq);
}

// one block: item i summed by wave i mod (waves per block), no block barriers,
// written to dst
__device__ __forceinline__ void reduce_items(const EF *in, int n_items, int n_parts, EF *dst) {
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int it = threadIdx.x >> 6; it < n_items; it += nw) {
        const EF *p = in + (size_t)it * n_parts;
        EF acc = ezero();
        int i = lane;
        // four loads in flight per lane: a large round leaves 16 parts per lane
        for (; i + 192 < n_parts; i += 256) {
            const EF a = p[i], b = p[i + 64], c = p[i + 128], d = p[i + 192];
            acc = eadd(acc, eadd(eadd(a, b), eadd(c, d)));
        }
        for (; i < n_parts; i += 64) acc = eadd(acc, p[i]);
        acc = wave_sum(acc);
        if (lane == 0) dst[it] = acc;
    }
}
// One block: the item sums written to host memory, then the flag.  Up to
// REDUCE_LDS_ITEMS sums go through LDS and wave 0 alone stores and fences them (a
// system fence also writes L2 back, ~0.15 us per wave that issues one); more sums
// are stored and fenced by the waves that made them.  Thread 0 then releases them
// with its own fence and raises the flag.
constexpr int REDUCE_LDS_ITEMS = 64;
__device__ __forceinline__ void reduce_and_publish(const EF *in, int n_items, int n_parts, EF *dst,
                                                   volatile uint32_t *flag, uint32_t seq) {
    __shared__ EF sums[REDUCE_LDS_ITEMS];
    if (n_items <= REDUCE_LDS_ITEMS) {
        reduce_items(in, n_items, n_parts, sums);
        __syncthreads();
        if (threadIdx.x < 64) {
            for (int i = threadIdx.x; i < n_items; i += 64) dst[i] = sums[i];
            __threadfence_system();
        }
    } else {
        reduce_items(in, n_items, n_parts, dst);
        __threadfence_system();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        *flag = seq;
    }
}
__global__ void k_reduce_publish(const EF *in, int n_items, int n_parts, EF *dst, volatile uint32_t *flag,
                                 uint32_t seq) {
    reduce_and_publish(in, n_items, n_parts, dst, flag, seq);
}
// thread 0 waits (bounded) until *aflag >= id, then the block copies `words` words
// of the host value into device memory; on timeout id goes to *err instead
__device__ void await_copy(const uint32_t *aflag, uint32_t id, const uint32_t *src, uint32_t *dst, int words,
                           uint64_t timeout_ticks, uint32_t *err) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        int good = 1;
        while ((int)(__hip_atomic_load(aflag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) - id) < 0) {
            __builtin_amdgcn_s_sleep(8);
            if (wall_clock64() - t0 > timeout_ticks) { good = 0; break; }
        }
        ok = good;
    }
    __syncthreads();
    if (!ok) {
        if (threadIdx.x == 0) __hip_atomic_store(err, id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    for (int i = threadIdx.x; i < words; i += blockDim.x)
        dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// k_reduce_publish followed by the wait for the host's answer (the next round's
// challenge), in one launch: a sumcheck round then costs one kernel besides its
// evaluation kernel.
__global__ void k_reduce_publish_await(const EF *in, int n_items, int n_parts, EF *dst, volatile uint32_t *flag,
                                       uint32_t seq, const uint32_t *aflag, uint32_t id, const uint32_t *src,
                                       uint32_t *adst, int words, uint64_t timeout_ticks, uint32_t *err) {
    reduce_and_publish(in, n_items, n_parts, dst, flag, seq);
    await_copy(aflag, id, src, adst, words, timeout_ticks, err);
}
void reduce_publish_await(const EF *in, int n_items, int n_parts, EF *host_dst, volatile uint32_t *flag,
                          uint32_t seq, const uint32_t *aflag, uint32_t id, const uint32_t *src, uint32_t *dst,
                          int words, uint64_t timeout_ticks, uint32_t *err, hipStream_t s) {
    k_reduce_publish_await<<<1, 256, 0, s>>>(in, n_items, n_parts, host_dst, flag, seq, aflag, id, src, dst, words,
                                             timeout_ticks, err);
}
void reduce_publish(const EF *in, int n_items, int n_parts, EF *host_dst, volatile uint32_t *flag,
                    uint32_t seq, hipStream_t s) {
    k_reduce_publish<<<1, 256, 0, s>>>(in, n_items, n_parts, host_dst, flag, seq);
}

// --------------------------------------------------------------- eq table --
// eq(point, .) * scale in one launch.  Block b covers entries b * 2^lo + t
// (lo = min(n, 8)): the table over the lo low variables is built by doublings in
// LDS while a product tree over 32 slots forms the block's factor from the high
// variables (bits of b); entry = factor * low[t].  Latency ~max(lo, 5) EF products.
__device__ __forceinline__ void eq_block(const EqPoint &pt, int lo, uint32_t b, EF *out) {
    __shared__ EF ta[256], tb[256], ha[32], hb[32], pl[32];
    const int t = threadIdx.x, hi = pt.n - lo;
    // every coordinate is loaded once, all in parallel, before the first barrier
    // (a load per level would put one memory round trip per level on the chain)
    const EF mine = t < pt.n ? pt.p[t] : ezero();
    if (t >= hi && t < pt.n) pl[t - hi] = mine;
    if (t < 32) {
        EF f = eone();
        if (t < hi) f = ((b >> (hi - 1 - t)) & 1) ? mine : esub(eone(), mine);
        if (t == 0) f = emul(f, pt.scale);
        ha[t] = f;
    }
    if (t == 0) ta[0] = eone();
    __syncthreads();
    EF *src = ta, *dst = tb, *hs = ha, *hd = hb;
    const int levels = lo > 5 ? lo : 5;
    for (int l = 0; l < levels; l++) {
        if (l < lo && t < (1 << l)) {
            const EF p = pl[l];
            const EF v = src[t];
            const EF h = emul(v, p);
            dst[2 * t] = esub(v, h);
            dst[2 * t + 1] = h;
        }
        if (l < 5 && t >= 128 && t < 128 + (16 >> l)) {
            const int i = t - 128;
            hd[i] = emul(hs[2 * i], hs[2 * i + 1]);
        }
        __syncthreads();
        if (l < lo) { EF *x = src; src = dst; dst = x; }
        if (l < 5) { EF *x = hs; hs = hd; hd = x; }
    }
    if (t < (1 << lo)) out[((size_t)b << lo) + t] = emul(hs[0], src[t]);
}
__global__ void __launch_bounds__(256) k_eq_one(EqPoint pt, int lo, EF *out) { eq_block(pt, lo, blockIdx.x, out); }
void eq_table_v(const EqPoint &pt, EF *out, hipStream_t s) {
    const int n = pt.n;
    const int lo = n < 8 ? n : 8;
    k_eq_one<<<1u << (n - lo), 256, 0, s>>>(pt, lo, out);
}
// several tables in one launch: table j uses blocks (x, j) with x < 2^(n_j - lo_j)
__global__ void __launch_bounds__(256) k_eq_batch(const EqPoint *pts, EF *const *outs) {
    const EqPoint &pt = pts[blockIdx.y];
    const int lo = pt.n < 8 ? pt.n : 8;
    if (blockIdx.x >= (1u << (pt.n - lo))) return;
    eq_block(pt, lo, blockIdx.x, outs[blockIdx.y]);
}
void eq_tables(const EqPoint *pts, EF *const *outs, int count, int max_n, hipStream_t s) {
    const int lo = max_n < 8 ? max_n : 8;
    if (count > 0) k_eq_batch<<<dim3(1u << (max_n - lo), count), 256, 0, s>>>(pts, outs);
}


// --------------------------------------------------------- AIR evaluation --
// Poseidon2Air<16, x^3, 4 + 20 + 4 rounds> constraints (p3-poseidon2-air):
// export*(export-1); full rounds: state_i - post_i; partial: s0^3 - post_sbox.
// `get(c)` yields column c of the local row.
template <class Get>
__device__ __forceinline__ EF air_eval_f(Get get, const AirConsts *air, const EF *ap) {
    AccF acc;
    int k = 0;
    {
        F e = get(0);
        acc.add(ap[k++], fmul(e, fsub(e, fone())));
    }
    F s[16];
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = get(1 + i);
    mds_light<16>(s);
    int col = 17;
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = fcube(fadd(s[i], air->rc[r * 16 + i]));
        mds_light<16>(s);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            F post = get(col + i);
            acc.add(ap[k + i], fsub(s[i], post));
            s[i] = post;
        }
        k += 16;
        col += 16;
    }
#pragma unroll 1
    for (int r = 0; r < 20; r++) {
        F t = fcube(fadd(s[0], air->rc[64 + r]));
        F post = get(col);
        acc.add(ap[k], fsub(t, post));
        k++;
        col++;
        s[0] = post;
        internal_layer_fast<16>(s);
    }
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = fcube(fadd(s[i], air->rc[84 + r * 16 + i]));
        mds_light<16>(s);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            F post = get(col + i);
            acc.add(ap[k + i], fsub(s[i], post));
            s[i] = post;
        }
        k += 16;
        col += 16;
    }
    return acc.get();
}

// -------------------------------------------------------- zerocheck round 0 --
__global__ void __launch_bounds__(256) k_zc_round0(const F *trace, int log_n, int skips,
                                                  const F *sel, const AirConsts *air,
                                                  const EF *ap, const EF *eq, EF *parts) {
    const size_t N = (size_t)1 << log_n;
    const int K = 1 << skips;
    const size_t M = N >> skips;
    const int zi = blockIdx.y;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    EF val = ezero();
    if (x < M) {
        const F *sv = sel + zi * K;
        auto get = [&](int c) -> F {
            const F *col = trace + (size_t)c * N;
            uint64_t a0 = 0;
            F r = fzero();
            for (int j = 0; j < K; j++) {
                size_t idx = (size_t)j * M + x;
                if (idx == N - 1) idx = N - 2; /* column_up: last row repeats the one before */
                a0 += (uint64_t)sv[j].v * col[idx].v;
                if ((j & 3) == 3) { r = fadd(r, F{redc64(a0)}); a0 = 0; }
            }
            if (K & 3) r = fadd(r, F{redc64(a0)});
            return r;
        };
        EF rr = air_eval_f(get, air, ap);
        val = emul(rr, eq[x]);
    }
    __shared__ EF lds[16];
    EF t = block_sum(val, lds);
    if (threadIdx.x == 0) parts[(size_t)zi * gridDim.x + blockIdx.x] = t;
}

int zc_round0(const F *trace, int log_n, int skips, const F *sel, int nz, const AirConsts *air,
              const EF *ap, const EF *eq, EF *parts, hipStream_t s) {
    size_t M = (size_t)1 << (log_n - skips);
    unsigned nb = nblk(M);
    k_zc_round0<<<dim3(nb, nz), 256, 0, s>>>(trace, log_n, skips, sel, air, ap, eq, parts);
    return (int)nb;
}

// Round 0 in two passes.  Pass 1: for each column c and x, the 2^skips values
// up_c[j*M + x] are those of a polynomial of degree < K = 2^skips in the skipped
// variable at j = 0..K-1; its value at z is sum_j L_j(z) up_c[j*M + x] with the
// Lagrange basis L_j of the nodes 0..K-1 (sel[zi][j] = L_j(K + zi), Montgomery):
// K products into one 64-bit sum (a fold after every four), one reduction.
// U0[zi][c][x] = f_c,x(K + zi), zi < nz.
template <int K>
__global__ void __launch_bounds__(256) k_zc_lde(const F *__restrict__ trace, int log_n, int nz,
                                                const F *__restrict__ sel, F *__restrict__ U0) {
    const size_t N = (size_t)1 << log_n;
    const size_t M = N / K;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (x >= M) return;
    const F *col = trace + (size_t)c * N;
    F v[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        size_t idx = (size_t)j * M + x;
        if (idx == N - 1) idx = N - 2;  // column_up: the last row repeats the one before
        v[j] = col[idx];
    }
    F *out = U0 + (size_t)c * M + x;
    const size_t zstride = (size_t)AIR_NCOLS * M;
#pragma unroll 2
    for (int zi = 0; zi < nz; zi++) {
        const F *L = sel + (size_t)zi * K;
        uint64_t a = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j && (j & 3) == 0) a = fold64(a);
            a += (uint64_t)v[j].v * L[j].v;
        }
        out[(size_t)zi * zstride] = F{redc64(a)};
    }
}

// Pass 2: one lane per (x, z): AIR on the extended row, times eq[x], block sums.
__global__ void __launch_bounds__(256) k_zc_air0(const F *U0, size_t M, const AirConsts *air,
                                                const EF *ap, const EF *eq, EF *parts) {
    const int zi = blockIdx.y;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    EF val = ezero();
    if (x < M) {
        const F *row = U0 + (size_t)zi * AIR_NCOLS * M + x;
        auto get = [&](int c) -> F { return row[(size_t)c * M]; };
        val = emul(air_eval_f(get, air, ap), eq[x]);
    }
    __shared__ EF lds[16];
    EF t = block_sum(val, lds);
    if (threadIdx.x == 0) parts[(size_t)zi * gridDim.x + blockIdx.x] = t;
}

int zc_round0_lde(const F *trace, int log_n, int skips, int nz, const F *sel, const AirConsts *air, const EF *ap,
                  const EF *eq, F *U0, EF *parts, hipStream_t s) {
    const size_t M = (size_t)1 << (log_n - skips);
    const dim3 g1(nblk(M), AIR_NCOLS);
    switch (skips) {
        case 1: k_zc_lde<2><<<g1, 256, 0, s>>>(trace, log_n, nz, sel, U0); break;
        case 2: k_zc_lde<4><<<g1, 256, 0, s>>>(trace, log_n, nz, sel, U0); break;
        case 3: k_zc_lde<8><<<g1, 256, 0, s>>>(trace, log_n, nz, sel, U0); break;
        case 4: k_zc_lde<16><<<g1, 256, 0, s>>>(trace, log_n, nz, sel, U0); break;
        default: return -1;
    }
    const unsigned nb = nblk(M);
    k_zc_air0<<<dim3(nb, nz), 256, 0, s>>>(U0, M, air, ap, eq, parts);
    return (int)nb;
}

__global__ void k_zc_fold0(const F *trace, int log_n, int skips, const EF *S, EF *U) {
    const size_t N = (size_t)1 << log_n;
    const int K = 1 << skips;
    const size_t M = N >> skips;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (x >= M) return;
    const F *col = trace + (size_t)c * N;
    AccF acc;
    int j = 0;
    for (; j + 8 <= K; j += 8) {  // eight loads in flight per thread
        F v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            size_t idx = (size_t)(j + q) * M + x;
            if (idx == N - 1) idx = N - 2;
            v[q] = col[idx];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc.add(S[j + q], v[q]);
    }
    for (; j < K; j++) {
        size_t idx = (size_t)j * M + x;
        if (idx == N - 1) idx = N - 2;
        acc.add(S[j], col[idx]);
    }
    U[(size_t)c * M + x] = acc.get();
}
void zc_fold0(const F *trace, int log_n, int skips, const EF *S, EF *U, hipStream_t s) {
    size_t M = (size_t)1 << (log_n - skips);
    k_zc_fold0<<<dim3(nblk(M), AIR_NCOLS), 256, 0, s>>>(trace, log_n, skips, S, U);
}

// Wait for a value the host writes after this kernel was enqueued: thread 0 polls
// *flag (host-mapped, system scope) until it reaches id, then the block copies
// the words into device memory for the kernels queued behind it.  The wait is
// bounded by timeout_ticks of the constant-rate clock; on expiry id is written
// to *err and nothing is copied.
__global__ void k_await_host(const uint32_t *flag, uint32_t id, const uint32_t *src, uint32_t *dst, int words,
                             uint64_t timeout_ticks, uint32_t *err) {
    await_copy(flag, id, src, dst, words, timeout_ticks, err);
}
void await_host(const uint32_t *flag, uint32_t id, const uint32_t *src, uint32_t *dst, int words,
                uint64_t timeout_ticks, uint32_t *err, hipStream_t s) {
    k_await_host<<<1, 64, 0, s>>>(flag, id, src, dst, words, timeout_ticks, err);
}

// column c < n_cols: fold by (1 - ch, ch), ch read from device memory (a value
// an await_host launch placed there); the extra row blockIdx.y == n_cols halves
// the eq table for the next round.
__global__ void k_fold_cols_p(EF *U, size_t stride, int l, int n_cols, const EF *chp, EF *E) {
    const size_t h = (size_t)1 << (l - 1);
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    if ((int)blockIdx.y == n_cols) {
        size_t q = h >> 1;
        if (x < q) E[x] = eadd(E[x], E[x + q]);
        return;
    }
    const EFm ch = emul_prep(*chp);
    EF *col = U + (size_t)blockIdx.y * stride;
    EF a = col[x];
    col[x] = eadd(a, emulm(esub(col[x + h], a), ch));
}
void fold_cols_p(EF *U, int n_cols, size_t stride, int l, const EF *c, EF *E, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_cols_p<<<dim3(nblk(h), n_cols + (E && h > 1 ? 1 : 0)), 256, 0, s>>>(U, stride, l, n_cols, c, E);
}

// down inner sums: a block takes DOWN_CPB columns of one segment j, so each eq
// weight it loads serves that many columns, with two elements per thread in flight
constexpr int DOWN_CPB = 4;
__global__ void __launch_bounds__(256) k_zc_down_parts(const F *trace, int log_n, int skips, const EF *eqr,
                                                       EF *parts) {
    const size_t N = (size_t)1 << log_n;
    const size_t M = N >> skips;
    const int c0 = blockIdx.x * DOWN_CPB, j = blockIdx.y, K = 1 << skips;
    AccF acc[DOWN_CPB];
    const size_t bd = blockDim.x;
    auto load = [&](size_t i, F *v) {
        size_t idx = (size_t)j * M + i + 1;
        if (idx > N - 1) idx = N - 1; /* column_down: last row repeats */
#pragma unroll
        for (int q = 0; q < DOWN_CPB; q++) v[q] = c0 + q < AIR_NCOLS ? trace[(size_t)(c0 + q) * N + idx] : F{0};
    };
    size_t i = threadIdx.x;
    for (; i + bd < M; i += 2 * bd) {
        F v0[DOWN_CPB], v1[DOWN_CPB];
        load(i, v0);
        load(i + bd, v1);
        const EF e0 = eqr[i], e1 = eqr[i + bd];
#pragma unroll
        for (int q = 0; q < DOWN_CPB; q++) {
            acc[q].add(e0, v0[q]);
            acc[q].add(e1, v1[q]);
        }
    }
    for (; i < M; i += bd) {
        F v[DOWN_CPB];
        load(i, v);
        const EF e = eqr[i];
#pragma unroll
        for (int q = 0; q < DOWN_CPB; q++) acc[q].add(e, v[q]);
    }
    __shared__ EF lds[16];
#pragma unroll
    for (int q = 0; q < DOWN_CPB; q++) {
        EF t = block_sum(acc[q].get(), lds);
        if (threadIdx.x == 0 && c0 + q < AIR_NCOLS) parts[(size_t)(c0 + q) * K + j] = t;
    }
}
void zc_down_parts(const F *trace, int log_n, int skips, const EF *eqr, EF *parts, hipStream_t s) {
    k_zc_down_parts<<<dim3((AIR_NCOLS + DOWN_CPB - 1) / DOWN_CPB, 1 << skips), 256, 0, s>>>(trace, log_n, skips,
                                                                                           eqr, parts);
}
__global__ void k_zc_down_combine(const EF *parts, int K, const EF *S, EF *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= AIR_NCOLS) return;
    EF acc = ezero();
    for (int j = 0; j < K; j++) acc = eadd(acc, emul(S[j], parts[(size_t)c * K + j]));
    out[c] = acc;
}
void zc_down_combine(const EF *parts, int skips, const EF *S, EF *out, hipStream_t s) {
    k_zc_down_combine<<<nblk(AIR_NCOLS), 256, 0, s>>>(parts, 1 << skips, S, out);
}

// ----------------------------------------------------------------- batching --
__global__ void k_lincomb(const F *trace, int log_n, int n_cols, const EF *eqc, EF *B) {
    const size_t N = (size_t)1 << log_n;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    AccF acc;
    // eight column loads in flight per thread (one at a time leaves the loop
    // waiting on memory latency)
    int c = 0;
    for (; c + 8 <= n_cols; c += 8) {
        F v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = trace[(size_t)(c + q) * N + x];
#pragma unroll
        for (int q = 0; q < 8; q++) acc.add(eqc[c + q], v[q]);
    }
    for (; c < n_cols; c++) acc.add(eqc[c], trace[(size_t)c * N + x]);
    B[x] = acc.get();
}
void lincomb_cols(const F *trace, int log_n, int n_cols, const EF *eqc, EF *B, hipStream_t s) {
    size_t N = (size_t)1 << log_n;
    k_lincomb<<<nblk(N), 256, 0, s>>>(trace, log_n, n_cols, eqc, B);
}

__global__ void k_mix(const EF *B, size_t N, const EF *a, EF *mixed) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    EF up = x == N - 1 ? B[N - 2] : B[x];
    EF dn = x == N - 1 ? B[N - 1] : B[x + 1];
    mixed[x] = eadd(up, emul(*a, dn));
}
void mix_updown(const EF *B, int log_n, const EF *a, EF *mixed, hipStream_t s) {
    size_t N = (size_t)1 << log_n;
    k_mix<<<nblk(N), 256, 0, s>>>(B, N, a, mixed);
}

__global__ void k_matrix(const EF *E, size_t N, const EF *a, EF *M) {
    const size_t y = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= N) return;
    EF r;
    if (y == 0) r = E[0];
    else if (y < N - 2) r = eadd(E[y], emul(*a, E[y - 1]));
    else if (y == N - 2) r = eadd(eadd(E[N - 2], E[N - 1]), emul(*a, E[N - 3]));
    else r = emul(*a, eadd(E[N - 2], E[N - 1]));
    M[y] = r;
}
void build_matrix(const EF *E, int log_n, const EF *a, EF *M, hipStream_t s) {
    size_t N = (size_t)1 << log_n;
    k_matrix<<<nblk(N), 256, 0, s>>>(E, N, a, M);
}

template <bool BASE>
__global__ void k_seg_dot(const void *Av, const EF *E, size_t len, EF *parts) {
    const int j = blockIdx.y;
    EF acc = ezero();
    AccF accf;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += st) {
        if (BASE) accf.add(E[i], ((const F *)Av)[(size_t)j * len + i]);
        else acc = eadd(acc, emul(((const EF *)Av)[(size_t)j * len + i], E[i]));
    }
    if (BASE) acc = accf.get();
    __shared__ EF lds[16];
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[(size_t)j * gridDim.x + blockIdx.x] = t;
}
int dot_blocks(size_t len) {
    // one element per thread up to 1024 blocks (EF products are latency-bound per
    // thread; wider grids hide it), grid-stride beyond
    unsigned nb = nblk(len, 256);
    if (nb > 1024) nb = 1024;
    if (nb == 0) nb = 1;
    return (int)nb;
}
int seg_dot(const EF *A, const EF *E, size_t len, int nseg, EF *parts, hipStream_t s) {
    unsigned nb = dot_blocks(len);
    k_seg_dot<false><<<dim3(nb, nseg), 256, 0, s>>>(A, E, len, parts);
    return (int)nb;
}

// One round of the product sumcheck on (A, B) of length 2h, fused with the fold
// by the previous challenge c when `fold` is set (the tables then hold 4h
// entries and are folded in place to 2h: thread x owns x, x+h, x+2h, x+3h), so
// each table is read once per round.  parts[z][block] holds the block sums of
// A(z) B(z) for z = 0 and 2 (z = 1 follows from the claimed sum).
__global__ void __launch_bounds__(256) k_prod_round_f(EF *A, EF *B, size_t h, int fold, const EF *cp,
                                                     EF *parts) {
    EF a0 = ezero(), a2 = ezero();
    const size_t st = (size_t)gridDim.x * blockDim.x;
    const EFm c = emul_prep(fold ? *cp : ezero());
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < h; x += st) {
        EF al = A[x], ah = A[x + h], bl = B[x], bh = B[x + h];
        if (fold) {
            al = eadd(al, emulm(esub(A[x + 2 * h], al), c));
            ah = eadd(ah, emulm(esub(A[x + 3 * h], ah), c));
            bl = eadd(bl, emulm(esub(B[x + 2 * h], bl), c));
            bh = eadd(bh, emulm(esub(B[x + 3 * h], bh), c));
            A[x] = al;
            A[x + h] = ah;
            B[x] = bl;
            B[x + h] = bh;
        }
        a0 = eadd(a0, emul(al, bl));
        a2 = eadd(a2, emul(esub(edbl(ah), al), esub(edbl(bh), bl)));
    }
    __shared__ EF lds[16];
    EF t0 = block_sum(a0, lds);
    EF t2 = block_sum(a2, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = t0;
        parts[gridDim.x + blockIdx.x] = t2;
    }
}
int prod_round_fused(EF *A, EF *B, int l, bool fold, const EF *c, EF *parts, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = dot_blocks(h);
    k_prod_round_f<<<nb, 256, 0, s>>>(A, B, h, fold ? 1 : 0, c, parts);
    return (int)nb;
}
__global__ void k_fold_ef2_p(EF *A, EF *B, size_t h, const EF *cp) {
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    const EFm c = emul_prep(*cp);
    EF *T = blockIdx.y ? B : A;
    EF a = T[x];
    T[x] = eadd(a, emulm(esub(T[x + h], a), c));
}
void fold_ef2_p(EF *A, EF *B, int l, const EF *c, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_ef2_p<<<dim3(nblk(h), B ? 2 : 1), 256, 0, s>>>(A, B, h, c);
}
__global__ void k_fold_ef(EF *A, size_t h, const EF *c) {
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    EF a = A[x];
    A[x] = eadd(a, emulm(esub(A[x + h], a), emul_prep(*c)));
}
void fold_ef(EF *A, int l, const EF *c, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_ef<<<nblk(h), 256, 0, s>>>(A, h, c);
}
// nf in-place folds of an EF table of 2^l <= 2^12 entries in one block: the first
// reads the table, the rest work from LDS; every fold's result is also written to
// A[0 .. len), the memory state a chain of fold_ef launches leaves.  Saves a launch
// (and its dispatch gap) per fold where each fold is a few microseconds of latency.
constexpr int FOLD_TAIL_LOG = 12;
__global__ void __launch_bounds__(1024) k_fold_ef_tail(EF *A, int l, int nf, const EF *c) {
    __shared__ EF buf[1 << (FOLD_TAIL_LOG - 1)];
    int h = 1 << (l - 1);
    const EFm c0 = emul_prep(c[0]);
    for (int x = threadIdx.x; x < h; x += blockDim.x) {
        const EF a = A[x], r = eadd(a, emulm(esub(A[x + h], a), c0));
        buf[x] = r;
        A[x] = r;
    }
    __syncthreads();
    for (int i = 1; i < nf; i++) {
        h >>= 1;
        const int x = threadIdx.x;
        EF r;
        if (x < h) {
            const EF a = buf[x];
            r = eadd(a, emulm(esub(buf[x + h], a), emul_prep(c[i])));
        }
        __syncthreads();
        if (x < h) {
            buf[x] = r;
            A[x] = r;
        }
        __syncthreads();
    }
}
void fold_ef_tail(EF *A, int l, int nf, const EF *c, hipStream_t s) {
    if (l < 1 || l > FOLD_TAIL_LOG || nf < 1 || nf > l) throw std::runtime_error("fold_ef_tail: bad table size");
    k_fold_ef_tail<<<1, 1024, 0, s>>>(A, l, nf, c);
}

// -------------------------------------------------------------------- micro --
__global__ void k_fold_f_large(const F *in, size_t M, const EF *sc, int K, EF *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    AccF acc;
    for (int j = 0; j < K; j++) acc.add(sc[j], in[i + (size_t)j * M]);
    out[i] = acc.get();
}
void fold_f_large(const F *in, int l, const EF *sc, int k, EF *out, hipStream_t s) {
    size_t M = (size_t)1 << (l - k);
    k_fold_f_large<<<nblk(M), 256, 0, s>>>(in, M, sc, 1 << k, out);
}
__global__ void k_fold_ef_large(const EF *in, size_t M, const EF *sc, int K, EF *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    EF acc = ezero();
    for (int j = 0; j < K; j++) acc = eadd(acc, emul(sc[j], in[i + (size_t)j * M]));
    out[i] = acc;
}
void fold_ef_large(const EF *in, int l, const EF *sc, int k, EF *out, hipStream_t s) {
    size_t M = (size_t)1 << (l - k);
    k_fold_ef_large<<<nblk(M), 256, 0, s>>>(in, M, sc, 1 << k, out);
}
__global__ void k_fold_f_to_ef(const F *in, size_t h, const EF *c, EF *out) {
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    F a = in[x], b = in[x + h];
    EF r = emulf(*c, fsub(b, a));
    r.c[0] = fadd(r.c[0], a);
    out[x] = r;
}
void fold_f_to_ef(const F *in, int l, const EF *c, EF *out, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_f_to_ef<<<nblk(h), 256, 0, s>>>(in, h, c, out);
}

// eq-weighted a^2 b: parts[zi][b] for z = 0, 2, 3 (gridDim.y == 3) or z = 0..3 (== 4)
template <class T>
__device__ __forceinline__ T zfold(T lo, T hi, int z);
template <>
__device__ __forceinline__ F zfold<F>(F lo, F hi, int z) {
    if (z == 0) return lo;
    if (z == 1) return hi;
    F d = fsub(hi, lo), r = fadd(lo, fdbl(d));
    return z == 2 ? r : fadd(r, d);
}
template <>
__device__ __forceinline__ EF zfold<EF>(EF lo, EF hi, int z) {
    if (z == 0) return lo;
    if (z == 1) return hi;
    EF d = esub(hi, lo), r = eadd(lo, edbl(d));
    return z == 2 ? r : eadd(r, d);
}

template <bool BASE>
__global__ void k_a2b_round(const void *av, const void *bv, size_t h, const EF *eq, EF *parts) {
    const int zi = blockIdx.y;
    const int z = gridDim.y == 4 ? zi : (zi == 0 ? 0 : zi + 1);
    EF acc = ezero();
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < h; x += st) {
        EF v;
        if (BASE) {
            const F *a = (const F *)av, *b = (const F *)bv;
            F az = zfold<F>(a[x], a[x + h], z), bz = zfold<F>(b[x], b[x + h], z);
            v = emulf(eq[x], fmul(fmul(az, az), bz));
        } else {
            const EF *a = (const EF *)av, *b = (const EF *)bv;
            EF az = zfold<EF>(a[x], a[x + h], z), bz = zfold<EF>(b[x], b[x + h], z);
            v = emul(eq[x], emul(esqr(az), bz));
        }
        acc = eadd(acc, v);
    }
    __shared__ EF lds[16];
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[(size_t)zi * gridDim.x + blockIdx.x] = t;
}
int a2b_round_f(const F *a, const F *b, int l, const EF *eq, EF *parts, hipStream_t s, bool all4) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = dot_blocks(h);
    k_a2b_round<true><<<dim3(nb, all4 ? 4 : 3), 256, 0, s>>>(a, b, h, eq, parts);
    return (int)nb;
}
// Rounds >= 1 of the a^2 b sumcheck in one pass per round: fold the previous
// round's tables (length 4h) by its challenge c into tables of length 2h (from
// the base-field inputs a0, b0 in round 1, in place afterwards), halve the eq
// table (eq(z[j+1..], x) = eq(z[j..], (0, x)) + eq(z[j..], (1, x))), and sum
// eq * A_z^2 B_z for z = 0, 2, 3 with one read of every table.
template <bool FIRST>
__global__ void __launch_bounds__(256) k_a2b_round_fused(const F *a0, const F *b0, EF *A, EF *B, size_t h,
                                                         EF c, EF *eq, EF *parts) {
    EF acc[3] = {ezero(), ezero(), ezero()};
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < h; x += st) {
        EF al, ah, bl, bh;
        if (FIRST) {
            auto fl = [&](const F *t, size_t i) { return eadd(efrom(t[i]), emulf(c, fsub(t[i + 2 * h], t[i]))); };
            al = fl(a0, x); ah = fl(a0, x + h); bl = fl(b0, x); bh = fl(b0, x + h);
        } else {
            const EFm cm = emul_prep(c);
            auto fl = [&](const EF *t, size_t i) { return eadd(t[i], emulm(esub(t[i + 2 * h], t[i]), cm)); };
            al = fl(A, x); ah = fl(A, x + h); bl = fl(B, x); bh = fl(B, x + h);
        }
        A[x] = al; A[x + h] = ah; B[x] = bl; B[x + h] = bh;
        const EF e = eadd(eq[x], eq[x + h]);
        eq[x] = e;
        const EF da = esub(ah, al), db = esub(bh, bl);
        // z = 0
        acc[0] = eadd(acc[0], emul(e, emul(esqr(al), bl)));
        // z = 2: lo + 2 (hi - lo); z = 3: one more difference
        EF a2 = eadd(al, edbl(da)), b2 = eadd(bl, edbl(db));
        acc[1] = eadd(acc[1], emul(e, emul(esqr(a2), b2)));
        EF a3 = eadd(a2, da), b3 = eadd(b2, db);
        acc[2] = eadd(acc[2], emul(e, emul(esqr(a3), b3)));
    }
    __shared__ EF lds[16];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        EF t = block_sum(acc[i], lds);
        if (threadIdx.x == 0) parts[(size_t)i * gridDim.x + blockIdx.x] = t;
    }
}
int a2b_round_fused(const F *a0, const F *b0, EF *A, EF *B, int l, bool first, EF c, EF *eq, EF *parts,
                    hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = dot_blocks(h);
    if (first) k_a2b_round_fused<true><<<nb, 256, 0, s>>>(a0, b0, A, B, h, c, eq, parts);
    else k_a2b_round_fused<false><<<nb, 256, 0, s>>>(a0, b0, A, B, h, c, eq, parts);
    return (int)nb;
}

// The same rounds with the eq weight split: eq(z[j+1..], x) = ehi[x >> L] elo[x mod
// 2^L], elo over the last L variables (read from cache, never halved) and ehi
// over the others.  Block b owns the h / gridDim.x consecutive x from b h /
// gridDim.x: a power of two no larger than 2^L, so one ehi entry, by which the
// block's sums are multiplied once.  No eq table of the round's length is built,
// read or halved.
template <bool FIRST>
__global__ void __launch_bounds__(256) k_a2b_round_fused_split(const F *a0, const F *b0, EF *A, EF *B, size_t h,
                                                               int L, EF c, const EF *elo, const EF *ehi,
                                                               EF *parts) {
    EF acc[3] = {ezero(), ezero(), ezero()};
    const size_t len = h / gridDim.x, x0 = (size_t)blockIdx.x * len, mask = ((size_t)1 << L) - 1;
    for (size_t i = threadIdx.x; i < len; i += blockDim.x) {
        const size_t x = x0 + i;
        EF al, ah, bl, bh;
        if (FIRST) {
            auto fl = [&](const F *t, size_t k) { return eadd(efrom(t[k]), emulf(c, fsub(t[k + 2 * h], t[k]))); };
            al = fl(a0, x); ah = fl(a0, x + h); bl = fl(b0, x); bh = fl(b0, x + h);
        } else {
            const EFm cm = emul_prep(c);
            auto fl = [&](const EF *t, size_t k) { return eadd(t[k], emulm(esub(t[k + 2 * h], t[k]), cm)); };
            al = fl(A, x); ah = fl(A, x + h); bl = fl(B, x); bh = fl(B, x + h);
        }
        A[x] = al; A[x + h] = ah; B[x] = bl; B[x + h] = bh;
        const EF e = elo[x & mask];
        const EF da = esub(ah, al), db = esub(bh, bl);
        acc[0] = eadd(acc[0], emul(e, emul(esqr(al), bl)));
        EF a2 = eadd(al, edbl(da)), b2 = eadd(bl, edbl(db));
        acc[1] = eadd(acc[1], emul(e, emul(esqr(a2), b2)));
        EF a3 = eadd(a2, da), b3 = eadd(b2, db);
        acc[2] = eadd(acc[2], emul(e, emul(esqr(a3), b3)));
    }
    __shared__ EF lds[16];
    const EF w = ehi[x0 >> L];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        EF t = block_sum(acc[i], lds);
        if (threadIdx.x == 0) parts[(size_t)i * gridDim.x + blockIdx.x] = emul(w, t);
    }
}
// round 0 of the split form on the base-field tables, z = 0..3 (gridDim.y == 4) or
// z = 0, 2, 3
__global__ void __launch_bounds__(256) k_a2b_round_split(const F *a, const F *b, size_t h, int L, const EF *elo,
                                                         const EF *ehi, EF *parts) {
    const int zi = blockIdx.y;
    const int z = gridDim.y == 4 ? zi : (zi == 0 ? 0 : zi + 1);
    const size_t len = h / gridDim.x, x0 = (size_t)blockIdx.x * len, mask = ((size_t)1 << L) - 1;
    EF acc = ezero();
    for (size_t i = threadIdx.x; i < len; i += blockDim.x) {
        const size_t x = x0 + i;
        F az = zfold<F>(a[x], a[x + h], z), bz = zfold<F>(b[x], b[x + h], z);
        acc = eadd(acc, emulf(elo[x & mask], fmul(fmul(az, az), bz)));
    }
    __shared__ EF lds[16];
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) parts[(size_t)zi * gridDim.x + blockIdx.x] = emul(ehi[x0 >> L], t);
}
// block counts as dot_blocks gives them; each block's range must fit inside one
// ehi entry (h / blocks <= 2^L) and h > 2^L
static unsigned split_blocks(size_t h, int L) {
    const unsigned nb = (unsigned)dot_blocks(h);
    if (h <= ((size_t)1 << L) || h % nb != 0 || h / nb > ((size_t)1 << L) || (h / nb) % 256 != 0)
        throw std::runtime_error("a2b split round: table length does not fit the block layout");
    return nb;
}
int a2b_round_f_split(const F *a, const F *b, int l, int L, const EF *elo, const EF *ehi, EF *parts, hipStream_t s,
                      bool all4) {
    const size_t h = (size_t)1 << (l - 1);
    const unsigned nb = split_blocks(h, L);
    k_a2b_round_split<<<dim3(nb, all4 ? 4 : 3), 256, 0, s>>>(a, b, h, L, elo, ehi, parts);
    return (int)nb;
}
int a2b_round_fused_split(const F *a0, const F *b0, EF *A, EF *B, int l, int L, bool first, EF c, const EF *elo,
                          const EF *ehi, EF *parts, hipStream_t s) {
    const size_t h = (size_t)1 << (l - 1);
    const unsigned nb = split_blocks(h, L);
    if (first) k_a2b_round_fused_split<true><<<nb, 256, 0, s>>>(a0, b0, A, B, h, L, c, elo, ehi, parts);
    else k_a2b_round_fused_split<false><<<nb, 256, 0, s>>>(a0, b0, A, B, h, L, c, elo, ehi, parts);
    return (int)nb;
}

}  // namespace wk
