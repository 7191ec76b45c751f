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

namespace wk {
using namespace kb;

// Lazy accumulator of EF x F products: 64-bit per coefficient, reduced every 4 terms.
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
    __device__ __forceinline__ void add(const EF &a, F b) {
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += (uint64_t)a.c[i].v * b.v;
        if (++n == 4) flush();
    }
    __device__ __forceinline__ EF get() { flush(); return r; }
};

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
    for (int i = threadIdx.x; i < n_parts; i += blockDim.x) acc = eadd(acc, p[i]);
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}
void reduce_parts(const EF *in, int n_items, int n_parts, EF *out, hipStream_t s) {
    k_reduce_parts<<<n_items, 256, 0, s>>>(in, n_parts, out);
}

// --------------------------------------------------------------- eq table --
// Builds T_k over (p_k..p_{n-1}) from T_{k+1}: out[S + j] = p_k T[j], out[j] = T[j] - out[S + j].
__global__ void k_eq_small(const EF *pt, int n, const EF *scale, EF *out, int levels) {
    // one block: first `levels` doublings (variables n-1 .. n-levels)
    if (threadIdx.x == 0) out[0] = *scale;
    __syncthreads();
    size_t S = 1;
    for (int t = 0; t < levels; t++) {
        EF p = pt[n - 1 - t];
        for (size_t j = threadIdx.x; j < S; j += blockDim.x) {
            EF v = out[j];
            EF hi = emul(v, p);
            out[S + j] = hi;
            out[j] = esub(v, hi);
        }
        __syncthreads();
        S <<= 1;
    }
}
__global__ void k_eq_level(const EF *pt, int k, size_t S, EF *out) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= S) return;
    EF p = pt[k];
    EF v = out[j];
    EF hi = emul(v, p);
    out[S + j] = hi;
    out[j] = esub(v, hi);
}
void eq_table(const EF *pt, int n, const EF *scale, EF *out, hipStream_t s) {
    int first = n < 10 ? n : 10;
    k_eq_small<<<1, 512, 0, s>>>(pt, n, scale, out, first);
    size_t S = (size_t)1 << first;
    for (int t = first; t < n; t++) {
        k_eq_level<<<nblk(S), 256, 0, s>>>(pt, n - 1 - t, S, out);
        S <<= 1;
    }
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
        internal_layer<16>(s, DDIAG.d16);
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

__device__ __forceinline__ void mat4_ef(EF *x) {
    EF t01 = eadd(x[0], x[1]), t23 = eadd(x[2], x[3]);
    EF t0123 = eadd(t01, t23);
    EF t01123 = eadd(t0123, x[1]);
    EF t01233 = eadd(t0123, x[3]);
    EF y3 = eadd(t01233, edbl(x[0]));
    EF y1 = eadd(t01123, edbl(x[2]));
    EF y0 = eadd(t01123, t01);
    EF y2 = eadd(t01233, t23);
    x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
}
__device__ __forceinline__ void mds_light_ef(EF *s) {
#pragma unroll
    for (int c = 0; c < 16; c += 4) mat4_ef(s + c);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        EF acc = eadd(eadd(s[k], s[k + 4]), eadd(s[k + 8], s[k + 12]));
        s[k] = eadd(s[k], acc);
        s[k + 4] = eadd(s[k + 4], acc);
        s[k + 8] = eadd(s[k + 8], acc);
        s[k + 12] = eadd(s[k + 12], acc);
    }
}
__device__ __forceinline__ void internal_ef(EF *s) {
    EF sum = s[0];
#pragma unroll
    for (int i = 1; i < 16; i++) sum = eadd(sum, s[i]);
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = eadd(sum, emulf(s[i], F{DDIAG.d16[i]}));
}

template <class Get>
__device__ __forceinline__ EF air_eval_ef(Get get, const AirConsts *air, const EF *ap) {
    EF acc;
    int k = 0;
    {
        EF e = get(0);
        acc = emul(ap[k++], emul(e, esub(e, eone())));
    }
    EF s[16];
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = get(1 + i);
    mds_light_ef(s);
    int col = 17;
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            EF t = s[i];
            t.c[0] = fadd(t.c[0], air->rc[r * 16 + i]);
            s[i] = ecube(t);
        }
        mds_light_ef(s);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            EF post = get(col + i);
            acc = eadd(acc, emul(ap[k + i], esub(s[i], post)));
            s[i] = post;
        }
        k += 16;
        col += 16;
    }
#pragma unroll 1
    for (int r = 0; r < 20; r++) {
        EF t = s[0];
        t.c[0] = fadd(t.c[0], air->rc[64 + r]);
        t = ecube(t);
        EF post = get(col);
        acc = eadd(acc, emul(ap[k], esub(t, post)));
        k++;
        col++;
        s[0] = post;
        internal_ef(s);
    }
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            EF t = s[i];
            t.c[0] = fadd(t.c[0], air->rc[84 + r * 16 + i]);
            s[i] = ecube(t);
        }
        mds_light_ef(s);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            EF post = get(col + i);
            acc = eadd(acc, emul(ap[k + i], esub(s[i], post)));
            s[i] = post;
        }
        k += 16;
        col += 16;
    }
    return acc;
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

__global__ void k_zc_fold0(const F *trace, int log_n, int skips, const EF *S, EF *U) {
    const size_t N = (size_t)1 << log_n;
    const int K = 1 << skips;
    const size_t M = N >> skips;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (x >= M) return;
    const F *col = trace + (size_t)c * N;
    AccF acc;
    for (int j = 0; j < K; j++) {
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

// ------------------------------------------------------ zerocheck rounds >= 1 --
__global__ void __launch_bounds__(256) k_zc_round(const EF *U, size_t stride, int l,
                                                 const AirConsts *air, const EF *ap,
                                                 const EF *eq, EF *parts) {
    const size_t h = (size_t)1 << (l - 1);
    const int zi = blockIdx.y;  // z = 0, 2, 3
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    EF val = ezero();
    if (x < h) {
        auto get = [&](int c) -> EF {
            const EF *col = U + (size_t)c * stride;
            EF a = col[x];
            if (zi == 0) return a;
            EF d = esub(col[x + h], a);
            EF r = eadd(a, edbl(d));
            if (zi == 2) r = eadd(r, d);
            return r;
        };
        EF rr = air_eval_ef(get, air, ap);
        val = emul(rr, eq[x]);
    }
    __shared__ EF lds[16];
    EF t = block_sum(val, lds);
    if (threadIdx.x == 0) parts[(size_t)zi * gridDim.x + blockIdx.x] = t;
}

int zc_round_strided(const EF *U, size_t stride, int l, const AirConsts *air, const EF *ap,
                     const EF *eq, EF *parts, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = nblk(h);
    k_zc_round<<<dim3(nb, 3), 256, 0, s>>>(U, stride, l, air, ap, eq, parts);
    return (int)nb;
}

__global__ void k_fold_cols(EF *U, size_t stride, int l, const EF *c) {
    const size_t h = (size_t)1 << (l - 1);
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    EF *col = U + (size_t)blockIdx.y * stride;
    EF a = col[x];
    col[x] = eadd(a, emul(*c, esub(col[x + h], a)));
}
void fold_cols_strided(EF *U, int n_cols, size_t stride, int l, const EF *c, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_cols<<<dim3(nblk(h), n_cols), 256, 0, s>>>(U, stride, l, c);
}

// down inner sums
__global__ void k_zc_down_parts(const F *trace, int log_n, int skips, const EF *eqr, EF *parts) {
    const size_t N = (size_t)1 << log_n;
    const size_t M = N >> skips;
    const int c = blockIdx.x, j = blockIdx.y, K = 1 << skips;
    const F *col = trace + (size_t)c * N;
    AccF acc;
    for (size_t i = threadIdx.x; i < M; i += blockDim.x) {
        size_t idx = (size_t)j * M + i + 1;
        if (idx > N - 1) idx = N - 1; /* column_down: last row repeats */
        acc.add(eqr[i], col[idx]);
    }
    __shared__ EF lds[16];
    EF t = block_sum(acc.get(), lds);
    if (threadIdx.x == 0) parts[(size_t)c * K + j] = t;
}
void zc_down_parts(const F *trace, int log_n, int skips, const EF *eqr, EF *parts, hipStream_t s) {
    k_zc_down_parts<<<dim3(AIR_NCOLS, 1 << skips), 256, 0, s>>>(trace, log_n, skips, eqr, parts);
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
    for (int c = 0; c < n_cols; c++) acc.add(eqc[c], trace[(size_t)c * N + x]);
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
    unsigned nb = nblk(len, 256 * 4);
    if (nb > 1024) nb = 1024;
    if (nb == 0) nb = 1;
    return (int)nb;
}
int seg_dot(const EF *A, const EF *E, size_t len, int nseg, EF *parts, hipStream_t s) {
    unsigned nb = dot_blocks(len);
    k_seg_dot<false><<<dim3(nb, nseg), 256, 0, s>>>(A, E, len, parts);
    return (int)nb;
}
int seg_dot_f(const F *A, const EF *E, size_t len, int nseg, EF *parts, hipStream_t s) {
    unsigned nb = dot_blocks(len);
    k_seg_dot<true><<<dim3(nb, nseg), 256, 0, s>>>(A, E, len, parts);
    return (int)nb;
}

// --------------------------------------------------------- product sumcheck --
__global__ void k_prod_round(const EF *A, const EF *B, size_t h, EF *parts) {
    EF a0 = ezero(), a2 = ezero();
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < h; x += st) {
        EF al = A[x], ah = A[x + h], bl = B[x], bh = B[x + h];
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
int prod_round(const EF *A, const EF *B, int l, EF *parts, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = dot_blocks(h);
    k_prod_round<<<nb, 256, 0, s>>>(A, B, h, parts);
    return (int)nb;
}
__global__ void k_fold_ef(EF *A, size_t h, const EF *c) {
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= h) return;
    EF a = A[x];
    A[x] = eadd(a, emul(*c, esub(A[x + h], a)));
}
void fold_ef(EF *A, int l, const EF *c, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    k_fold_ef<<<nblk(h), 256, 0, s>>>(A, h, c);
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
int a2b_round_ef(const EF *a, const EF *b, int l, const EF *eq, EF *parts, hipStream_t s) {
    size_t h = (size_t)1 << (l - 1);
    unsigned nb = dot_blocks(h);
    k_a2b_round<false><<<dim3(nb, 3), 256, 0, s>>>(a, b, h, eq, parts);
    return (int)nb;
}

}  // namespace wk
