// utils.hip -- crates/utils counterparts on the device and their C-ABI entry
// points: fold_multilinear_in_small_field (crates/utils/src/multilinear.rs:10-52),
// multilinears_linear_combination (:125-147) and multilinear evaluation at an
// extension point (EvaluationsList::evaluate, called at crates/air/src/prove.rs:190
// and verify.rs:114-136).
//
// Evaluation splits eq(point, .) into the tensor product of two tables of at
// most 2^14 entries, eq_hi over the top h variables and eq_lo over the bottom l:
//   f(point) = sum_hi eq_hi[hi] * sum_lo eq_lo[lo] * f(hi 2^l + lo).
// A block walks whole rows hi, so each lane multiplies its row sum by eq_hi once;
// the input is read once, in canonical form, and converted on load.
#include <hip/hip_runtime.h>
#include <mutex>
#include "../../include/whirlaway_hip.h"
#include "ctx.h"
#include "trace.h"

using namespace wl;

namespace wu {

constexpr int MAX_HALF = 14;     // eq_hi and eq_lo hold at most 2^14 entries each
constexpr int MAX_VARS = 2 * MAX_HALF;
constexpr int DOT_BLOCKS = 1024; // at most this many block partials

static inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// split of n variables: l bottom (row) variables, h = n - l top ones
inline void split_vars(int n, int *h, int *l) {
    *l = n <= 8 ? n : std::max(8, (n + 1) / 2);
    *h = n - *l;
}

struct Point { EF p[MAX_VARS]; int n; int h; };

// eq_hi[i] for i < 2^h (coordinates p[0..h)) and eq_lo[j] for j < 2^l (p[h..n)),
// each entry the product over its bits (big-endian: coordinate 0 is the top bit)
__global__ void k_eq_split(Point pt, EF *ehi, EF *elo) {
    const int h = pt.h, l = pt.n - pt.h;
    const size_t nh = (size_t)1 << h, nl = (size_t)1 << l;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh + nl) return;
    const bool hi = i < nh;
    const size_t x = hi ? i : i - nh;
    const int nv = hi ? h : l, base = hi ? 0 : h;
    EF acc = eone();
    for (int k = 0; k < nv; k++) {
        const EF &p = pt.p[base + k];
        const bool bit = (x >> (nv - 1 - k)) & 1;
        acc = emul(acc, bit ? p : esub(eone(), p));
    }
    (hi ? ehi : elo)[x] = acc;
}

// Block partials of sum_x eq_hi[x >> l] eq_lo[x mod 2^l] f(x), f canonical (F when
// !EFIN, EF when EFIN); block b takes rows hi = b, b + gridDim.x, ...
template <bool EFIN>
__global__ void __launch_bounds__(256) k_mle_dot(const uint32_t *f, int h, int l, const EF *ehi, const EF *elo,
                                                 EF *parts) {
    __shared__ EF lds[16];
    const size_t nh = (size_t)1 << h, nl = (size_t)1 << l;
    EF total = ezero();
    for (size_t row = blockIdx.x; row < nh; row += gridDim.x) {
        EF rs;
        if (!EFIN) {
            AccF acc;
            for (size_t lo = threadIdx.x; lo < nl; lo += blockDim.x)
                acc.add(elo[lo], to_mont(f[(row << l) + lo]));
            rs = acc.get();
        } else {
            rs = ezero();
            for (size_t lo = threadIdx.x; lo < nl; lo += blockDim.x) {
                const uint32_t *e = f + ((row << l) + lo) * 8;
                EF v;
#pragma unroll
                for (int c = 0; c < 8; c++) v.c[c] = to_mont(e[c]);
                rs = eadd(rs, emul(elo[lo], v));
            }
        }
        total = eadd(total, emul(ehi[row], rs));
    }
    const EF s = block_sum(total, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// out = sum of n parts, written as 8 canonical words
__global__ void k_sum_canon(const EF *parts, int n, uint32_t *out) {
    __shared__ EF lds[16];
    EF acc = ezero();
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc = eadd(acc, parts[i]);
    const EF s = block_sum(acc, lds);
    if (threadIdx.x == 0)
        for (int c = 0; c < 8; c++) out[c] = from_mont(s.c[c]);
}

// Enqueue the evaluation: scratch holds eq_hi, eq_lo and the block partials
// (scratch_bytes(n) bytes); out receives 8 canonical words on the device.
inline size_t scratch_bytes(int n) {
    int h, l;
    split_vars(n, &h, &l);
    return (((size_t)1 << h) + ((size_t)1 << l) + DOT_BLOCKS) * sizeof(EF);
}
inline void mle_eval_enqueue(const uint32_t *f, bool is_ef, int n, const EF *point, void *scratch, uint32_t *out,
                             hipStream_t s) {
    Point pt;
    memset(&pt, 0, sizeof pt);
    int h, l;
    split_vars(n, &h, &l);
    pt.n = n;
    pt.h = h;
    for (int i = 0; i < n; i++) pt.p[i] = point[i];
    EF *ehi = (EF *)scratch, *elo = ehi + ((size_t)1 << h), *parts = elo + ((size_t)1 << l);
    const size_t ntab = ((size_t)1 << h) + ((size_t)1 << l);
    k_eq_split<<<nblk(ntab), 256, 0, s>>>(pt, ehi, elo);
    const int nb = (int)std::min<size_t>((size_t)1 << h, DOT_BLOCKS);
    if (is_ef) k_mle_dot<true><<<nb, 256, 0, s>>>(f, h, l, ehi, elo, parts);
    else k_mle_dot<false><<<nb, 256, 0, s>>>(f, h, l, ehi, elo, parts);
    k_sum_canon<<<1, 256, 0, s>>>(parts, nb, out);
}

// out[i] = sum_j sc[j] in[i + j M], M = 2^(l - k); F scalars, F or EF entries
// (Montgomery form in and out)
template <bool EFIN>
__global__ void k_fold_small(const F *in, size_t M, const F *sc, int K, F *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    if (!EFIN) {
        uint64_t acc = 0;
        for (int j = 0; j < K; j++) {
            acc += (uint64_t)sc[j].v * in[i + (size_t)j * M].v;
            if ((j & 3) == 3) acc = fold64(acc);
        }
        out[i] = F{redc64(acc)};
    } else {
        uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < K; j++) {
            const F *e = in + (i + (size_t)j * M) * 8;
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] += (uint64_t)sc[j].v * e[c].v;
            if ((j & 3) == 3)
#pragma unroll
                for (int c = 0; c < 8; c++) acc[c] = fold64(acc[c]);
        }
#pragma unroll
        for (int c = 0; c < 8; c++) out[i * 8 + c] = F{redc64(acc[c])};
    }
}

// Scratch of wl_mle_evaluate_noalloc: eq_hi, eq_lo and the block partials of the
// largest evaluation (28 variables), one copy per device for the whole process.
// Host code uses the device address that hipGetSymbolAddress returns for the
// current device, never the name.
constexpr size_t NOALLOC_SCRATCH_BYTES = (((size_t)1 << MAX_HALF) * 2 + DOT_BLOCKS) * sizeof(EF);
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES];  // per device, set on the first call there

// device address of this device's scratch (the device must be current)
void *device_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("no device address for the evaluation scratch");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace wu

namespace {

template <class Fn>
int guard(wl_ctx *c, Fn fn) {
    if (!c) { wl_set_error("null context"); return WL_ERR_ARG; }
    try {
        HIP_CHECK(hipSetDevice(c->device));
        return fn();
    } catch (const std::exception &e) {
        wl_set_error(e.what());
        return WL_ERR_DEVICE;
    }
}

}  // namespace

extern "C" {

int wl_fold_small(wl_ctx *c, const uint32_t *in, int in_is_ef, uint32_t log_n, const uint32_t *scalars,
                  uint32_t log_k, uint32_t *out) {
    if (!in || !scalars || !out || log_k > log_n || log_n > 32) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        const size_t N = (size_t)1 << log_n, M = N >> log_k, K = (size_t)1 << log_k;
        const size_t w = in_is_ef ? 8 : 1;
        DBuf<uint32_t> din(c->pool, N * w), dsc(c->pool, K), dout(c->pool, M * w);
        HIP_CHECK(hipMemcpyAsync(din.p, in, N * w * 4, hipMemcpyHostToDevice, c->stream));
        wk::convert_to_mont(din.p, N * w, c->stream);
        HIP_CHECK(hipMemcpyAsync(dsc.p, scalars, K * 4, hipMemcpyHostToDevice, c->stream));
        wk::convert_to_mont(dsc.p, K, c->stream);
        if (in_is_ef)
            wu::k_fold_small<true><<<wu::nblk(M), 256, 0, c->stream>>>((const F *)din.p, M, (const F *)dsc.p, (int)K,
                                                                       (F *)dout.p);
        else
            wu::k_fold_small<false><<<wu::nblk(M), 256, 0, c->stream>>>((const F *)din.p, M, (const F *)dsc.p,
                                                                        (int)K, (F *)dout.p);
        HIP_CHECK(hipGetLastError());
        wk::convert_from_mont(dout.p, M * w, c->stream);
        HIP_CHECK(hipMemcpyAsync(out, dout.p, M * w * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return WL_OK;
    });
}

int wl_multilinears_linear_combination(wl_ctx *c, const uint32_t *cols, uint32_t n_cols, uint32_t log_n,
                                       const uint32_t *scalars, uint32_t *out) {
    if (!cols || !scalars || !out || n_cols < 1 || log_n > 30) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        const size_t N = (size_t)1 << log_n;
        DBuf<uint32_t> din(c->pool, N * n_cols);
        HIP_CHECK(hipMemcpyAsync(din.p, cols, N * n_cols * 4, hipMemcpyHostToDevice, c->stream));
        wk::convert_to_mont(din.p, N * n_cols, c->stream);
        std::vector<EF> sc(n_cols);
        for (uint32_t i = 0; i < n_cols; i++) sc[i] = wh::to_mont_ef(scalars + 8 * (size_t)i);
        DBuf<EF> dsc = upload_ef(c, sc);
        DBuf<EF> dout(c->pool, N);
        wl::Span span("multilinears_linear_combination");  // crates/utils/src/multilinear.rs:125
        wk::lincomb_cols((const F *)din.p, (int)log_n, (int)n_cols, dsc, dout, c->stream);
        HIP_CHECK(hipGetLastError());
        wk::convert_from_mont((uint32_t *)dout.p, N * 8, c->stream);
        HIP_CHECK(hipMemcpyAsync(out, dout.p, N * 32, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return WL_OK;
    });
}

int wl_mle_evaluate(wl_ctx *c, const uint32_t *evals, int is_ef, uint32_t n_vars, const uint32_t *point,
                    uint32_t *out) {
    if (!evals || !out || (n_vars && !point) || n_vars > (uint32_t)wu::MAX_VARS) {
        wl_set_error("bad arguments (at most 28 variables)");
        return WL_ERR_ARG;
    }
    return guard(c, [&]() {
        const size_t N = (size_t)1 << n_vars, w = is_ef ? 8 : 1;
        std::vector<EF> pt(n_vars);
        for (uint32_t i = 0; i < n_vars; i++) pt[i] = wh::to_mont_ef(point + 8 * (size_t)i);
        DBuf<uint32_t> din(c->pool, N * w), dout(c->pool, 8);
        DBuf<uint8_t> scratch(c->pool, wu::scratch_bytes((int)n_vars));
        HIP_CHECK(hipMemcpyAsync(din.p, evals, N * w * 4, hipMemcpyHostToDevice, c->stream));
        wu::mle_eval_enqueue(din.p, is_ef != 0, (int)n_vars, pt.data(), scratch.p, dout.p, c->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(out, dout.p, 32, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return WL_OK;
    });
}

int wl_mle_evaluate_noalloc(wl_ctx *c, const uint32_t *evals_dev, int is_ef, uint32_t n_vars, const uint32_t *point,
                            uint32_t *out_dev) {
    if (!evals_dev || !out_dev || (n_vars && !point)) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    if (n_vars > (uint32_t)wu::MAX_VARS || wu::scratch_bytes((int)n_vars) > wu::NOALLOC_SCRATCH_BYTES) {
        wl_set_error("wl_mle_evaluate_noalloc: at most 28 variables (its scratch holds 2 x 2^14 + 1024 EF)");
        return WL_ERR_ARG;
    }
    return guard(c, [&]() {
        EF pt[wu::MAX_VARS];  // the point travels in the kernel arguments
        for (uint32_t i = 0; i < n_vars; i++) pt[i] = wh::to_mont_ef(point + 8 * (size_t)i);
        void *scr = wu::device_scratch(c->device);
        // every scratch entry that is read is written first in the same call, so
        // the previous call's contents need no clearing
        wu::mle_eval_enqueue(evals_dev, is_ef != 0, (int)n_vars, pt, scr, out_dev, c->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return WL_OK;
    });
}

}  // extern "C"
