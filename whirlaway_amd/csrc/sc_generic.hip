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
#include <type_traits>
#include "../../include/whirlaway_hip.h"
#include "sc_generic.h"
#include "trace.h"

using namespace wl;

namespace wg {

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ F d_add(F a, F b) { return fadd(a, b); }
__device__ __forceinline__ EF d_add(const EF &a, const EF &b) { return eadd(a, b); }
__device__ __forceinline__ F d_sub(F a, F b) { return fsub(a, b); }
__device__ __forceinline__ EF d_sub(const EF &a, const EF &b) { return esub(a, b); }
__device__ __forceinline__ F d_mul(F a, F b) { return fmul(a, b); }
__device__ __forceinline__ EF d_mul(const EF &a, const EF &b) { return emul(a, b); }
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
// evaluation point blockIdx.y (scalars zsc[blockIdx.y][0..K)); tables: n inputs of
// `len` entries each, contiguous.
template <class T, int MAXR>
__global__ void __launch_bounds__(128) k_eval(const T *tables, size_t len, int k, const F *zsc, const uint32_t *prog,
                                              int n_prog, const EF *batching, const EF *eq, EF *parts) {
    __shared__ EF lds[16];
    const int K = 1 << k;
    const size_t half = len >> k;
    const F *s = zsc + (size_t)blockIdx.y * K;
    T regs[MAXR];
    EF total = ezero();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < half; x += stride) {
        EF acc = ezero();
        int ka = 0;
        for (int pc = 0; pc < n_prog; pc++) {
            const uint32_t w0 = prog[2 * pc], w1 = prog[2 * pc + 1];
            const uint32_t op = w0 & 0xff, dst = w0 >> 8, a = w1 & 0xffff, b = w1 >> 16;
            switch (op) {
            case OP_LOAD: regs[dst] = load_folded(tables + (size_t)a * len, half, x, s, K); break;
            case OP_CONST: regs[dst] = d_const<T>(F{w1}); break;
            case OP_ADD: regs[dst] = d_add(regs[a], regs[b]); break;
            case OP_SUB: regs[dst] = d_sub(regs[a], regs[b]); break;
            case OP_MUL: regs[dst] = d_mul(regs[a], regs[b]); break;
            case OP_NEG: regs[dst] = d_neg(regs[a]); break;
            default: {  // OP_ASSERT
                if constexpr (std::is_same<T, F>::value) acc = eadd(acc, emulf(batching[ka], regs[a]));
                else acc = eadd(acc, emul(batching[ka], regs[a]));
                ka++;
            }
            }
        }
        if (eq) acc = emul(acc, eq[x]);
        total = eadd(total, acc);
    }
    const EF r = block_sum(total, lds);
    if (threadIdx.x == 0) parts[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
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
void compile(const uint32_t *code, size_t n_instr, int n_inputs, Program &out) {
    if (n_inputs < 1 || n_inputs > 65535) throw std::runtime_error("a program reads 1 to 65535 inputs");
    const size_t nv = (size_t)n_inputs + n_instr;
    // last use of every value (inputs and instruction results)
    std::vector<long> last(nv, -1);
    std::vector<char> is_assert(n_instr, 0);
    for (size_t j = 0; j < n_instr; j++) {
        const uint32_t *w = code + 4 * j;
        const uint64_t lim = (uint64_t)n_inputs + j;
        auto use = [&](uint32_t v) {
            if (v >= lim) throw std::runtime_error("program operand refers to a later value");
            if (v >= (uint32_t)n_inputs && is_assert[v - n_inputs])
                throw std::runtime_error("program operand refers to an ASSERT, which defines no value");
            last[v] = (long)j;
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
    // linear scan: a slot per live value, operands freed before the result is placed
    std::vector<int> slot(nv, -1), free_slots;
    int n_slots = 0, n_asserts = 0;
    auto take = [&]() {
        if (!free_slots.empty()) { int s = free_slots.back(); free_slots.pop_back(); return s; }
        return n_slots++;
    };
    std::vector<uint32_t> dev;
    auto emit = [&](uint32_t op, uint32_t dst, uint32_t w1) { dev.push_back(op | (dst << 8)); dev.push_back(w1); };
    for (size_t j = 0; j < n_instr; j++) {
        const uint32_t *w = code + 4 * j;
        const int nops = (w[0] == OP_ADD || w[0] == OP_SUB || w[0] == OP_MUL) ? 2 : (w[0] == OP_CONST ? 0 : 1);
        for (int o = 0; o < nops; o++) {
            const uint32_t v = w[1 + o];
            if (v < (uint32_t)n_inputs && slot[v] < 0) {
                slot[v] = take();
                emit(OP_LOAD, (uint32_t)slot[v], v);
            }
        }
        const uint32_t a = nops >= 1 ? (uint32_t)slot[w[1]] : 0, b = nops >= 2 ? (uint32_t)slot[w[2]] : 0;
        if (w[0] == OP_ASSERT) {
            emit(OP_ASSERT, 0, a);
            n_asserts++;
        }
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
        if (w[0] == OP_CONST) emit(OP_CONST, (uint32_t)slot[vid], to_mont(w[1]).v);
        else emit(w[0], (uint32_t)slot[vid], a | (b << 16));
    }
    if (n_slots > 256) throw std::runtime_error("program keeps more than 256 values live");
    out.n_inputs = n_inputs;
    out.n_asserts = n_asserts;
    out.n_slots = std::max(1, n_slots);
    out.dev = std::move(dev);
}

void upload(wl_ctx *c, Program &p) {
    p.dev_buf = DBuf<uint32_t>(c->pool, std::max<size_t>(2, p.dev.size()));
    if (!p.dev.empty()) c->h2d(p.dev_buf.p, p.dev.data(), p.dev.size());
}

template <class T>
int launch_eval(const Program &p, const T *tab, size_t len, int k, const F *zsc, int nz, const EF *bat, const EF *eq,
                EF *parts, hipStream_t s) {
    const size_t half = len >> k;
    const int gx = (int)std::min<size_t>(nblk(half, 128), 1024);
    const dim3 grid(gx, nz);
    const int n = (int)(p.dev.size() / 2);
    if (p.n_slots <= 16) k_eval<T, 16><<<grid, 128, 0, s>>>(tab, len, k, zsc, p.dev_buf.p, n, bat, eq, parts);
    else if (p.n_slots <= 64) k_eval<T, 64><<<grid, 128, 0, s>>>(tab, len, k, zsc, p.dev_buf.p, n, bat, eq, parts);
    else k_eval<T, 256><<<grid, 128, 0, s>>>(tab, len, k, zsc, p.dev_buf.p, n, bat, eq, parts);
    return gx;
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
            const size_t half = len >> k;
            const int gx = (int)std::min<size_t>(nblk(half, 128), 1024);
            DBuf<EF> parts(c->pool, (size_t)gx * zs.size());
            if (cur_base) launch_eval(prog, (const F *)cur, len, k, dz, (int)zs.size(), dbat, eqm.p, parts, c->stream);
            else launch_eval(prog, (const EF *)cur, len, k, dz, (int)zs.size(), dbat, eqm.p, parts, c->stream);
            HIP_CHECK(hipGetLastError());
            vals = reduce_many(c, parts, (int)zs.size(), gx);
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
