// micro.hip -- C-ABI entry points for the kernel-level operations: folds
// (crates/utils/src/multilinear.rs:90-123, benches/folding_benchmarks.rs), the
// eq-weighted product sumcheck (crates/sumcheck/src/prove.rs:22-254 with
// skips = 1), Poseidon2 permutations and the Reed-Solomon/Merkle commitment core.
#include <hip/hip_runtime.h>
#include "../../include/whirlaway_hip.h"
#include "ctx.h"
#include "tune.h"

using namespace wl;

namespace {

void upload_canon(wl_ctx *c, uint32_t *dst, const uint32_t *src, size_t n) {
    HIP_CHECK(hipMemcpyAsync(dst, src, n * 4, hipMemcpyHostToDevice, c->stream));
    wk::convert_to_mont(dst, n, c->stream);
}
void download_canon(wl_ctx *c, uint32_t *dst, uint32_t *src_dev, size_t n) {
    wk::convert_from_mont(src_dev, n, c->stream);
    c->d2h(dst, src_dev, n);
}
std::vector<EF> host_ef(const uint32_t *canon, size_t n) {
    std::vector<EF> v(n);
    for (size_t i = 0; i < n; i++) v[i] = wh::to_mont_ef(canon + 8 * i);
    return v;
}

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

// Throughput probe of the Montgomery product: four independent chains per
// thread, 2^20 threads (16 waves per SIMD on 256 CUs).
__global__ void __launch_bounds__(256) k_fmul_probe(uint32_t *out, int iters) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    F x{g % P}, y{(g * 7u + 3u) % P}, z{(g ^ 0x55u) % P}, w{(g + 99u) % P};
    for (int i = 0; i < iters; i++) {
        x = fmul(x, y);
        y = fmul(y, z);
        z = fmul(z, w);
        w = fmul(w, x);
    }
    out[g] = x.v ^ y.v ^ z.v ^ w.v;
}

// Host-side check of the Poseidon2 partial rounds by convolution against the
// round-by-round form, for one width: random round constants (xorshift from
// `seed`), `n` random states plus field-extreme states.  Returns the number of
// states whose permutations differ.
template <int W, int RP, typename C>
uint64_t conv_mismatches(const PartialConv<W, RP> &pc, const uint32_t *diag, uint32_t n, uint64_t &rng) {
    auto next = [&]() {
        rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
        return (uint32_t)(rng % P);
    };
    C k;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < W; i++) k.ext_init[r][i] = to_mont(next()), k.ext_term[r][i] = to_mont(next());
    for (int r = 0; r < RP; r++) k.internal[r] = to_mont(next());
    const uint32_t extremes[] = {0u, 1u, P - 1, P - 2, P / 2, (P + 1) / 2, 1u << 24, (1u << 24) - 1, 1u << 30};
    const uint32_t n_ext = sizeof(extremes) / sizeof(extremes[0]);
    uint64_t bad = 0;
    for (uint32_t st = 0; st < n + n_ext * 3; st++) {
        F a[W], b[W];
        for (int i = 0; i < W; i++) {
            uint32_t v;
            if (st < n) v = next();
            else if (st < n + n_ext) v = extremes[st - n];  // constant state
            else if (st < n + 2 * n_ext) v = i & 1 ? extremes[st - n - n_ext] : 0u;
            else v = (i == (int)(st % W)) ? extremes[st - n - 2 * n_ext] : P - 1;
            a[i] = b[i] = to_mont(v);
        }
        F h[W];
        for (int i = 0; i < W; i++) h[i] = a[i];
        p2_permute<W, RP>(a, k, diag);
        p2_permute_conv<W, RP>(b, k, pc);
        bool differ = false;
        for (int i = 0; i < W; i++) differ |= a[i].v != b[i].v;
#if !defined(__HIP_DEVICE_COMPILE__)
        if constexpr (W == 16) {  // the transcript's AVX2 permutation on the same state
            wh::HostP16::permute(h, k);
            for (int i = 0; i < W; i++) differ |= a[i].v != h[i].v;
        }
#endif
        bad += differ;
    }
#if !defined(__HIP_DEVICE_COMPILE__)
    if constexpr (W == 16) {  // its partial-round cube on field extremes and random inputs
        for (uint32_t j = 0; j < n + 64; j++) {
            const uint32_t t = j < 32 ? j : (j < 64 ? P - 1 - (j - 32) : next());
            bad += wh::HostP16::cube64(t) != fcube(F{t}).v;
        }
    }
#endif
    return bad;
}

}  // namespace

extern "C" {

int wl_p2_conv_check(uint32_t n_states, uint64_t seed, uint64_t *mismatches) {
    if (!mismatches) { wl_set_error("null argument"); return WL_ERR_ARG; }
    uint64_t rng = seed * 0x9e3779b97f4a7c15ull + 1;
    mismatches[0] = conv_mismatches<16, 20, P2Consts16>(PCONV16, DIAG.d16, n_states, rng);
    mismatches[1] = conv_mismatches<24, 23, P2Consts24>(PCONV24, DIAG.d24, n_states, rng);
    return WL_OK;
}

int wl_fmul_rate(wl_ctx *c, double *fmul_per_s) {
    if (!fmul_per_s) { wl_set_error("null argument"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        const int threads = 1 << 20, iters = 1024;
        DBuf<uint32_t> out(c->pool, threads);
        hipLaunchKernelGGL(k_fmul_probe, dim3(threads / 256), dim3(256), 0, c->stream, out.p, iters);  // warm-up
        HIP_CHECK(hipEventRecord(c->ev0, c->stream));
        hipLaunchKernelGGL(k_fmul_probe, dim3(threads / 256), dim3(256), 0, c->stream, out.p, iters);
        HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        HIP_CHECK(hipEventSynchronize(c->ev1));
        HIP_CHECK(hipGetLastError());
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *fmul_per_s = (double)threads * iters * 4 / (ms * 1e-3);
        return WL_OK;
    });
}

int wl_fold_large(wl_ctx *c, const uint32_t *in, int in_is_ef, uint32_t log_n, const uint32_t *scalars,
                  uint32_t log_k, uint32_t *out) {
    if (!in || !scalars || !out || log_k > log_n) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        size_t N = (size_t)1 << log_n, M = N >> log_k, K = (size_t)1 << log_k;
        size_t w = in_is_ef ? 8 : 1;
        DBuf<uint32_t> din(c->pool, N * w);
        upload_canon(c, din, in, N * w);
        DBuf<EF> dsc = upload_ef(c, host_ef(scalars, K));
        DBuf<EF> dout(c->pool, M);
        if (in_is_ef) wk::fold_ef_large((const EF *)din.p, (int)log_n, dsc, (int)log_k, dout, c->stream);
        else wk::fold_f_large((const F *)din.p, (int)log_n, dsc, (int)log_k, dout, c->stream);
        download_canon(c, out, (uint32_t *)dout.p, M * 8);
        return WL_OK;
    });
}

int wl_fold_chain(wl_ctx *c, const uint32_t *in, uint32_t log_n, const uint32_t *challenges,
                  uint32_t n_folds, uint32_t *out_all, uint32_t *final_out, double *ms) {
    if (!in || !challenges || n_folds < 1 || n_folds > log_n) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        size_t N = (size_t)1 << log_n;
        DBuf<F> din(c->pool, N);
        upload_canon(c, (uint32_t *)din.p, in, N);
        std::vector<EF> ch = host_ef(challenges, n_folds);
        DBuf<EF> dch = upload_ef(c, ch);
        DBuf<EF> cur(c->pool, N / 2);
        HIP_CHECK(hipEventRecord(c->ev0, c->stream));
        wk::fold_f_to_ef(din, (int)log_n, dch.p, cur, c->stream);
        for (uint32_t i = 1; i < n_folds; i++) {
            if (log_n - i <= 12) {  // the rest in one launch
                wk::fold_ef_tail(cur, (int)(log_n - i), (int)(n_folds - i), dch.p + i, c->stream);
                break;
            }
            wk::fold_ef(cur, (int)(log_n - i), dch.p + i, c->stream);
        }
        HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        HIP_CHECK(hipEventSynchronize(c->ev1));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, c->ev0, c->ev1));
        if (ms) *ms = t;
        if (out_all) {
            // recompute stage by stage so every intermediate table is returned
            DBuf<EF> st(c->pool, N / 2);
            wk::fold_f_to_ef(din, (int)log_n, dch.p, st, c->stream);
            size_t off = 0;
            for (uint32_t i = 0; i < n_folds; i++) {
                if (i > 0) wk::fold_ef(st, (int)(log_n - i), dch.p + i, c->stream);
                size_t len = (size_t)1 << (log_n - 1 - i);
                DBuf<EF> cp(c->pool, len);
                HIP_CHECK(hipMemcpyAsync(cp.p, st.p, len * sizeof(EF), hipMemcpyDeviceToDevice, c->stream));
                download_canon(c, out_all + off * 8, (uint32_t *)cp.p, len * 8);
                off += len;
            }
        }
        if (final_out) download_canon(c, final_out, (uint32_t *)cur.p, ((size_t)1 << (log_n - n_folds)) * 8);
        return WL_OK;
    });
}

int wl_sumcheck_a2b(wl_ctx *c, const uint32_t *p16, const uint32_t *a, const uint32_t *b, uint32_t n,
                    const uint32_t *z, uint32_t *challenges, uint32_t *final_evals, uint32_t *final_sum,
                    uint32_t *proof_out, size_t proof_cap, size_t *proof_len, double *ms) {
    if (!p16 || !a || !b || !z || n < 1) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    return guard(c, [&]() {
        size_t N = (size_t)1 << n;
        P2Dev ph;
        load_p2(p16, nullptr, ph);
        DBuf<P2Dev> pd = upload_p2(c, ph);
        wh::Transcript ts(&ph.p16);
        std::vector<EF> zz = host_ef(z, n);
        DBuf<F> da(c->pool, N), db(c->pool, N);
        upload_canon(c, (uint32_t *)da.p, a, N);
        upload_canon(c, (uint32_t *)db.p, b, N);
        HIP_CHECK(hipEventRecord(c->ev0, c->stream));
        // Round j weighs x by eq(z[j+1..], x).  While that has more than L + 1
        // variables it is split as ehi_j[x >> L] elo[x mod 2^L]: elo = eq over the
        // last L variables, ehi_j over z[j+1 .. n-1-L] (all ehi_j built at once, 2^(n-L)
        // entries in all), so no table of the round's length is built or read.  From
        // the round with L + 1 variables on, one eq table of 2^(L+1) entries that
        // every later round halves inside its kernel.  L >= n - 11 keeps a block's
        // range (at most 2^(n-1) / 1024 values) inside one ehi entry.
        const int L = std::max(13, (int)n - 11);
        const bool split = (int)n - 1 > L;
        std::vector<DBuf<EF>> ehi;
        DBuf<EF> elo, eq;
        if (split) {
            elo = DBuf<EF>(c->pool, (size_t)1 << L);
            eq = DBuf<EF>(c->pool, (size_t)1 << (L + 1));
            std::vector<EqJob> jobs;
            jobs.push_back(EqJob{std::vector<EF>(zz.end() - L, zz.end()), eone(), elo.p});
            jobs.push_back(EqJob{std::vector<EF>(zz.end() - (L + 1), zz.end()), eone(), eq.p});
            for (int j = 0; j + 1 + L < (int)n; j++) {  // ehi_j over z[j+1 .. n-1-L]
                ehi.emplace_back(c->pool, (size_t)1 << (n - 1 - L - j));
                jobs.push_back(EqJob{std::vector<EF>(zz.begin() + j + 1, zz.end() - L), eone(), ehi.back().p});
            }
            eq_batch(c, jobs);
        } else {
            // eq(z[1..], .) once; every later round halves it inside its kernel
            eq = eq_dev(c, std::vector<EF>(zz.begin() + 1, zz.end()), eone());
        }
        // round 0 on the base-field tables at z = 0..3: z = 0, 1 give the claimed sum
        // (1 - z0) S(0) + z0 S(1), z = 0, 2, 3 the round's evaluations
        EF sum;
        std::vector<EF> v3(3);
        {
            int nb = wk::dot_blocks(N / 2);
            DBuf<EF> parts(c->pool, 4 * (size_t)nb);
            if (split) wk::a2b_round_f_split(da, db, (int)n, L, elo, ehi[0], parts, c->stream, true);
            else wk::a2b_round_f(da, db, (int)n, eq, parts, c->stream, true);
            std::vector<EF> s4 = reduce_many(c, parts, 4, nb);
            sum = eadd(emul(esub(eone(), zz[0]), s4[0]), emul(zz[0], s4[1]));
            v3 = {s4[0], s4[2], s4[3]};
        }
        EF missing = eone();
        bool has_missing = false;
        std::vector<EF> ch;
        DBuf<EF> ea(c->pool, std::max<size_t>(N / 2, 1)), eb(c->pool, std::max<size_t>(N / 2, 1));
        for (uint32_t j = 0; j < n; j++) {
            int l = (int)(n - j);
            EF r = zz[j];
            if (j > 0) {
                // fold by the previous challenge, halve eq, evaluate: one pass
                int nb = wk::dot_blocks((size_t)1 << (l - 1));
                DBuf<EF> parts(c->pool, 3 * (size_t)nb);
                if (split && l - 1 > L)
                    wk::a2b_round_fused_split(da, db, ea, eb, l, L, j == 1, ch.back(), elo, ehi[j], parts,
                                              c->stream);
                else
                    wk::a2b_round_fused(da, db, ea, eb, l, j == 1, ch.back(), eq, parts, c->stream);
                v3 = reduce_many(c, parts, 3, nb);
            }
            EF e0 = v3[0], e2 = v3[1], e3 = v3[2];
            if (has_missing) { e0 = emul(e0, missing); e2 = emul(e2, missing); e3 = emul(e3, missing); }
            EF e1 = emul(esub(sum, emul(e0, esub(eone(), r))), wh::einv(r));
            std::vector<F> xs = {fsmall(0), fsmall(1), fsmall(2), fsmall(3)};
            std::vector<EF> p = wh::lagrange(xs, {e0, e1, e2, e3});
            p = wh::pmul(p, {esub(eone(), r), esub(edbl(r), eone())});
            ts.add_ef(p.data(), p.size());
            EF cj = ts.sample_ef();
            ch.push_back(cj);
            sum = wh::peval(p, cj);
            grind(c, ts, pd.p, pow_bits_sumcheck(128, 3 + 1));
            EF f = eadd(emul(esub(eone(), r), esub(eone(), cj)), emul(r, cj));
            missing = has_missing ? emul(f, missing) : f;
            has_missing = true;
        }
        // the last fold (tables of length 2) on the host
        EF fa, fb;
        {
            EF lo[2], hi[2];
            if (n == 1) {
                F t[4];
                c->d2h(t, da.p, 2);
                c->d2h(t + 2, db.p, 2);
                lo[0] = efrom(t[0]); hi[0] = efrom(t[1]); lo[1] = efrom(t[2]); hi[1] = efrom(t[3]);
            } else {
                EF t[4];
                c->d2h(t, ea.p, 2);
                c->d2h(t + 2, eb.p, 2);
                lo[0] = t[0]; hi[0] = t[1]; lo[1] = t[2]; hi[1] = t[3];
            }
            fa = eadd(lo[0], emul(ch.back(), esub(hi[0], lo[0])));
            fb = eadd(lo[1], emul(ch.back(), esub(hi[1], lo[1])));
        }
        HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        HIP_CHECK(hipEventSynchronize(c->ev1));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, c->ev0, c->ev1));
        if (ms) *ms = t;
        for (uint32_t j = 0; j < n; j++) wh::from_mont_ef(ch[j], challenges + 8 * j);
        wh::from_mont_ef(fa, final_evals);
        wh::from_mont_ef(fb, final_evals + 8);
        wh::from_mont_ef(sum, final_sum);
        if (proof_len) *proof_len = ts.data.size();
        if (proof_out) {
            if (proof_cap < ts.data.size()) { wl_set_error("proof buffer too small"); return WL_ERR_BUFFER; }
            memcpy(proof_out, ts.data.data(), ts.data.size() * 4);
        }
        return WL_OK;
    });
}

int wl_poseidon2_permute(wl_ctx *c, const uint32_t *p16, const uint32_t *p24, int width, uint32_t *states,
                         size_t n) {
    if ((width != 16 && width != 24) || !states || (width == 16 ? !p16 : !p24)) {
        wl_set_error("bad arguments");
        return WL_ERR_ARG;
    }
    return guard(c, [&]() {
        P2Dev ph;
        load_p2(p16, p24, ph);
        DBuf<P2Dev> pd = upload_p2(c, ph);
        DBuf<uint32_t> ds(c->pool, n * width);
        upload_canon(c, ds, states, n * width);
        wk::p2_permute_batch(pd.p, width, (F *)ds.p, (int)n, c->stream);
        download_canon(c, states, ds, n * width);
        return WL_OK;
    });
}

int wl_rs_commit(wl_ctx *c, const uint32_t *p16, const uint32_t *p24, const uint32_t *cols, uint32_t n_cols,
                 uint32_t log_m, uint32_t log_rows, uint32_t *codeword_out, uint32_t *root_out) {
    if (!p16 || !p24 || !cols || !root_out || log_rows < log_m || n_cols < 1) {
        wl_set_error("bad arguments");
        return WL_ERR_ARG;
    }
    return guard(c, [&]() {
        P2Dev ph;
        load_p2(p16, p24, ph);
        DBuf<P2Dev> pd = upload_p2(c, ph);
        size_t m = (size_t)1 << log_m, rows = (size_t)1 << log_rows;
        DBuf<uint32_t> dcols(c->pool, n_cols * m);
        upload_canon(c, dcols, cols, n_cols * m);
        DBuf<F> cw(c->pool, n_cols * rows), tmp(c->pool, n_cols * rows), nodes(c->pool, 2 * rows * 8);
        int la = wt::rs_encode(c, (const F *)dcols.p, (int)n_cols, (int)n_cols, (int)log_m, (int)log_rows, cw, tmp);
        HIP_CHECK(hipGetLastError());
        wt::merkle_leaves(c, cw, (int)n_cols, (int)n_cols, (int)log_rows, la, pd.p, nodes);
        wt::merkle_nodes(c, (int)log_rows, pd.p, nodes);
        download_canon(c, root_out, (uint32_t *)nodes.p + 8, 8);
        if (codeword_out) {
            // storage order -> natural row order
            std::vector<uint32_t> st(n_cols * rows);
            download_canon(c, st.data(), (uint32_t *)cw.p, n_cols * rows);
            const int lb = (int)log_rows - la;
            for (size_t col = 0; col < n_cols; col++)
                for (size_t r = 0; r < rows; r++) {
                    size_t at = ((r & (((size_t)1 << la) - 1)) << lb) | (r >> la);
                    codeword_out[col * rows + r] = st[col * rows + at];
                }
        }
        return WL_OK;
    });
}

}  // extern "C"
