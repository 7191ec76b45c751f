// host.h -- host-side protocol logic of the MI355X prover: extension-field
// inverse, dense univariate polynomials, the duplex-sponge transcript and the
// WHIR round schedule.  Everything here is O(polylog) in the trace size; every
// pass over trace-sized data is a HIP kernel.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <stdexcept>
#include <vector>
#include "common.h"
#include "host_p16.h"

namespace wh {
using namespace kb;

// ------------------------------------------------------------------ field --
inline EF efrob(const EF &a, int times) {
    // X^p = X * W^((p-1)/8): coefficient k scales by (W^((p-1)/8))^(k*times)
    F z = fpow(fsmall(W_EXT), (uint64_t)(P - 1) / 8);
    F zt = fpow(z, (uint64_t)times), m = fone();
    EF r;
    for (int k = 0; k < 8; k++) { r.c[k] = fmul(a.c[k], m); m = fmul(m, zt); }
    return r;
}
inline EF einv(const EF &a) {
    EF acc = eone();
    for (int i = 1; i < 8; i++) acc = emul(acc, efrob(a, i));
    EF nrm = emul(acc, a);
    return emulf(acc, finv(nrm.c[0]));
}
inline EF esmall(uint32_t c) { return efrom(fsmall(c)); }
inline EF to_mont_ef(const uint32_t *canon) { EF r; for (int i = 0; i < 8; i++) r.c[i] = to_mont(canon[i]); return r; }
inline void from_mont_ef(const EF &a, uint32_t *out) { for (int i = 0; i < 8; i++) out[i] = from_mont(a.c[i]); }

inline size_t log2_up(size_t x) {  // crates/utils/src/misc.rs:1-8
    if (x <= 1) return 0;
    size_t r = 0, v = x - 1;
    while (v) { r++; v >>= 1; }
    return r;
}

// ------------------------------------------------------------ polynomials --
// Lagrange interpolation over distinct F nodes; coefficients low degree first.
inline std::vector<EF> lagrange(const std::vector<F> &xs, const std::vector<EF> &ys) {
    int n = (int)xs.size();
    std::vector<F> m(n + 1, fzero()), q(n);
    m[0] = fone();
    for (int i = 0; i < n; i++) {
        for (int d = i + 1; d > 0; d--) m[d] = fsub(m[d - 1], fmul(xs[i], m[d]));
        m[0] = fneg(fmul(xs[i], m[0]));
    }
    std::vector<EF> co(n, ezero());
    for (int i = 0; i < n; i++) {
        F carry = m[n];
        for (int d = n - 1; d >= 0; d--) { q[d] = carry; carry = fadd(m[d], fmul(carry, xs[i])); }
        F den = fone();
        for (int j = 0; j < n; j++) if (j != i) den = fmul(den, fsub(xs[i], xs[j]));
        EF s = emulf(ys[i], finv(den));
        for (int d = 0; d < n; d++) co[d] = eadd(co[d], emulf(s, q[d]));
    }
    return co;
}
inline EF peval(const std::vector<EF> &c, const EF &x) {
    EF acc = ezero();
    for (int i = (int)c.size() - 1; i >= 0; i--) acc = eadd(emul(acc, x), c[i]);
    return acc;
}
inline std::vector<EF> pmul(const std::vector<EF> &a, const std::vector<EF> &b) {
    std::vector<EF> o(a.size() + b.size() - 1, ezero());
    for (size_t i = 0; i < a.size(); i++)
        for (size_t j = 0; j < b.size(); j++) o[i + j] = eadd(o[i + j], emul(a[i], b[j]));
    return o;
}
// univariate_selectors(k) (crates/utils/src/univariate.rs:5-15) evaluated at x
inline std::vector<F> selectors_f(int k, F x) {
    int N = 1 << k;
    std::vector<F> out(N);
    for (int i = 0; i < N; i++) {
        F num = fone(), den = fone();
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            num = fmul(num, fsub(x, fsmall(j)));
            den = fmul(den, fsub(fsmall(i), fsmall(j)));
        }
        out[i] = fmul(num, finv(den));
    }
    return out;
}
inline std::vector<EF> selectors_ef(int k, const EF &x) {
    int N = 1 << k;
    std::vector<EF> out(N);
    for (int i = 0; i < N; i++) {
        EF num = eone();
        F den = fone();
        for (int j = 0; j < N; j++) {
            if (j == i) continue;
            num = emul(num, esub(x, esmall(j)));
            den = fmul(den, fsub(fsmall(i), fsmall(j)));
        }
        out[i] = emulf(num, finv(den));
    }
    return out;
}
// eq(point, .) big-endian (host, small tables only)
inline std::vector<EF> eq_host(const std::vector<EF> &p) {
    std::vector<EF> t(1, eone());
    for (size_t k = 0; k < p.size(); k++) {
        std::vector<EF> n(t.size() * 2);
        for (size_t j = 0; j < t.size(); j++) {
            EF hi = emul(t[j], p[k]);
            n[2 * j] = esub(t[j], hi);
            n[2 * j + 1] = hi;
        }
        t.swap(n);
    }
    return t;
}
inline EF eq_outside(const EF *a, const EF *b, int n) {
    EF acc = eone();
    for (int i = 0; i < n; i++)
        acc = emul(acc, eadd(emul(a[i], b[i]), emul(esub(eone(), a[i]), esub(eone(), b[i]))));
    return acc;
}
inline std::vector<EF> expand_univ(const EF &x, int n) {
    std::vector<EF> out(n);
    EF cur = x;
    for (int i = 0; i < n; i++) { out[n - 1 - i] = cur; cur = esqr(cur); }
    return out;
}
// h(r) for the degree-2 polynomial through (0,h0),(1,h1),(2,h2)
inline EF interp3(const EF h[3], const EF &r) {
    EF one = eone(), two = esmall(2);
    EF rm1 = esub(r, one), rm2 = esub(r, two);
    F inv2 = finv(fsmall(2));
    EF t0 = emulf(emul(emul(h[0], rm1), rm2), inv2);
    EF t1 = eneg(emul(emul(h[1], r), rm2));
    EF t2 = emulf(emul(emul(h[2], r), rm1), inv2);
    return eadd(eadd(t0, t1), t2);
}

// -------------------------------------------------------------- transcript --
// DuplexChallenger<F, Poseidon16, 16, 8> + proof_data (canonical u32).
struct Transcript {
    const P2Consts16 *p16;
    F state[16];
    F in_buf[8];
    int n_in = 0;
    F out_buf[8];
    int n_out = 0;
    std::vector<uint32_t> data;

    explicit Transcript(const P2Consts16 *p) : p16(p) {
        for (int i = 0; i < 16; i++) state[i] = fzero();
    }
    void duplex() {
        for (int i = 0; i < n_in; i++) state[i] = in_buf[i];
        n_in = 0;
#if !defined(__HIP_DEVICE_COMPILE__)
        HostP16::permute(state, *p16);
#else
        p2_permute<16, 20>(state, *p16, DIAG.d16);
#endif
        for (int i = 0; i < 8; i++) out_buf[i] = state[i];
        n_out = 8;
    }
    void observe(F v) {
        n_out = 0;
        in_buf[n_in++] = v;
        if (n_in == 8) duplex();
    }
    F sample() {
        if (n_in > 0 || n_out == 0) duplex();
        return out_buf[--n_out];
    }
    uint32_t sample_bits(int bits) { return from_mont(sample()) & (uint32_t)((1ull << bits) - 1); }
    EF sample_ef() { EF r; for (int i = 0; i < 8; i++) r.c[i] = sample(); return r; }
    void add_base(const F *v, size_t n) {
        for (size_t i = 0; i < n; i++) { data.push_back(from_mont(v[i])); observe(v[i]); }
    }
    void add_ef(const EF *v, size_t n) { for (size_t i = 0; i < n; i++) add_base(v[i].c, 8); }
    void hint_base(const F *v, size_t n) { for (size_t i = 0; i < n; i++) data.push_back(from_mont(v[i])); }
    void hint_canonical(const uint32_t *v, size_t n) { data.insert(data.end(), v, v + n); }
};

// ----------------------------------------------------------- WHIR schedule --
// WhirConfig::new for ConstantFromSecondRound(f0, f), CapacityBound soundness,
// 248-bit extension field (whir-p3 @ 9caef91; restated, parity unpinned).
struct WhirRound {
    int pow_bits, folding_pow_bits, num_queries, ood_samples, log_inv_rate, num_variables,
        folding_factor, log_domain_size;
};
struct WhirCfg {
    int num_variables, security_level, max_pow_bits, f0, f, starting_log_inv_rate, rs_reduction0;
    int n_rounds, final_sumcheck_rounds, commit_ood_samples, starting_folding_pow_bits;
    std::vector<WhirRound> rounds;
    int final_queries, final_pow_bits, final_folding_pow_bits, final_log_inv_rate, final_log_domain_size;
    int folding_at(int r) const { return r == 0 ? f0 : f; }
};

namespace cfgmath {
constexpr double FIELD_BITS = 248.0;
inline double log_eta(int r) { return -((double)r + 1.0); }
inline double list_size_bits(int d, int r) { return (double)(d + r) - log_eta(r); }
inline double prox_gaps_error(int d, int r, int nf) {
    return FIELD_BITS - ((double)(d + 2 * r) - log_eta(r) + log2((double)nf - 1.0));
}
inline int ood_samples(int sec, int d, int r) {
    for (int s = 1; s < 64; s++) {
        double err = 2.0 * list_size_bits(d, r) + (double)(d * s);
        if ((double)s * FIELD_BITS + 1.0 - err >= (double)sec) return s;
    }
    throw std::runtime_error("no OOD sample count reaches the security level");
}
inline int queries(int psec, int r) { return (int)ceil((double)psec / (double)r); }
inline int folding_pow(int sec, int nv, int r) {
    double e1 = prox_gaps_error(nv, r, 2);
    double e2 = FIELD_BITS - (list_size_bits(nv, r) + 1.0);
    double v = (double)sec - std::min(e1, e2);
    return v > 0 ? (int)ceil(v) : 0;
}
}  // namespace cfgmath

inline WhirCfg whir_config(int v, int sec, int max_pow, int f0, int f, int rate, int rs0) {
    using namespace cfgmath;
    if (f0 > v || f <= 0 || rs0 > f0) throw std::runtime_error("invalid WHIR folding parameters");
    WhirCfg c{};
    c.num_variables = v; c.security_level = sec; c.max_pow_bits = max_pow; c.f0 = f0; c.f = f;
    c.starting_log_inv_rate = rate; c.rs_reduction0 = rs0;
    int nv1 = v - f0;
    const int MAXC = 6;
    if (nv1 < MAXC) { c.n_rounds = 0; c.final_sumcheck_rounds = nv1; }
    else { c.n_rounds = (nv1 - MAXC + f - 1) / f; c.final_sumcheck_rounds = nv1 - c.n_rounds * f; }
    int psec = std::max(0, sec - max_pow);
    c.commit_ood_samples = ood_samples(sec, v, rate);
    c.starting_folding_pow_bits = folding_pow(sec, v, rate);
    int nv = v - f0, lir = rate, logd = v + rate;
    for (int r = 0; r < c.n_rounds; r++) {
        int rs = r == 0 ? rs0 : 1;
        int fold = c.folding_at(r);
        int next = lir + (fold - rs);
        int q = queries(psec, lir);
        int ood = ood_samples(sec, nv, next);
        double qe = (double)q * lir;
        double ce = FIELD_BITS - (log2((double)(ood + q)) + list_size_bits(nv, next) + 1.0);
        double pw = (double)sec - std::min(qe, ce);
        WhirRound rp{};
        rp.pow_bits = pw > 0 ? (int)ceil(pw) : 0;
        rp.folding_pow_bits = folding_pow(sec, nv, next);
        rp.num_queries = q; rp.ood_samples = ood; rp.log_inv_rate = lir; rp.num_variables = nv;
        rp.folding_factor = fold; rp.log_domain_size = logd;
        c.rounds.push_back(rp);
        nv -= c.folding_at(r + 1);
        lir = next;
        logd -= rs;
    }
    c.final_queries = queries(psec, lir);
    double fp = (double)sec - (double)c.final_queries * lir;
    c.final_pow_bits = fp > 0 ? (int)ceil(fp) : 0;
    double ffp = (double)sec - (FIELD_BITS - 1.0);
    c.final_folding_pow_bits = ffp > 0 ? (int)ceil(ffp) : 0;
    c.final_log_inv_rate = lir;
    c.final_log_domain_size = logd;
    return c;
}

// Domain separator pattern: (tag, count) per transcript operation; tags
// 1 observe, 2 sample, 3 hint, 4 pow.
inline std::vector<uint32_t> whir_domsep(const WhirCfg &c) {
    std::vector<uint32_t> o;
    auto op = [&](uint32_t t, uint32_t k) { o.push_back(t); o.push_back(k); };
    op(1, 8);
    if (c.commit_ood_samples) { op(2, 8 * c.commit_ood_samples); op(1, 8 * c.commit_ood_samples); }
    op(2, 8);
    for (int k = 0; k < c.f0; k++) { op(1, 24); if (c.starting_folding_pow_bits) op(4, c.starting_folding_pow_bits); op(2, 8); }
    for (int r = 0; r < c.n_rounds; r++) {
        const WhirRound &rp = c.rounds[r];
        op(1, 8);
        if (rp.ood_samples) { op(2, 8 * rp.ood_samples); op(1, 8 * rp.ood_samples); }
        if (rp.pow_bits) op(4, rp.pow_bits);
        op(2, rp.num_queries);
        op(3, 0);
        op(2, 8);
        for (int k = 0; k < c.folding_at(r + 1); k++) { op(1, 24); if (rp.folding_pow_bits) op(4, rp.folding_pow_bits); op(2, 8); }
    }
    op(1, 8u << c.final_sumcheck_rounds);
    if (c.final_pow_bits) op(4, c.final_pow_bits);
    op(2, c.final_queries);
    op(3, 0);
    for (int k = 0; k < c.final_sumcheck_rounds; k++) { op(1, 24); if (c.final_folding_pow_bits) op(4, c.final_folding_pow_bits); op(2, 8); }
    return o;
}

}  // namespace wh
