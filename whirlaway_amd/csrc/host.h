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
// ------------------------------------------ interpolation on 0, 1, .., n-1 --
// Coefficient rows of the Lagrange basis on the integer nodes 0..n-1 (n <= 64):
// row j is L_j(X) = prod_{i != j} (X - i) / (j - i), low degree first, from the
// node polynomial M(X) = prod_i (X - i) divided by (X - j).  Built once per n and
// thread (the batch engine proves on several threads).
inline const std::vector<F> &int_basis(int n) {
    static thread_local std::vector<std::vector<F>> cache(65);
    if (n < 1 || n > 64) throw std::runtime_error("interpolation on more than 64 nodes");
    std::vector<F> &B = cache[n];
    if (!B.empty()) return B;
    std::vector<F> m(n + 1, fzero());
    m[0] = fone();
    for (int i = 0; i < n; i++) {
        const F xi = fsmall(i);
        for (int d = i + 1; d > 0; d--) m[d] = fsub(m[d - 1], fmul(xi, m[d]));
        m[0] = fneg(fmul(xi, m[0]));
    }
    B.assign((size_t)n * n, fzero());
    for (int j = 0; j < n; j++) {
        F *row = &B[(size_t)j * n];
        const F xj = fsmall(j);
        F carry = m[n], den = fone();
        for (int d = n - 1; d >= 0; d--) { row[d] = carry; carry = fadd(m[d], fmul(carry, xj)); }
        for (int i = 0; i < n; i++)
            if (i != j) den = fmul(den, fsub(xj, fsmall(i)));
        const F inv = finv(den);
        for (int d = 0; d < n; d++) row[d] = fmul(row[d], inv);
    }
    return B;
}
// sum_i w[i stride] v[i]: EF x F products summed per coefficient in 64 bits (folded
// after every four, as in emul), one reduction per coefficient
inline EF lin_ef(const F *w, size_t stride, const EF *v, int n) {
    uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        const uint64_t wi = w[(size_t)i * stride].v;
        for (int k = 0; k < 8; k++) acc[k] += wi * v[i].c[k].v;
        if ((i & 3) == 3)
            for (int k = 0; k < 8; k++) acc[k] = fold64(acc[k]);
    }
    EF r;
    for (int k = 0; k < 8; k++) r.c[k] = F{redc64(acc[k])};
    return r;
}
// coefficients of the polynomial of degree < n through (i, ys[i]), i = 0..n-1
inline std::vector<EF> interp_int(const std::vector<EF> &ys) {
    const int n = (int)ys.size();
    const std::vector<F> &B = int_basis(n);
    std::vector<EF> co(n);
    for (int d = 0; d < n; d++) co[d] = lin_ef(&B[d], (size_t)n, ys.data(), n);
    return co;
}
// value at the integer z of basis polynomial j on the nodes 0..n-1
inline F int_basis_at(int n, int j, int z) {
    const F *row = &int_basis(n)[(size_t)j * n];
    const F x = fsmall((uint32_t)z);
    F acc = fzero();
    for (int d = n - 1; d >= 0; d--) acc = fadd(fmul(acc, x), row[d]);
    return acc;
}

// univariate_selectors(k) at an extension point: L_i(x) on the nodes 0..2^k-1 as
// (prefix product) (suffix product) / prod_{j != i} (i - j), the denominators'
// inverses cached per k
inline std::vector<EF> selectors_ef(int k, const EF &x) {
    const int N = 1 << k;
    static thread_local std::vector<std::vector<F>> dinv_cache(32);
    std::vector<F> &dinv = dinv_cache[k];
    if (dinv.empty()) {
        dinv.resize(N);
        for (int i = 0; i < N; i++) {
            F den = fone();
            for (int j = 0; j < N; j++)
                if (j != i) den = fmul(den, fsub(fsmall(i), fsmall(j)));
            dinv[i] = finv(den);
        }
    }
    std::vector<EF> d(N), pre(N + 1), suf(N + 1), out(N);
    for (int j = 0; j < N; j++) d[j] = esub(x, esmall(j));
    pre[0] = eone();
    for (int j = 0; j < N; j++) pre[j + 1] = emul(pre[j], d[j]);
    suf[N] = eone();
    for (int j = N - 1; j >= 0; j--) suf[j] = emul(suf[j + 1], d[j]);
    for (int i = 0; i < N; i++) out[i] = emulf(emul(pre[i], suf[i + 1]), dinv[i]);
    return out;
}

// Round 0 of the zerocheck with k univariate skips, K = 2^k: the round polynomial
// is f = p q, where p (degree 3(K-1)) is 0 on 0..K-1 and vals on K..3(K-1), and q
// (degree K-1) takes sel[i] at i = 0..K-1.  f is formed from its values on
// 0..4(K-1) (zero below K; p extrapolated past 3(K-1), q evaluated past K-1 through
// cached basis values) and interpolated there: the same coefficients as
// interpolating p and q and multiplying them.
struct SkipTables {
    std::vector<F> q_at;   // [z - K][i]: L_i on 0..K-1 at z = K..4(K-1)
    std::vector<F> p_at;   // [z - D - 1][j - K]: L_j on 0..D at z = D+1..4(K-1), j = K..D
    std::vector<F> coef;   // [d][z - K]: coefficient d of L_z on 0..4(K-1), z = K..4(K-1)
};
inline const SkipTables &skip_tables(int k) {
    static thread_local std::vector<SkipTables> cache(8);
    if (k < 1 || k > 4) throw std::runtime_error("univariate skips beyond 4 use the generic path");
    SkipTables &t = cache[k];
    if (!t.coef.empty()) return t;
    const int K = 1 << k, D = 3 * (K - 1), Df = 4 * (K - 1), NZ = Df - K + 1;
    t.q_at.resize((size_t)NZ * K);
    for (int z = K; z <= Df; z++)
        for (int i = 0; i < K; i++) t.q_at[(size_t)(z - K) * K + i] = int_basis_at(K, i, z);
    t.p_at.resize((size_t)(Df - D) * (D - K + 1));
    for (int z = D + 1; z <= Df; z++)
        for (int j = K; j <= D; j++) t.p_at[(size_t)(z - D - 1) * (D - K + 1) + (j - K)] = int_basis_at(D + 1, j, z);
    const std::vector<F> &B = int_basis(Df + 1);
    t.coef.resize((size_t)(Df + 1) * NZ);
    for (int d = 0; d <= Df; d++)
        for (int z = K; z <= Df; z++) t.coef[(size_t)d * NZ + (z - K)] = B[(size_t)z * (Df + 1) + d];
    return t;
}
inline std::vector<EF> skip_round_poly(int k, const std::vector<EF> &vals, const std::vector<EF> &sel) {
    const SkipTables &t = skip_tables(k);
    const int K = 1 << k, D = 3 * (K - 1), Df = 4 * (K - 1), NZ = Df - K + 1;
    if ((int)vals.size() != D - K + 1 || (int)sel.size() != K) throw std::runtime_error("skip round sizes");
    std::vector<EF> fz(NZ);
    for (int z = K; z <= Df; z++) {
        const EF q = lin_ef(&t.q_at[(size_t)(z - K) * K], 1, sel.data(), K);
        const EF p = z <= D ? vals[z - K] : lin_ef(&t.p_at[(size_t)(z - D - 1) * (D - K + 1)], 1, vals.data(), D - K + 1);
        fz[z - K] = emul(p, q);
    }
    std::vector<EF> co(Df + 1);
    for (int d = 0; d <= Df; d++) co[d] = lin_ef(&t.coef[(size_t)d * NZ], 1, fz.data(), NZ);
    return co;
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
        data.reserve((size_t)1 << 16);  // a 2^17-row proof holds ~55K words
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
// WhirConfig::new for ConstantFromSecondRound(f0, f) and a SecurityAssumption,
// 248-bit extension field (whir-p3 @ 9caef91 parameters/errors.rs; restated,
// parity unpinned).
enum Soundness { SA_UNIQUE_DECODING = 0, SA_JOHNSON_BOUND = 1, SA_CAPACITY_BOUND = 2 };
struct WhirRound {
    int pow_bits, folding_pow_bits, num_queries, ood_samples, log_inv_rate, num_variables,
        folding_factor, log_domain_size;
};
struct WhirCfg {
    int num_variables, security_level, max_pow_bits, f0, f, starting_log_inv_rate, rs_reduction0, soundness;
    int n_rounds, final_sumcheck_rounds, commit_ood_samples, starting_folding_pow_bits;
    std::vector<WhirRound> rounds;
    int final_queries, final_pow_bits, final_folding_pow_bits, final_log_inv_rate, final_log_domain_size;
    int folding_at(int r) const { return r == 0 ? f0 : f; }
};

// Error terms of the round-by-round soundness analysis per decoding regime, in
// bits.  eta, the proximity slack: sqrt(rho)/20 under the Johnson bound, rho/2
// at capacity, none for unique decoding; r = log2(1/rho).
namespace cfgmath {
constexpr double FIELD_BITS = 248.0, LOG2_10 = 3.321928094887362;
inline double log_eta(int sa, int r) {
    if (sa == SA_UNIQUE_DECODING) return 0.0;
    if (sa == SA_JOHNSON_BOUND) return -(0.5 * r + LOG2_10 + 1.0);
    return -((double)r + 1.0);
}
inline double list_size_bits(int sa, int d, int r) {
    if (sa == SA_UNIQUE_DECODING) return 0.0;
    if (sa == SA_JOHNSON_BOUND) return 0.5 * r - (1.0 + log_eta(sa, r));
    return (double)(d + r) - log_eta(sa, r);
}
inline double prox_gaps_error(int sa, int d, int r, int nf) {
    double err;
    if (sa == SA_UNIQUE_DECODING) err = (double)(d + r);
    else if (sa == SA_JOHNSON_BOUND) err = 2.0 * d + 7.0 * (std::min(1.0 + LOG2_10 + 0.5 * r, -log_eta(sa, r)) - 1.0);
    else err = (double)(d + 2 * r) - log_eta(sa, r);
    return FIELD_BITS - (err + log2((double)nf - 1.0));
}
inline int ood_samples(int sa, int sec, int d, int r) {
    if (sa == SA_UNIQUE_DECODING) return 0;
    for (int s = 1; s < 64; s++) {
        double err = 2.0 * list_size_bits(sa, d, r) + (double)(d * s);
        if ((double)s * FIELD_BITS + 1.0 - err >= (double)sec) return s;
    }
    throw std::runtime_error("no OOD sample count reaches the security level");
}
// bits of soundness per query: -log2(1 - delta)
inline double query_bits(int sa, int r) {
    if (sa == SA_UNIQUE_DECODING) return -log2(0.5 * (1.0 + 1.0 / (double)(1 << r)));
    if (sa == SA_JOHNSON_BOUND) return 0.5 * r;
    return (double)r;
}
inline int queries(int sa, int psec, int r) { return (int)ceil((double)psec / query_bits(sa, r)); }
inline int folding_pow(int sa, int sec, int nv, int r) {
    double e1 = prox_gaps_error(sa, nv, r, 2);
    double e2 = FIELD_BITS - (list_size_bits(sa, nv, r) + 1.0);
    double v = (double)sec - std::min(e1, e2);
    return v > 0 ? (int)ceil(v) : 0;
}
}  // namespace cfgmath

inline WhirCfg whir_config(int v, int sec, int max_pow, int f0, int f, int rate, int rs0,
                           int sa = SA_CAPACITY_BOUND) {
    using namespace cfgmath;
    if (f0 > v || f <= 0 || rs0 > f0 || rate < 1) throw std::runtime_error("invalid WHIR folding parameters");
    if (sa < SA_UNIQUE_DECODING || sa > SA_CAPACITY_BOUND) throw std::runtime_error("unknown soundness assumption");
    WhirCfg c{};
    c.num_variables = v; c.security_level = sec; c.max_pow_bits = max_pow; c.f0 = f0; c.f = f;
    c.starting_log_inv_rate = rate; c.rs_reduction0 = rs0; c.soundness = sa;
    int nv1 = v - f0;
    const int MAXC = 6;
    if (nv1 < MAXC) { c.n_rounds = 0; c.final_sumcheck_rounds = nv1; }
    else { c.n_rounds = (nv1 - MAXC + f - 1) / f; c.final_sumcheck_rounds = nv1 - c.n_rounds * f; }
    int psec = std::max(0, sec - max_pow);
    c.commit_ood_samples = ood_samples(sa, sec, v, rate);
    c.starting_folding_pow_bits = folding_pow(sa, sec, v, rate);
    int nv = v - f0, lir = rate, logd = v + rate;
    for (int r = 0; r < c.n_rounds; r++) {
        int rs = r == 0 ? rs0 : 1;
        int fold = c.folding_at(r);
        int next = lir + (fold - rs);
        int q = queries(sa, psec, lir);
        int ood = ood_samples(sa, sec, nv, next);
        double qe = (double)q * query_bits(sa, lir);
        double ce = FIELD_BITS - (log2((double)(ood + q)) + list_size_bits(sa, nv, next) + 1.0);
        double pw = (double)sec - std::min(qe, ce);
        WhirRound rp{};
        rp.pow_bits = pw > 0 ? (int)ceil(pw) : 0;
        rp.folding_pow_bits = folding_pow(sa, sec, nv, next);
        rp.num_queries = q; rp.ood_samples = ood; rp.log_inv_rate = lir; rp.num_variables = nv;
        rp.folding_factor = fold; rp.log_domain_size = logd;
        c.rounds.push_back(rp);
        nv -= c.folding_at(r + 1);
        lir = next;
        logd -= rs;
    }
    c.final_queries = queries(sa, psec, lir);
    double fp = (double)sec - (double)c.final_queries * query_bits(sa, lir);
    c.final_pow_bits = fp > 0 ? (int)ceil(fp) : 0;
    double ffp = (double)sec - (FIELD_BITS - 1.0);
    c.final_folding_pow_bits = ffp > 0 ? (int)ceil(ffp) : 0;
    c.final_log_inv_rate = lir;
    c.final_log_domain_size = logd;
    // every codeword is evaluated on a power-of-two subgroup of F*, at most 2^24
    bool fits = v + rate - f0 <= TWO_ADICITY;
    for (int r = 0; r < c.n_rounds; r++)
        fits = fits && c.rounds[r].log_domain_size - (r == 0 ? rs0 : 1) - c.folding_at(r + 1) <= TWO_ADICITY;
    if (!fits) throw std::runtime_error("a codeword domain exceeds 2^24, the field's two-adicity");
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
