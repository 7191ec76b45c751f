// verifier.hip -- AirTable::verify (crates/air/src/verify.rs:55-249) on the host:
// the sumcheck verifiers (crates/sumcheck/src/verify.rs:23-113), the WHIR
// opening check, a batch verifier over host threads and the proof wire format.
//
// Verification reads proof_data through a verifier transcript (wl_transcript in
// verifier mode) in the prover's order.  Failures are reported with the kinds of
// AirVerifError (verify.rs:26-34): 1 InvalidPcsCommitment, 2 InvalidPcsOpening,
// 3 Fs, 4 Sumcheck, 5 InvalidBoundaryCondition, 6 SumMismatch.  The work is
// polylogarithmic in the trace (a few thousand Poseidon2 permutations for the
// Merkle paths), so it runs on CPU cores; batches spread proofs over threads.
#include <hip/hip_runtime.h>
#include <atomic>
#include <thread>
#include "../../include/whirlaway_hip.h"
#include "sc_generic.h"
#include "trace.h"

using namespace wl;

namespace wv {
using namespace kb;

enum Kind { E_PCS_COMMIT = 1, E_PCS_OPEN = 2, E_FS = 3, E_SUMCHECK = 4, E_BOUNDARY = 5, E_SUM = 6 };
struct Reject { int kind; };

// whir-p3 VerifierState operations on a verifier transcript; a failure rejects
// with the caller's error kind
struct VS {
    wl_transcript *t;
    void next_f(F *out, size_t n, int kind) {
        if (t->pos + n > t->proof.size()) throw Reject{kind};
        for (size_t i = 0; i < n; i++) {
            const uint32_t w = t->proof[t->pos + i];
            if (w >= P) throw Reject{kind};
            out[i] = to_mont(w);
            t->ts.observe(out[i]);
        }
        t->pos += n;
    }
    void next_ef(EF *out, size_t n, int kind) {
        for (size_t i = 0; i < n; i++) next_f(out[i].c, 8, kind);
    }
    std::vector<EF> next_ef(size_t n, int kind) {
        std::vector<EF> v(n);
        next_ef(v.data(), n, kind);
        return v;
    }
    void hint(F *out, size_t n, int kind) {  // read, not observed
        if (t->pos + n > t->proof.size()) throw Reject{kind};
        for (size_t i = 0; i < n; i++) {
            const uint32_t w = t->proof[t->pos + i];
            if (w >= P) throw Reject{kind};
            out[i] = to_mont(w);
        }
        t->pos += n;
    }
    void check_pow(int bits, int kind) {
        if (bits <= 0) return;
        F w;
        next_f(&w, 1, kind);
        if (t->ts.sample_bits(bits) != 0) throw Reject{kind};
    }
    EF sample() { return t->ts.sample_ef(); }
    uint32_t sample_bits(int b) { return t->ts.sample_bits(b); }
};

// ---------------------------------------------------------- polynomials --
EF mle_eval(const std::vector<EF> &evals, const EF *pt, int n) {
    const std::vector<EF> eq = wh::eq_host(std::vector<EF>(pt, pt + n));
    EF acc = ezero();
    for (size_t x = 0; x < evals.size(); x++) acc = eadd(acc, emul(eq[x], evals[x]));
    return acc;
}
EF poly_at_int(const std::vector<EF> &co, uint32_t s) { return wh::peval(co, wh::esmall(s)); }

// matrix_up_lde / matrix_down_lde / next_mle (crates/air/src/utils.rs:12-138) at
// point = (x, y), 2n coordinates, big-endian
EF matrix_up_lde(const std::vector<EF> &p) {
    const int n = (int)p.size() / 2;
    EF prod = eone();
    for (size_t i = 0; i + 1 < p.size(); i++) prod = emul(prod, p[i]);
    const EF last = p.back();
    return eadd(wh::eq_outside(p.data(), p.data() + n, n), emul(prod, esub(eone(), edbl(last))));
}
EF next_mle(const std::vector<EF> &p) {
    const int n = (int)p.size() / 2;
    const EF *x = p.data(), *y = p.data() + n;
    const EF one = eone();
    EF total = ezero();
    for (int k = 0; k < n; k++) {
        EF high = one;  // bits above k equal
        for (int i = k + 1; i < n; i++) {
            const EF a = x[n - 1 - i], b = y[n - 1 - i];
            high = emul(high, eadd(emul(a, b), emul(esub(one, a), esub(one, b))));
        }
        const EF carry = emul(esub(one, x[n - 1 - k]), y[n - 1 - k]);  // x_k = 0, y_k = 1
        EF low = one;  // bits below k: 1 in x, 0 in y
        for (int i = 0; i < k; i++) low = emul(low, emul(x[n - 1 - i], esub(one, y[n - 1 - i])));
        total = eadd(total, emul(emul(high, carry), low));
    }
    return total;
}
EF matrix_down_lde(const std::vector<EF> &p) {
    EF prod = eone();
    for (const EF &v : p) prod = emul(prod, v);
    return eadd(next_mle(p), prod);
}

// ------------------------------------------------------------- sumcheck --
// verify_core (crates/sumcheck/src/verify.rs:71-113): round r's polynomial has
// degs[r] + 1 coefficients and sums over {0 .. sets[r] - 1}; AUTO grinding
struct Evaluation { std::vector<EF> point; EF value; };
EF sumcheck_core(VS &v, const std::vector<int> &degs, const std::vector<int> &sets, int sec, Evaluation &out) {
    EF sum = ezero(), target = ezero();
    for (size_t r = 0; r < degs.size(); r++) {
        std::vector<EF> co = v.next_ef((size_t)degs[r] + 1, E_SUMCHECK);
        EF cs = ezero();
        for (int s = 0; s < sets[r]; s++) cs = eadd(cs, poly_at_int(co, (uint32_t)s));
        if (r == 0) sum = cs;
        else if (!eeq(target, cs)) throw Reject{E_SUMCHECK};
        const EF c = v.sample();
        v.check_pow(pow_bits_sumcheck(sec, degs[r]), E_SUMCHECK);
        target = wh::peval(co, c);
        out.point.push_back(c);
    }
    out.value = target;
    return sum;
}

// ----------------------------------------------------------------- WHIR --
void p16_compress(const P2Consts16 &p16, const F *l, const F *r, F *out) {
    F s[16];
    for (int i = 0; i < 8; i++) { s[i] = l[i]; s[8 + i] = r[i]; }
    p2_permute<16, 20>(s, p16, DIAG.d16);
    for (int i = 0; i < 8; i++) out[i] = s[i];
}
// PaddingFreeSponge<P24, 24, 16, 8>: each chunk of 16 overwrites the rate, then permute
void p24_leaf(const P2Consts24 &p24, const F *row, size_t len, F *out) {
    F s[24];
    for (int i = 0; i < 24; i++) s[i] = fzero();
    for (size_t off = 0; off < len; off += 16) {
        const size_t m = std::min<size_t>(16, len - off);
        for (size_t i = 0; i < m; i++) s[i] = row[off + i];
        p2_permute<24, 23>(s, p24, DIAG.d24);
    }
    for (int i = 0; i < 8; i++) out[i] = s[i];
}
bool merkle_ok(const P2Consts16 &p16, const P2Consts24 &p24, const F *root, int log_rows, uint32_t r, const F *row,
               int width, const F *path) {
    F cur[8], nxt[8];
    p24_leaf(p24, row, (size_t)width, cur);
    uint32_t idx = r;
    for (int l = 0; l < log_rows; l++) {
        if (idx & 1) p16_compress(p16, path + 8 * l, cur, nxt);
        else p16_compress(p16, cur, path + 8 * l, nxt);
        memcpy(cur, nxt, sizeof cur);
        idx >>= 1;
    }
    for (int i = 0; i < 8; i++)
        if (cur[i].v != root[i].v) return false;
    return true;
}
// a leaf (row of 2^fold F values, or 2^fold EF values flattened coefficient-wise)
// folded by eq(r, .)
EF fold_leaf(const F *row, bool base, int fold, const EF *r) {
    const std::vector<EF> eq = wh::eq_host(std::vector<EF>(r, r + fold));
    EF acc = ezero();
    for (int b = 0; b < (1 << fold); b++) {
        if (base) acc = eadd(acc, emulf(eq[b], row[b]));
        else {
            EF v;
            for (int c = 0; c < 8; c++) v.c[c] = row[8 * b + c];
            acc = eadd(acc, emul(eq[b], v));
        }
    }
    return acc;
}
std::vector<uint32_t> queries(VS &v, int n, int log_rows) {
    std::vector<uint32_t> q(n);
    for (int i = 0; i < n; i++) q[i] = v.sample_bits(log_rows);
    std::sort(q.begin(), q.end());
    q.erase(std::unique(q.begin(), q.end()), q.end());
    return q;
}
void fold_rounds(VS &v, int k, int pow_bits, EF &sum, std::vector<EF> &rs) {
    for (int i = 0; i < k; i++) {
        EF h[3];
        v.next_ef(h, 3, E_PCS_OPEN);
        if (!eeq(eadd(h[0], h[1]), sum)) throw Reject{E_PCS_OPEN};
        v.check_pow(pow_bits, E_PCS_OPEN);
        const EF r = v.sample();
        sum = wh::interp3(h, r);
        rs.push_back(r);
    }
}

struct Commitment { F root[8]; std::vector<EF> ood_points, ood_answers; };

// CommitmentReader::parse_commitment (crates/air/src/verify.rs:72-76)
Commitment parse_commitment(VS &v, const wh::WhirCfg &c) {
    Commitment cm;
    v.next_f(cm.root, 8, E_PCS_COMMIT);
    for (int i = 0; i < c.commit_ood_samples; i++) cm.ood_points.push_back(v.sample());
    cm.ood_answers = v.next_ef((size_t)c.commit_ood_samples, E_PCS_COMMIT);
    return cm;
}

// Verifier::verify of one evaluation claim f(point) = value (verify.rs:239-246):
// the folding sumcheck, the STIR queries of every round against the previous
// commitment, the final polynomial, then the combined weight at the folding point.
void whir_verify(VS &v, const wh::WhirCfg &c, const P2Consts16 &p16, const P2Consts24 &p24, const Commitment &cm,
                 const std::vector<EF> &point, const EF &value) {
    const int nv = c.num_variables;
    struct Set { size_t offset; int n; EF gamma; std::vector<std::vector<EF>> pts; };
    std::vector<Set> sets;
    std::vector<EF> rg;
    EF sum = ezero();
    {
        Set s0{0, nv, v.sample(), {}};
        EF gp = eone();
        for (size_t i = 0; i < cm.ood_points.size(); i++) {
            s0.pts.push_back(wh::expand_univ(cm.ood_points[i], nv));
            sum = eadd(sum, emul(gp, cm.ood_answers[i]));
            gp = emul(gp, s0.gamma);
        }
        s0.pts.push_back(point);
        sum = eadd(sum, emul(gp, value));
        sets.push_back(std::move(s0));
    }
    fold_rounds(v, c.f0, c.starting_folding_pow_bits, sum, rg);
    int prev_fold = c.f0, prev_log_rows = nv + c.starting_log_inv_rate - c.f0, m = nv - c.f0;
    bool prev_base = true;
    F root[8];
    memcpy(root, cm.root, sizeof root);
    std::vector<F> row, path;
    auto open = [&](const std::vector<uint32_t> &q, int log_rows, std::vector<EF> &vals) {
        const int width = prev_base ? (1 << prev_fold) : (8 << prev_fold);
        row.resize(width);
        path.resize((size_t)log_rows * 8);
        for (uint32_t x : q) {
            v.hint(row.data(), row.size(), E_PCS_OPEN);
            v.hint(path.data(), path.size(), E_PCS_OPEN);
            if (!merkle_ok(p16, p24, root, log_rows, x, row.data(), width, path.data())) throw Reject{E_PCS_OPEN};
            vals.push_back(fold_leaf(row.data(), prev_base, prev_fold, rg.data() + rg.size() - prev_fold));
        }
    };
    for (int r = 0; r < c.n_rounds; r++) {
        const wh::WhirRound &rp = c.rounds[r];
        const int rs = r == 0 ? c.rs_reduction0 : 1, kn = c.folding_at(r + 1);
        F nroot[8];
        v.next_f(nroot, 8, E_PCS_OPEN);
        std::vector<EF> oodp;
        for (int i = 0; i < rp.ood_samples; i++) oodp.push_back(v.sample());
        std::vector<EF> ooda = v.next_ef((size_t)rp.ood_samples, E_PCS_OPEN);
        v.check_pow(rp.pow_bits, E_PCS_OPEN);
        const int log_rows = rp.log_domain_size - prev_fold;
        if (log_rows != prev_log_rows) throw Reject{E_PCS_OPEN};
        const std::vector<uint32_t> q = queries(v, rp.num_queries, log_rows);
        std::vector<EF> vals;
        open(q, log_rows, vals);
        Set cs{rg.size(), m, v.sample(), {}};
        EF gp = eone();
        for (int i = 0; i < rp.ood_samples; i++) {
            cs.pts.push_back(wh::expand_univ(oodp[i], m));
            sum = eadd(sum, emul(gp, ooda[i]));
            gp = emul(gp, cs.gamma);
        }
        const F gen = fpow(fsmall(3), (uint64_t)(P - 1) >> log_rows);
        for (size_t i = 0; i < q.size(); i++) {
            cs.pts.push_back(wh::expand_univ(efrom(fpow(gen, q[i])), m));
            sum = eadd(sum, emul(gp, vals[i]));
            gp = emul(gp, cs.gamma);
        }
        sets.push_back(std::move(cs));
        fold_rounds(v, kn, rp.folding_pow_bits, sum, rg);
        m -= kn;
        memcpy(root, nroot, sizeof root);
        prev_log_rows = rp.log_domain_size - rs - kn;
        prev_fold = kn;
        prev_base = false;
    }
    const int fv = c.final_sumcheck_rounds;
    if (fv != m) throw Reject{E_PCS_OPEN};
    const std::vector<EF> fin = v.next_ef((size_t)1 << fv, E_PCS_OPEN);
    v.check_pow(c.final_pow_bits, E_PCS_OPEN);
    const int log_rows = c.final_log_domain_size - prev_fold;
    if (log_rows != prev_log_rows) throw Reject{E_PCS_OPEN};
    const std::vector<uint32_t> q = queries(v, c.final_queries, log_rows);
    std::vector<EF> vals;
    open(q, log_rows, vals);
    const F gen = fpow(fsmall(3), (uint64_t)(P - 1) >> log_rows);
    for (size_t i = 0; i < q.size(); i++) {
        const std::vector<EF> pt = wh::expand_univ(efrom(fpow(gen, q[i])), fv);
        if (!eeq(vals[i], mle_eval(fin, pt.data(), fv))) throw Reject{E_PCS_OPEN};
    }
    fold_rounds(v, fv, c.final_folding_pow_bits, sum, rg);
    if ((int)rg.size() != nv) throw Reject{E_PCS_OPEN};
    const EF fval = mle_eval(fin, rg.data() + rg.size() - fv, fv);
    EF W = ezero();
    for (const Set &s : sets) {
        EF gp = eone();
        for (const auto &pt : s.pts) {
            W = eadd(W, emul(gp, wh::eq_outside(pt.data(), rg.data() + s.offset, s.n)));
            gp = emul(gp, s.gamma);
        }
    }
    if (!eeq(sum, emul(fval, W))) throw Reject{E_PCS_OPEN};
}

// ------------------------------------------------------------ constraints --
// The Poseidon2 AIR's constraints at an EF point (local row x[0..165)), in the
// AIR's constraint order, folded by the batching scalars ap.
void mds_light(EF *s) {
    for (int c = 0; c < 16; c += 4) {
        EF *x = s + c;
        const EF t01 = eadd(x[0], x[1]), t23 = eadd(x[2], x[3]), t0123 = eadd(t01, t23);
        const EF t01123 = eadd(t0123, x[1]), t01233 = eadd(t0123, x[3]);
        const EF y0 = eadd(t01123, t01), y1 = eadd(t01123, edbl(x[2]));
        const EF y2 = eadd(t01233, t23), y3 = eadd(t01233, edbl(x[0]));
        x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
    }
    EF col[4];
    for (int i = 0; i < 4; i++) col[i] = eadd(eadd(s[i], s[i + 4]), eadd(s[i + 8], s[i + 12]));
    for (int i = 0; i < 16; i++) s[i] = eadd(s[i], col[i & 3]);
}
EF p2_constraints(const wk::AirConsts &air, const EF *x, const std::vector<EF> &ap) {
    EF acc = ezero();
    size_t k = 0;
    auto emit = [&](const EF &v) { acc = eadd(acc, emul(ap[k++], v)); };
    emit(emul(x[0], esub(x[0], eone())));
    EF s[16];
    for (int i = 0; i < 16; i++) s[i] = x[1 + i];
    mds_light(s);
    int col = 17;
    const F *rc = air.rc;
    auto full = [&](const F *rcr) {
        for (int i = 0; i < 16; i++) {
            const EF t = eadd(s[i], efrom(rcr[i]));
            s[i] = emul(emul(t, t), t);
        }
        mds_light(s);
        for (int i = 0; i < 16; i++) { emit(esub(s[i], x[col + i])); s[i] = x[col + i]; }
        col += 16;
    };
    for (int r = 0; r < 4; r++) full(rc + 16 * r);
    for (int r = 0; r < 20; r++) {
        const EF t = eadd(s[0], efrom(rc[64 + r]));
        emit(esub(emul(emul(t, t), t), x[col]));
        s[0] = x[col++];
        EF tot = ezero();
        for (int i = 0; i < 16; i++) tot = eadd(tot, s[i]);
        for (int i = 0; i < 16; i++) s[i] = eadd(tot, emulf(s[i], F{DIAG.d16[i]}));
    }
    for (int r = 0; r < 4; r++) full(rc + 84 + 16 * r);
    return acc;
}
// a constraint program (SSA) at an EF point
EF program_constraints(const std::vector<uint32_t> &code, const EF *x, int n_inputs, const std::vector<EF> &ap) {
    const size_t ni = code.size() / 4;
    std::vector<EF> val((size_t)n_inputs + ni);
    for (int i = 0; i < n_inputs; i++) val[i] = x[i];
    EF acc = ezero();
    size_t k = 0;
    for (size_t j = 0; j < ni; j++) {
        const uint32_t *w = &code[4 * j];
        EF r = ezero();
        switch (w[0]) {
        case wg::OP_CONST: r = efrom(to_mont(w[1])); break;
        case wg::OP_ADD: r = eadd(val[w[1]], val[w[2]]); break;
        case wg::OP_SUB: r = esub(val[w[1]], val[w[2]]); break;
        case wg::OP_MUL: r = emul(val[w[1]], val[w[2]]); break;
        case wg::OP_NEG: r = eneg(val[w[1]]); break;
        default: acc = eadd(acc, emul(ap[k++], val[w[1]])); break;
        }
        val[n_inputs + j] = r;
    }
    return acc;
}

// column_up / column_down of a canonical column, folded by the outer selectors and
// evaluated at the rest of the outer point (verify.rs:114-136)
EF preprocessed_eval(const uint32_t *col, int n, int skips, bool down, const std::vector<EF> &osel, const EF *rest) {
    const size_t N = (size_t)1 << n, M = N >> skips;
    std::vector<EF> folded(M);
    for (size_t i = 0; i < M; i++) {
        EF acc = ezero();
        for (int j = 0; j < (1 << skips); j++) {
            const size_t x = (size_t)j * M + i;
            const uint32_t w = down ? col[x == N - 1 ? N - 1 : x + 1] : col[x == N - 1 ? N - 2 : x];
            acc = eadd(acc, emulf(osel[j], to_mont(w)));
        }
        folded[i] = acc;
    }
    return mle_eval(folded, rest, n - skips);
}

// ------------------------------------------------------------- AIR verify --
void air_verify(const wl_table *t, const wl_air_settings *st, const P2Consts16 &p16, const P2Consts24 &p24, VS &v) {
    wl::Span span("air table: verify");  // crates/air/src/verify.rs:54
    const int n = (int)t->log_length, skips = (int)st->univariate_skips, sec = (int)st->security_bits;
    const wg::AirSpec *g = t->gen;
    const int C = g ? g->n_columns : wk::AIR_NCOLS, np = g ? g->n_preprocessed : 0, nw = C - np;
    const int degree = g ? g->degree : 3;
    int ncons = wk::AIR_NCONS;
    if (g) {
        ncons = 0;
        for (size_t j = 0; j < g->code.size(); j += 4) ncons += g->code[j] == wg::OP_ASSERT;
    }
    const int lc = (int)wh::log2_up((size_t)nw), K = 1 << skips;
    const wh::WhirCfg cfg = wh::whir_config(n + lc, sec, 16, (int)st->folding_factor_first, (int)st->folding_factor,
                                        (int)st->log_inv_rate, (int)st->rs_domain_initial_reduction,
                                        (int)st->soundness);
    const Commitment cm = parse_commitment(v, cfg);
    v.check_pow(pow_bits_air(sec, (int)wh::log2_up((size_t)std::max(1, ncons))), E_FS);
    const EF alpha = v.sample();
    v.check_pow(pow_bits_air(sec, n), E_FS);
    std::vector<EF> zc(n - skips + 1);
    for (auto &z : zc) z = v.sample();
    // zerocheck (verify_with_univariate_skip, degree constraint_degree + 1)
    std::vector<int> degs(1, (degree + 1) * (K - 1)), sets(1, K);
    for (int i = 0; i < n - skips; i++) { degs.push_back(degree + 1); sets.push_back(2); }
    Evaluation outer;
    const EF scsum = sumcheck_core(v, degs, sets, sec, outer);
    if (!eeq(scsum, ezero())) throw Reject{E_SUM};
    const std::vector<EF> wup = v.next_ef((size_t)nw, E_FS), wdn = v.next_ef((size_t)nw, E_FS);
    const std::vector<EF> osel = wh::selectors_ef(skips, outer.point[0]);
    std::vector<EF> gp(2 * (size_t)C);
    for (int c = 0; c < np; c++) {
        const uint32_t *col = g->pre_host.data() + (size_t)c * ((size_t)1 << n);
        gp[c] = preprocessed_eval(col, n, skips, false, osel, outer.point.data() + 1);
        gp[C + c] = preprocessed_eval(col, n, skips, true, osel, outer.point.data() + 1);
    }
    for (int c = 0; c < nw; c++) { gp[np + c] = wup[c]; gp[C + np + c] = wdn[c]; }
    std::vector<EF> ap(std::max(1, ncons));
    ap[0] = eone();
    for (int i = 1; i < ncons; i++) ap[i] = emul(ap[i - 1], alpha);
    const EF gce = g ? program_constraints(g->code, gp.data(), 2 * C, ap) : p2_constraints(t->air_host, gp.data(), ap);
    const std::vector<EF> zsel = wh::selectors_ef(skips, zc[0]);
    EF dot = ezero();
    for (int i = 0; i < K; i++) dot = eadd(dot, emul(zsel[i], osel[i]));
    const EF lhs = emul(emul(dot, wh::eq_outside(zc.data() + 1, outer.point.data() + 1, n - skips)), gce);
    if (!eeq(lhs, outer.value)) throw Reject{E_SUM};
    v.check_pow(pow_bits_air(sec, lc), E_FS);
    std::vector<EF> cbs(lc);
    for (auto &x : cbs) x = v.sample();
    const EF alpha2 = v.sample();
    const std::vector<EF> sub = v.next_ef((size_t)K, E_FS);
    const std::vector<EF> eqc = wh::eq_host(cbs);
    EF l = ezero(), ru = ezero(), rd = ezero();
    for (int i = 0; i < K; i++) l = eadd(l, emul(sub[i], osel[i]));
    for (int c = 0; c < nw; c++) { ru = eadd(ru, emul(wup[c], eqc[c])); rd = eadd(rd, emul(wdn[c], eqc[c])); }
    if (!eeq(l, eadd(ru, emul(rd, alpha2)))) throw Reject{E_SUM};
    std::vector<EF> eps(skips);
    for (auto &x : eps) x = v.sample();
    Evaluation inner;
    const EF isum = sumcheck_core(v, std::vector<int>(n, 2), std::vector<int>(n, 2), sec, inner);
    if (!eeq(isum, mle_eval(sub, eps.data(), skips))) throw Reject{E_SUM};
    std::vector<EF> lde(eps);
    lde.insert(lde.end(), outer.point.begin() + 1, outer.point.end());
    lde.insert(lde.end(), inner.point.begin(), inner.point.end());
    const EF den = eadd(matrix_up_lde(lde), emul(alpha2, matrix_down_lde(lde)));
    if (eeq(den, ezero())) throw Reject{E_SUM};
    const EF expected = emul(inner.value, wh::einv(den));
    std::vector<EF> fp(cbs);
    fp.insert(fp.end(), inner.point.begin(), inner.point.end());
    whir_verify(v, cfg, p16, p24, cm, fp, expected);
}

int verify_one(const wl_table *t, const wl_air_settings *st, const P2Dev &ph, wl_transcript *vt, int *error) {
    VS v{vt};
    try {
        air_verify(t, st, ph.p16, ph.p24, v);
    } catch (const Reject &r) {
        if (error) *error = r.kind;
        wl_set_error("proof rejected");
        return WL_ERR_VERIFY;
    }
    if (error) *error = 0;
    return WL_OK;
}

// ------------------------------------------------------------ wire format --
constexpr uint32_t MAGIC = 0x31504c57u;  // "WLP1"
constexpr size_t HEADER_WORDS = 12;      // magic, version, log_length, n_witness_columns, settings[7], n_words
uint32_t fnv1a(const uint8_t *p, size_t n) {
    uint32_t h = 2166136261u;
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 16777619u; }
    return h;
}
size_t packed_bytes(size_t n_words) { return (31 * n_words + 7) / 8; }

}  // namespace wv

extern "C" {

int wl_air_verify(const wl_table *t, const wl_air_settings *st, const uint32_t *p16c, const uint32_t *p24c,
                  wl_transcript *verifier_state, int *error) {
    if (error) *error = 0;
    if (!t || !st || !p16c || !p24c || !verifier_state || !verifier_state->verifier) {
        wl_set_error("bad arguments (a verifier transcript is required)");
        return WL_ERR_ARG;
    }
    const uint32_t nw = t->gen ? (uint32_t)(t->gen->n_columns - t->gen->n_preprocessed) : (uint32_t)wk::AIR_NCOLS;
    if (st->univariate_skips != t->univariate_skips || wl_air_settings_validate(t->log_length, nw, st) != WL_OK) {
        wl_set_error("unsupported settings");
        return WL_ERR_ARG;
    }
    P2Dev ph;
    load_p2(p16c, p24c, ph);
    return wv::verify_one(t, st, ph, verifier_state, error);
}

int wl_air_verify_batch(const wl_table *t, const wl_air_settings *st, const uint32_t *p16c, const uint32_t *p24c,
                        const uint32_t *const *proofs, const size_t *proof_lens, size_t n_proofs, int n_threads,
                        int *results) {
    if (!t || !st || !p16c || !p24c || (n_proofs && (!proofs || !proof_lens || !results))) {
        wl_set_error("null argument");
        return WL_ERR_ARG;
    }
    const uint32_t nw = t->gen ? (uint32_t)(t->gen->n_columns - t->gen->n_preprocessed) : (uint32_t)wk::AIR_NCOLS;
    if (st->univariate_skips != t->univariate_skips || wl_air_settings_validate(t->log_length, nw, st) != WL_OK) {
        wl_set_error("unsupported settings");
        return WL_ERR_ARG;
    }
    P2Dev ph;
    load_p2(p16c, p24c, ph);
    const wh::WhirCfg cfg = wh::whir_config((int)t->log_length + (int)wh::log2_up(nw), (int)st->security_bits, 16,
                                        (int)st->folding_factor_first, (int)st->folding_factor, (int)st->log_inv_rate,
                                        (int)st->rs_domain_initial_reduction, (int)st->soundness);
    const std::vector<uint32_t> domsep = wh::whir_domsep(cfg);
    const int nt = std::max(1, std::min<int>(n_threads > 0 ? n_threads : 1, (int)std::max<size_t>(1, n_proofs)));
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n_proofs;) {
            wl_transcript *vt = wl_transcript_new_verifier(p16c, domsep.data(), domsep.size(), proofs[i], proof_lens[i]);
            if (!vt) { results[i] = wv::E_FS; continue; }
            int err = 0;
            if (wv::verify_one(t, st, ph, vt, &err) == WL_OK && vt->pos != vt->proof.size()) err = wv::E_FS;
            results[i] = err;
            wl_transcript_free(vt);
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < nt; k++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return WL_OK;
}

size_t wl_proof_serialized_size(size_t proof_len) {
    return wv::HEADER_WORDS * 4 + wv::packed_bytes(proof_len) + 4;
}

int wl_proof_serialize(const wl_air_settings *st, uint32_t log_length, uint32_t n_witness_columns,
                       const uint32_t *proof, size_t proof_len, uint8_t *out, size_t cap, size_t *out_len) {
    if (!st || (proof_len && !proof) || proof_len > 0xffffffffu) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    const size_t total = wl_proof_serialized_size(proof_len);
    if (out_len) *out_len = total;
    if (!out) return WL_OK;
    if (cap < total) { wl_set_error("buffer too small"); return WL_ERR_BUFFER; }
    const uint32_t head[wv::HEADER_WORDS] = {wv::MAGIC, 1u, log_length, n_witness_columns, st->security_bits,
                                             st->folding_factor_first, st->folding_factor, st->log_inv_rate,
                                             st->univariate_skips, st->rs_domain_initial_reduction, st->soundness,
                                             (uint32_t)proof_len};
    for (size_t i = 0; i < wv::HEADER_WORDS; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(head[i] >> (8 * b));
    // 31 bits per field element, little-endian bit order
    uint8_t *p = out + wv::HEADER_WORDS * 4;
    const size_t nb = wv::packed_bytes(proof_len);
    memset(p, 0, nb);
    uint64_t acc = 0;
    int bits = 0;
    size_t o = 0;
    for (size_t i = 0; i < proof_len; i++) {
        if (proof[i] >= P) { wl_set_error("proof word is not a canonical field element"); return WL_ERR_ARG; }
        acc |= (uint64_t)proof[i] << bits;
        bits += 31;
        while (bits >= 8) { p[o++] = (uint8_t)acc; acc >>= 8; bits -= 8; }
    }
    if (bits) p[o++] = (uint8_t)acc;
    const uint32_t h = wv::fnv1a(out, total - 4);
    for (int b = 0; b < 4; b++) out[total - 4 + b] = (uint8_t)(h >> (8 * b));
    return WL_OK;
}

int wl_proof_deserialize(const uint8_t *in, size_t in_len, wl_air_settings *st, uint32_t *log_length,
                         uint32_t *n_witness_columns, uint32_t *proof, size_t cap, size_t *proof_len) {
    if (!in || in_len < wv::HEADER_WORDS * 4 + 4) { wl_set_error("truncated proof"); return WL_ERR_ARG; }
    uint32_t head[wv::HEADER_WORDS];
    for (size_t i = 0; i < wv::HEADER_WORDS; i++) {
        head[i] = 0;
        for (int b = 0; b < 4; b++) head[i] |= (uint32_t)in[4 * i + b] << (8 * b);
    }
    if (head[0] != wv::MAGIC || head[1] != 1u) { wl_set_error("not a serialized proof (magic or version)"); return WL_ERR_ARG; }
    const size_t n = head[11];
    if (in_len != wl_proof_serialized_size(n)) { wl_set_error("serialized proof length mismatch"); return WL_ERR_ARG; }
    uint32_t h = 0;
    for (int b = 0; b < 4; b++) h |= (uint32_t)in[in_len - 4 + b] << (8 * b);
    if (h != wv::fnv1a(in, in_len - 4)) { wl_set_error("serialized proof checksum mismatch"); return WL_ERR_ARG; }
    if (proof_len) *proof_len = n;
    if (log_length) *log_length = head[2];
    if (n_witness_columns) *n_witness_columns = head[3];
    if (st) *st = wl_air_settings{head[4], head[5], head[6], head[7], head[8], head[9], head[10]};
    if (!proof) return WL_OK;
    if (cap < n) { wl_set_error("buffer too small"); return WL_ERR_BUFFER; }
    const uint8_t *p = in + wv::HEADER_WORDS * 4;
    uint64_t acc = 0;
    int bits = 0;
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        while (bits < 31) { acc |= (uint64_t)p[o++] << bits; bits += 8; }
        proof[i] = (uint32_t)(acc & 0x7fffffffu);
        acc >>= 31;
        bits -= 31;
        if (proof[i] >= P) { wl_set_error("serialized word is not a canonical field element"); return WL_ERR_ARG; }
    }
    return WL_OK;
}

}  // extern "C"
