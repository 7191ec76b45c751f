// prover.hip -- MI355X prover for the Poseidon2 AIR (Whirlaway hot path) and
// the C ABI declared in include/whirlaway_hip.h.
//
// Protocol order follows crates/air/src/prove.rs:49-226 step by step:
//   WHIR commit (prove.rs:80-91) -> alpha (:93-100) -> zerocheck challenges
//   (:102-107) -> zerocheck sumcheck with univariate skip (:114-131,
//   crates/sumcheck/src/prove.rs:22-254) -> inner sums (:135-145) -> column
//   batching (:147-160) -> alpha2, sub_evals (:162-173) -> epsilons, inner
//   sumcheck (:175-208) -> WHIR opening (:216-225).
// The host side runs only the transcript and polylog-sized algebra; every pass
// over trace- or codeword-sized data is a HIP kernel on the context's stream.
#include <hip/hip_runtime.h>
#include <chrono>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include "../../include/whirlaway_hip.h"
#include "ctx.h"
#include "devts.h"
#include "sc_generic.h"
#include "trace.h"
#include "tune.h"

using namespace kb;
using wh::Transcript;
using wh::WhirCfg;

using namespace wl;

namespace {

using wk::AIR_NCOLS;
using wk::AIR_NCONS;

// ------------------------------------------------------------------ WHIR --
struct Codeword {
    DBuf<F> cw;      // column-major [width][rows]
    DBuf<F> nodes;   // heap [2 rows][8]
    int width = 0;   // F per leaf
    int nnz = 0;     // columns [nnz, width) are zero and not stored
    int log_rows = 0;
    int la = 0;      // storage order of the rows (rs_encode)
    bool base = true;
    int fold = 0;
};

void commit_codeword(wl_ctx *c, const P2Dev *pdev, Codeword &cwd, const F *cols, int n_cols,
                     int n_nonzero, int log_m, int log_rows) {
    size_t rows = (size_t)1 << log_rows;
    cwd.nodes = DBuf<F>(c->pool, 2 * rows * 8);
    cwd.width = n_cols;
    cwd.nnz = n_nonzero;
    cwd.log_rows = log_rows;
    wl_ctx::PreEncoded &pre = c->pre0;
    if (pre.cw.p && pre.cols == cols && pre.log_rows == log_rows && pre.n_cols == n_cols && pre.nnz == n_nonzero) {
        // encoded while the witness uploaded
        cwd.cw = std::move(pre.cw);
        cwd.la = pre.la;
        pre = wl_ctx::PreEncoded();
    } else {
        cwd.cw = DBuf<F>(c->pool, (size_t)std::max(1, n_nonzero) * rows);  // zero columns are not stored
        DBuf<F> tmp(c->pool, (size_t)std::max(1, n_nonzero) * rows);
        // algorithmic HBM bytes: read the evaluations, write the codeword columns
        KTimer kt(c, KT_NTT, 4.0 * (double)n_nonzero * (double)(rows + ((size_t)1 << log_m)));
        cwd.la = wt::rs_encode(c, cols, n_cols, n_nonzero, log_m, log_rows, cwd.cw, tmp);
        HIP_CHECK(hipGetLastError());
    }
    {
        // one Poseidon2-24 permutation per 16 leaf elements (PaddingFreeSponge rate 16)
        KTimer kt(c, KT_MERKLE_LEAVES, (double)rows * ((n_cols + 15) / 16));
        wt::merkle_leaves(c, cwd.cw, n_cols, n_nonzero, log_rows, cwd.la, pdev, cwd.nodes);
    }
    {
        KTimer kt(c, KT_MERKLE_NODES, (double)(rows - 1));
        wt::merkle_nodes(c, log_rows, pdev, cwd.nodes);
    }
}

std::vector<uint32_t> sample_queries(Transcript &ts, int n, int log_rows) {
    std::vector<uint32_t> q(n);
    for (int i = 0; i < n; i++) q[i] = ts.sample_bits(log_rows);
    std::sort(q.begin(), q.end());
    q.erase(std::unique(q.begin(), q.end()), q.end());
    return q;
}

// Open rows + sibling paths of the sorted query indices as hints; when eqr is
// given, also return the rows folded by it (computed on the device).
std::vector<EF> open_queries(wl_ctx *c, Transcript &ts, const Codeword &cwd, const std::vector<uint32_t> &q,
                             const std::vector<EF> *eqr) {
    int nq = (int)q.size();
    std::vector<EF> vals;
    if (nq == 0) return vals;
    for (uint32_t x : q)
        if ((size_t)x >= ((size_t)1 << cwd.log_rows)) throw std::runtime_error("query index outside the codeword");
    size_t nrow = (size_t)nq * cwd.width, npath = (size_t)nq * cwd.log_rows * 8;
    size_t nval = eqr ? (size_t)nq * 8 : 0;
    // one device block [rows | paths | folded values] so that one fetch brings it all
    DBuf<F> buf(c->pool, nrow + npath + nval);
    Uploads up(c);
    const size_t o_idx = up.add(q), o_eq = eqr ? up.add(*eqr) : 0;
    up.send();
    wk::merkle_open(cwd.cw, cwd.nodes, cwd.width, cwd.nnz, cwd.log_rows, cwd.la, up.at<uint32_t>(o_idx), nq, buf.p,
                    buf.p + nrow, c->stream);
    if (eqr)
        wk::fold_rows(buf.p, nq, cwd.fold, cwd.base, up.at<EF>(o_eq), (EF *)(buf.p + nrow + npath), c->stream);
    // rows and paths arrive in canonical form (converted by the copy kernel), the
    // folded values in Montgomery form
    std::vector<uint32_t> host(nrow + npath + nval);
    c->fetch(host.data(), buf.p, host.size() * sizeof(uint32_t), nrow + npath);
    for (int i = 0; i < nq; i++) {
        ts.hint_canonical(&host[(size_t)i * cwd.width], cwd.width);
        ts.hint_canonical(&host[nrow + (size_t)i * cwd.log_rows * 8], (size_t)cwd.log_rows * 8);
    }
    if (eqr) {
        vals.resize(nq);
        memcpy(vals.data(), &host[nrow + npath], nval * sizeof(F));
    }
    return vals;
}

// A WHIR equality constraint on the current polynomial (m variables):
// scal * eq(point, x); point kinds: univariate F (expand(y, m)), univariate EF
// (expand(z, m)) or an explicit EF point.
struct Constraint {
    EF scal;
    int kind;          // 0 explicit, 1 univariate EF, 2 univariate F
    std::vector<EF> pt;
    EF z;
    F y;
};

std::vector<EF> cons_point(const Constraint &k, int m) {
    if (k.kind == 0) return k.pt;
    if (k.kind == 1) return wh::expand_univ(k.z, m);
    return wh::expand_univ(efrom(k.y), m);
}

// After binding the top `r.size()` variables to r: scal *= eq(prefix, r), point -> suffix.
void fold_constraint(Constraint &k, int m, const std::vector<EF> &r) {
    int kk = (int)r.size();
    if (k.kind == 2) {
        // prefix of expand(y, m): y^(2^(m-1)), ..., y^(2^(m-kk)), computed in F;
        // eq factor per coordinate: (1 - p) + r_t (2p - 1)
        std::vector<F> pw(m);
        F cur = k.y;
        for (int i = 0; i < m; i++) { pw[m - 1 - i] = cur; cur = fmul(cur, cur); }
        EF acc = eone();
        for (int t = 0; t < kk; t++) {
            F p = pw[t];
            EF f = emulf(r[t], fsub(fdbl(p), fone()));
            f.c[0] = fadd(f.c[0], fsub(fone(), p));
            acc = emul(acc, f);
        }
        k.scal = emul(k.scal, acc);
        return;  // expand(y, m)[kk..] = expand(y, m - kk): the parameter is unchanged
    }
    std::vector<EF> pt = cons_point(k, m);
    k.scal = emul(k.scal, wh::eq_outside(pt.data(), r.data(), kk));
    if (k.kind == 0) k.pt.erase(k.pt.begin(), k.pt.begin() + kk);
}

// W(x) += sum_k scal_k eq(point_k, x) over m variables (hi/lo tensor split).
void add_constraints_w(wl_ctx *c, EF *W, int m, const std::vector<Constraint> &cs, bool zero_first) {
    size_t Nw = (size_t)1 << m;
    if (cs.empty()) {
        if (zero_first) wk::fill_zero(W, Nw * sizeof(EF), c->stream);
        return;
    }
    int lo = m / 2, hi = m - lo;
    size_t LO = (size_t)1 << lo, HI = (size_t)1 << hi;
    std::vector<EF> sF, sE;
    std::vector<F> yl, yh;
    std::vector<const Constraint *> efc;
    for (const auto &k : cs) {
        if (k.kind == 2) {
            sF.push_back(k.scal);
            yl.push_back(k.y);
            F yy = k.y;
            for (int i = 0; i < lo; i++) yy = fmul(yy, yy);
            yh.push_back(yy);
        } else {
            sE.push_back(k.scal);
            efc.push_back(&k);
        }
    }
    int nF = (int)sF.size(), nE = (int)sE.size();
    DBuf<F> hiF(c->pool, std::max<size_t>(1, nF * HI)), loF(c->pool, std::max<size_t>(1, nF * LO));
    DBuf<EF> hiE(c->pool, std::max<size_t>(1, nE * HI)), loE(c->pool, std::max<size_t>(1, nE * LO));
    Uploads up(c);
    const size_t o_sF = up.add(sF), o_yl = up.add(yl), o_yh = up.add(yh);
    up.send();
    const EF *dsF = up.at<EF>(o_sF);
    if (nF) {
        wk::eq_univ_f_tables(up.at<F>(o_yh), nF, hi, hiF, c->stream);
        wk::eq_univ_f_tables(up.at<F>(o_yl), nF, lo, loF, c->stream);
    }
    std::vector<EqJob> eqs;
    for (int j = 0; j < nE; j++) {
        std::vector<EF> pt = cons_point(*efc[j], m);
        eqs.push_back(EqJob{std::vector<EF>(pt.begin(), pt.begin() + hi), sE[j], hiE.p + (size_t)j * HI});
        eqs.push_back(EqJob{std::vector<EF>(pt.begin() + hi, pt.end()), eone(), loE.p + (size_t)j * LO});
    }
    eq_batch(c, eqs);
    wk::whir_build_w(W, m, lo, nF, dsF, hiF, loF, nE, hiE, loE, !zero_first, c->stream);
}

struct SumcheckState {
    DBuf<EF> f, W;
    int m = 0;  // tables hold 2^m entries
    EF sum;
};

// kn GPU sumcheck rounds on (f, W), evaluations at 0, 1, 2 sent (whir-p3
// convention).  Every round after the first is fused with the fold by the
// previous challenge and, unless grinding sits between a round and its
// challenge, enqueued behind a device wait for that challenge before the host
// computes it.  The tables end folded by the last challenge.
std::vector<EF> whir_sc_rounds(wl_ctx *c, Transcript &ts, const P2Dev *pdev, SumcheckState &s, int kn,
                               int pow_bits) {
    std::vector<EF> rs;
    // grinding launches device work before the challenge exists; a context without
    // device waits enqueues each round once its challenge is known
    const bool pipe = pow_bits <= 0 && !c->host_rounds;
    auto enqueue = [&](int m, const EF *cdev, DBuf<EF> &parts, uint32_t &seq) {
        int nb = wk::dot_blocks((size_t)1 << (m - 1));
        parts = DBuf<EF>(c->pool, 2 * (size_t)nb);
        wk::prod_round_fused(s.f, s.W, m, cdev != nullptr, cdev, parts, c->stream);
        wl_ctx::HostSlot sl{};
        if (pipe) sl = reduce_async_await(c, parts, 2, nb, 8, seq);
        else seq = reduce_async(c, parts, 2, nb);
        return sl;
    };
    DBuf<EF> pcur, pnext;
    uint32_t seq = 0;
    wl_ctx::HostSlot slot = enqueue(s.m, nullptr, pcur, seq);
    for (int k = 0; k < kn; k++) {
        wl_ctx::HostSlot slot_next{};
        uint32_t seq_next = 0;
        if (pipe) {
            if (k + 1 < kn) slot_next = enqueue(s.m - 1, (const EF *)slot.dev, pnext, seq_next);
            else wk::fold_ef2_p(s.f, s.W, s.m, (const EF *)slot.dev, c->stream);
        }
        std::vector<EF> h02 = collect(c, seq, 2);
        EF h[3] = {h02[0], esub(s.sum, h02[0]), h02[1]};
        ts.add_ef(h, 3);
        grind(c, ts, pdev, pow_bits);
        EF r = ts.sample_ef();
        s.sum = wh::interp3(h, r);
        if (pipe) {
            c->release(slot, &r, sizeof(EF));
        } else {
            DBuf<EF> dr = upload_ef(c, std::vector<EF>(1, r));
            if (k + 1 < kn) slot_next = enqueue(s.m - 1, dr.p, pnext, seq_next);
            else wk::fold_ef2_p(s.f, s.W, s.m, dr.p, c->stream);
        }
        rs.push_back(r);
        seq = seq_next;
        slot = slot_next;
        std::swap(pcur, pnext);
        s.m--;
    }
    return rs;
}

// Host sumcheck round on small tables (final rounds).
EF host_sc_round(Transcript &ts, wl_ctx *c, const P2Dev *pdev, std::vector<EF> &f, std::vector<EF> &W,
                 EF &sum, int pow_bits) {
    size_t h = f.size() / 2;
    EF h0 = ezero(), h2 = ezero();
    for (size_t i = 0; i < h; i++) {
        h0 = eadd(h0, emul(f[i], W[i]));
        h2 = eadd(h2, emul(esub(edbl(f[i + h]), f[i]), esub(edbl(W[i + h]), W[i])));
    }
    EF hv[3] = {h0, esub(sum, h0), h2};
    ts.add_ef(hv, 3);
    grind(c, ts, pdev, pow_bits);
    EF r = ts.sample_ef();
    sum = wh::interp3(hv, r);
    for (size_t i = 0; i < h; i++) {
        f[i] = eadd(f[i], emul(r, esub(f[i + h], f[i])));
        W[i] = eadd(W[i], emul(r, esub(W[i + h], W[i])));
    }
    f.resize(h);
    W.resize(h);
    return r;
}

struct WhirWitness {
    const F *evals;   // device, 2^v
    size_t nz_len;    // evals[i] = 0 for i >= nz_len
    Codeword cw0;
    std::vector<EF> ood_points, ood_answers;
    std::vector<std::vector<EF>> ood_T;  // T tables of the OOD points (whir_T_tables)
};

// T_j(u) = sum_x f(u 2^(v-f0) + x) eq(suf_j, x) for u < 2^f0: the packed
// polynomial summed against the eq tables of the variables after the first f0,
// for several points in one pass (one eq launch, one pass over f, one reduction).
// f(pre_j, suf_j) is then sum_u eq(pre_j, u) T_j(u).
std::vector<std::vector<EF>> whir_T_tables(wl_ctx *c, const F *f, int v, int f0, size_t nz_len,
                                           const std::vector<std::vector<EF>> &sufs,
                                           const std::function<void()> &meanwhile = nullptr) {
    const int X = v - f0, U = 1 << f0, nC = (int)sufs.size();
    std::vector<std::vector<EF>> out(nC);
    if (nC == 0) return out;
    DBuf<EF> E(c->pool, (size_t)nC << X);
    std::vector<EqJob> eqs;
    for (int j = 0; j < nC; j++) eqs.push_back(EqJob{sufs[j], eone(), E.p + ((size_t)j << X)});
    eq_batch(c, eqs);
    const int nch = wk::whir_T_chunks(v, f0);
    DBuf<EF> parts(c->pool, (size_t)nC * U * nch);
    wk::whir_T_f(f, v, f0, E, nC, parts, nz_len, c->stream);
    std::vector<EF> T = reduce_many(c, parts, nC * U, nch, meanwhile);
    for (int j = 0; j < nC; j++) out[j].assign(T.begin() + (size_t)j * U, T.begin() + (size_t)(j + 1) * U);
    return out;
}

EF eval_ef_univ(wl_ctx *c, const EF *f, int m, const EF &z) {
    std::vector<EF> pt = wh::expand_univ(z, m);
    DBuf<EF> E = eq_dev(c, pt, eone());
    int nb = wk::dot_blocks((size_t)1 << m);
    DBuf<EF> parts(c->pool, nb);
    wk::seg_dot(f, E, (size_t)1 << m, 1, parts, c->stream);
    return reduce_one(c, parts, nb);
}

// CommitmentWriter::commit: RS-encode the 2^f0 columns of the packed
// polynomial, Merkle-commit, absorb the root, answer OOD samples.
// `absorb_first` (optional): elements the transcript observes before the root,
// absorbed while the device encodes and hashes.
void whir_commit(wl_ctx *c, const WhirCfg &cfg, const P2Dev *pdev, Transcript &ts, const F *evals,
                 size_t nonzero_len, WhirWitness &w, const std::vector<uint32_t> *absorb_first = nullptr) {
    int v = cfg.num_variables, f0 = cfg.f0;
    int log_m = v - f0;
    int n_cols = 1 << f0;
    size_t colen = (size_t)1 << log_m;
    int n_nonzero = (int)std::min<size_t>(n_cols, (nonzero_len + colen - 1) / colen);
    w.evals = evals;
    w.nz_len = nonzero_len;
    commit_codeword(c, pdev, w.cw0, evals, n_cols, n_nonzero, log_m, v + cfg.starting_log_inv_rate - f0);
    if (absorb_first)
        for (uint32_t x : *absorb_first) ts.observe(to_mont(x));
    w.cw0.base = true;
    w.cw0.fold = f0;
    F root[8];
    c->d2h(root, w.cw0.nodes.p + 8, 8);
    ts.add_base(root, 8);
    for (int i = 0; i < cfg.commit_ood_samples; i++) w.ood_points.push_back(ts.sample_ef());
    // the OOD answers f(expand(z_i)) through the T tables that the opening's first
    // f0 sumcheck rounds use again
    std::vector<std::vector<EF>> pres, sufs;
    for (const EF &z : w.ood_points) {
        std::vector<EF> pt = wh::expand_univ(z, v);
        pres.emplace_back(pt.begin(), pt.begin() + cfg.f0);
        sufs.emplace_back(pt.begin() + cfg.f0, pt.end());
    }
    std::vector<std::vector<EF>> eqp(pres.size());
    w.ood_T = whir_T_tables(c, evals, v, cfg.f0, w.nz_len, sufs, [&] {
        for (size_t i = 0; i < pres.size(); i++) eqp[i] = wh::eq_host(pres[i]);
    });
    for (size_t i = 0; i < w.ood_points.size(); i++) {
        const std::vector<EF> &a = eqp[i];
        EF ans = ezero();
        for (size_t u = 0; u < a.size(); u++) ans = eadd(ans, emul(a[u], w.ood_T[i][u]));
        w.ood_answers.push_back(ans);
    }
    ts.add_ef(w.ood_answers.data(), w.ood_answers.size());
}

// Prover::prove (crates/air/src/prove.rs:216-225)
void whir_prove(wl_ctx *c, const WhirCfg &cfg, const P2Dev *pdev, Transcript &ts, WhirWitness &w,
                const std::vector<EF> &point, const EF &value) {
    int v = cfg.num_variables, f0 = cfg.f0;
    std::vector<Constraint> cs;
    EF g = ts.sample_ef(), gp = eone(), sum = ezero();
    for (size_t i = 0; i < w.ood_points.size(); i++) {
        Constraint k{gp, 1, {}, w.ood_points[i], fzero()};
        cs.push_back(k);
        sum = eadd(sum, emul(gp, w.ood_answers[i]));
        gp = emul(gp, g);
    }
    {
        Constraint k{gp, 0, point, ezero(), fzero()};
        cs.push_back(k);
        sum = eadd(sum, emul(gp, value));
    }
    // initial f0 rounds: T_j(u) = sum_x f(u, x) eq(pt_j[f0..], x) on the device,
    // then the f0 round polynomials from the 2^f0-entry tables.
    // the OOD constraints come first in cs; their T tables were built at commit time
    // Constraint j adds scal_j sum_u eq(p_j, u) T_j(u) over the 2^f0 entries u.  With
    // E_j = eq(p_j[t+1..f0], .) and q = p_j[t], round t's values are a0 = (1 - q) d0,
    // a1 = q d1, a2 = (3q - 1)(2 d1 - d0) for d0, d1 = sum_i E_j(i) T_j(i), T_j(i + h);
    // folding by r multiplies the constraint by eq(q, r), halves T_j, and the next
    // E_j is the sum of E_j's halves.
    const int X = v - f0, nC = (int)cs.size(), nood = (int)w.ood_T.size();
    std::vector<std::vector<EF>> Q(nC), E(nC), sufs;
    std::vector<EF> fac(nC);
    for (int j = 0; j < nC; j++) {
        std::vector<EF> pt = cons_point(cs[j], v);
        Q[j].assign(pt.begin(), pt.begin() + f0);
        fac[j] = cs[j].scal;
        if (j >= nood) sufs.emplace_back(pt.begin() + f0, pt.end());
    }
    std::vector<std::vector<EF>> Tj = w.ood_T;
    auto eq_tables = [&] {
        for (int j = 0; j < nC; j++) E[j] = wh::eq_host(std::vector<EF>(Q[j].begin() + 1, Q[j].end()));
    };
    for (auto &t : whir_T_tables(c, w.evals, v, f0, w.nz_len, sufs, eq_tables)) Tj.push_back(std::move(t));
    if (sufs.empty()) eq_tables();
    std::vector<EF> rf;
    const EF one = eone(), three = wh::esmall(3);
    for (int t = 0; t < f0; t++) {
        size_t h = Tj[0].size() / 2;
        EF h0 = ezero(), h1 = ezero(), h2 = ezero();
        for (int j = 0; j < nC; j++) {
            EF d0 = ezero(), d1 = ezero();
            for (size_t i = 0; i < h; i++) {
                d0 = eadd(d0, emul(E[j][i], Tj[j][i]));
                d1 = eadd(d1, emul(E[j][i], Tj[j][i + h]));
            }
            const EF q = Q[j][t];
            h0 = eadd(h0, emul(fac[j], emul(esub(one, q), d0)));
            h1 = eadd(h1, emul(fac[j], emul(q, d1)));
            h2 = eadd(h2, emul(fac[j], emul(esub(emul(three, q), one), esub(edbl(d1), d0))));
        }
        EF hv[3] = {h0, h1, h2};
        ts.add_ef(hv, 3);
        grind(c, ts, pdev, cfg.starting_folding_pow_bits);
        EF r = ts.sample_ef();
        rf.push_back(r);
        sum = wh::interp3(hv, r);
        for (int j = 0; j < nC; j++) {
            const EF q = Q[j][t];
            fac[j] = emul(fac[j], eadd(emul(esub(one, q), esub(one, r)), emul(q, r)));
            for (size_t i = 0; i < h; i++) Tj[j][i] = eadd(Tj[j][i], emul(r, esub(Tj[j][i + h], Tj[j][i])));
            Tj[j].resize(h);
            const size_t hh = h / 2;
            for (size_t i = 0; i < hh; i++) E[j][i] = eadd(E[j][i], E[j][i + hh]);
            E[j].resize(hh ? hh : 1);
        }
    }
    for (auto &k : cs) fold_constraint(k, v, rf);
    SumcheckState s;
    s.m = X;
    s.sum = sum;
    s.f = DBuf<EF>(c->pool, (size_t)1 << X);
    {
        std::vector<EF> eqr = wh::eq_host(rf);
        DBuf<EF> deq = upload_ef(c, eqr);
        wk::whir_fold_f(w.evals, v, f0, deq, s.f, w.nz_len, c->stream);
    }
    s.W = DBuf<EF>(c->pool, (size_t)1 << X);
    add_constraints_w(c, s.W, X, cs, true);

    Codeword *prev = &w.cw0;
    Codeword own[2];
    int cur = 0;
    for (int r = 0; r < cfg.n_rounds; r++) {
        const wh::WhirRound &rp = cfg.rounds[r];
        int m = s.m;
        int rs = r == 0 ? cfg.rs_reduction0 : 1;
        int new_logd = rp.log_domain_size - rs;
        int kn = cfg.folding_at(r + 1);
        Codeword &nc = own[cur];
        {
            DBuf<F> cols(c->pool, (size_t)8 << m);
            wk::ef_to_cols(s.f, m, kn, cols, c->stream);
            commit_codeword(c, pdev, nc, cols, 8 << kn, 8 << kn, m - kn, new_logd - kn);
        }
        nc.base = false;
        nc.fold = kn;
        F root[8];
        c->d2h(root, nc.nodes.p + 8, 8);
        ts.add_base(root, 8);
        std::vector<EF> oodp, ooda;
        for (int i = 0; i < rp.ood_samples; i++) oodp.push_back(ts.sample_ef());
        for (int i = 0; i < rp.ood_samples; i++) ooda.push_back(eval_ef_univ(c, s.f, m, oodp[i]));
        ts.add_ef(ooda.data(), ooda.size());
        grind(c, ts, pdev, rp.pow_bits);
        int log_rows = rp.log_domain_size - prev->fold;
        std::vector<uint32_t> q = sample_queries(ts, rp.num_queries, log_rows);
        std::vector<EF> eqr = wh::eq_host(rf);
        std::vector<EF> qvals = open_queries(c, ts, *prev, q, &eqr);
        EF gg = ts.sample_ef(), gpp = eone();
        std::vector<Constraint> fresh;
        for (int i = 0; i < rp.ood_samples; i++) {
            fresh.push_back(Constraint{gpp, 1, {}, oodp[i], fzero()});
            s.sum = eadd(s.sum, emul(gpp, ooda[i]));
            gpp = emul(gpp, gg);
        }
        F gen = fpow(fsmall(3), (uint64_t)(P - 1) >> log_rows);
        for (size_t i = 0; i < q.size(); i++) {
            EF val = qvals[i];
            fresh.push_back(Constraint{gpp, 2, {}, ezero(), fpow(gen, q[i])});
            s.sum = eadd(s.sum, emul(gpp, val));
            gpp = emul(gpp, gg);
        }
        add_constraints_w(c, s.W, m, fresh, false);
        for (auto &k : fresh) cs.push_back(k);
        rf = whir_sc_rounds(c, ts, pdev, s, kn, rp.folding_pow_bits);
        for (auto &k : cs) fold_constraint(k, m, rf);
        prev = &nc;
        cur ^= 1;
        own[cur] = Codeword();  // release the codeword two rounds back
    }
    // final round
    size_t FN = (size_t)1 << s.m;
    std::vector<EF> ff(FN), Wf(FN);
    c->fetch2(ff.data(), s.f.p, FN * sizeof(EF), Wf.data(), s.W.p, FN * sizeof(EF));
    ts.add_ef(ff.data(), FN);
    grind(c, ts, pdev, cfg.final_pow_bits);
    {
        int log_rows = cfg.final_log_domain_size - prev->fold;
        std::vector<uint32_t> q = sample_queries(ts, cfg.final_queries, log_rows);
        open_queries(c, ts, *prev, q, nullptr);
    }
    for (int k = 0; k < cfg.final_sumcheck_rounds; k++) host_sc_round(ts, c, pdev, ff, Wf, s.sum, cfg.final_folding_pow_bits);
}

// --------------------------------------------------------------- AIR prove --
// the host transcript's duplex sponge <-> its device copy
void sponge_to_device(const Transcript &ts, wk::DevSponge &d) {
    for (int i = 0; i < 16; i++) d.st[i] = ts.state[i];
    for (int i = 0; i < 8; i++) { d.inb[i] = ts.in_buf[i]; d.outb[i] = ts.out_buf[i]; }
    d.n_in = (uint32_t)ts.n_in;
    d.n_out = (uint32_t)ts.n_out;
}
void sponge_from_device(const wk::DevSponge &d, Transcript &ts) {
    for (int i = 0; i < 16; i++) ts.state[i] = d.st[i];
    for (int i = 0; i < 8; i++) { ts.in_buf[i] = d.inb[i]; ts.out_buf[i] = d.outb[i]; }
    ts.n_in = (int)d.n_in;
    ts.n_out = (int)d.n_out;
}
// the Lagrange basis on the nodes 0..3 (coefficient rows) for the device rounds
wk::Basis4 basis4() {
    wk::Basis4 b;
    const std::vector<F> &B = wh::int_basis(4);
    for (int i = 0; i < 16; i++) b.b[i] = B[i];
    return b;
}

// AirTable::build_whir_params (crates/air/src/table.rs:72-98): the committed
// polynomial has log_length + log2_up(n_witness_columns) variables
WhirCfg air_whir_config(uint32_t log_length, uint32_t n_witness_columns, const wl_air_settings *st) {
    const int v = (int)log_length + (int)wh::log2_up(n_witness_columns);
    return wh::whir_config(v, (int)st->security_bits, 16, (int)st->folding_factor_first, (int)st->folding_factor,
                           (int)st->log_inv_rate, (int)st->rs_domain_initial_reduction, (int)st->soundness);
}

// trace: device, Montgomery, 165 columns x 2^n inside a zero-padded 2^(n+8) buffer.
// Appends to the transcript ts; absorb_first (optional): domain-separator words the
// transcript observes before the first message, absorbed while the first commit runs.
void air_prove(wl_table *t, const wl_air_settings *st, const P2Dev *pdev, const F *packed, Transcript &ts,
               const std::vector<uint32_t> *absorb_first) {
    wl_ctx *c = t->ctx;
    const int n = (int)t->log_length, skips = (int)t->univariate_skips, sec = (int)st->security_bits;
    const size_t N = (size_t)1 << n;
    const int lc = (int)wh::log2_up(AIR_NCOLS);
    WhirCfg cfg = air_whir_config(t->log_length, AIR_NCOLS, st);
    c->ring_reset();

    auto tp = std::chrono::steady_clock::now();
    // phase boundaries on the host's timeline (no stream synchronization: the
    // next phase's launches queue behind the current one's)
    auto lap = [&](int i) {
        auto now = std::chrono::steady_clock::now();
        c->phase_ms[i] += std::chrono::duration<double, std::milli>(now - tp).count();
        tp = now;
    };

    // tracing spans of crates/air/src/prove.rs:48,114,133 (whir-p3's commit and
    // prove under this build's names)
    Span span_prove("air: prove");
    PhaseSpan phase("whir: commit");
    WhirWitness w;
    whir_commit(c, cfg, pdev, ts, packed, (size_t)AIR_NCOLS * N, w, absorb_first);
    lap(0);
    phase.next("zerocheck");
    grind(c, ts, pdev, pow_bits_air(sec, (int)wh::log2_up(AIR_NCONS)));
    EF alpha = ts.sample_ef();
    std::vector<EF> ap(AIR_NCONS);
    ap[0] = eone();
    for (int i = 1; i < AIR_NCONS; i++) ap[i] = emul(ap[i - 1], alpha);
    DBuf<EF> dap = upload_ef(c, ap);
    grind(c, ts, pdev, pow_bits_air(sec, n));
    const int nz = n + 1 - skips;
    std::vector<EF> zc(nz);
    for (int i = 0; i < nz; i++) zc[i] = ts.sample_ef();

    // ---- zerocheck round 0 (univariate skip) ----
    const int K = 1 << skips;
    const size_t M = N >> skips;
    std::vector<EF> ch;
    std::vector<EF> zc_inv(nz), sc0;  // 1 / zc[j] for the later rounds; the selectors at round 0's challenge
    EF sum = ezero(), missing = eone();
    DBuf<EF> U(c->pool, (size_t)AIR_NCOLS * M);
    // eq(zc[1..], .) for round 0 and eq(zc[2..], .) for round 1 in one launch
    DBuf<EF> eq1(c->pool, (size_t)1 << (nz - 2));
    {
        DBuf<EF> eq(c->pool, (size_t)1 << (nz - 1));
        eq_batch(c, {EqJob{std::vector<EF>(zc.begin() + 1, zc.end()), eone(), eq.p},
                     EqJob{std::vector<EF>(zc.begin() + 2, zc.end()), eone(), eq1.p}});
        int zmax = 3 * (K - 1);
        int nzp = zmax - K + 1;  // z = K .. zmax
        size_t nbx = (M + 255) / 256;
        DBuf<EF> parts(c->pool, (size_t)nzp * nbx);
        {
            KTimer kt(c, KT_ZC_ROUND0, (double)nzp * M);
            // selector rows of the nzp points z = K .. zmax (fixed by skips)
            static thread_local std::vector<F> sel_cache[32];
            std::vector<F> &sel = sel_cache[skips];
            if (sel.empty())
                for (int z = K; z <= zmax; z++) {
                    std::vector<F> s = wh::selectors_f(skips, fsmall(z));
                    sel.insert(sel.end(), s.begin(), s.end());
                }
            DBuf<F> dsel(c->pool, sel.size());
            c->h2d_async(dsel.p, sel.data(), sel.size());
            if (skips <= 4) {
                // the rows extended to every z first, then the AIR per (x, z)
                DBuf<F> U0(c->pool, (size_t)nzp * AIR_NCOLS * M);
                wk::zc_round0_lde(packed, n, skips, nzp, dsel, t->air_dev, dap, eq, U0, parts, c->stream);
            } else {
                wk::zc_round0(packed, n, skips, dsel, nzp, t->air_dev, dap, eq, parts, c->stream);
            }
        }
        // host values the round-0 reply needs, computed while the device works:
        // the selectors at zc[0] and the inverses of the later round points
        std::vector<EF> sr = wh::selectors_ef(skips, zc[0]);
        if (skips <= 4) wh::skip_tables(skips);
        for (int j = 1; j < nz; j++) zc_inv[j] = wh::einv(zc[j]);
        std::vector<EF> vals = reduce_many(c, parts, nzp, (int)nbx);
        std::vector<EF> p;
        if (skips <= 4) {
            p = wh::skip_round_poly(skips, vals, sr);
        } else {
            std::vector<F> xs;
            std::vector<EF> ys;
            for (int i = 0; i < K; i++) { xs.push_back(fsmall(i)); ys.push_back(ezero()); }
            for (int i = 0; i < nzp; i++) { xs.push_back(fsmall(K + i)); ys.push_back(vals[i]); }
            std::vector<F> qx;
            for (int i = 0; i < K; i++) qx.push_back(fsmall(i));
            p = wh::pmul(wh::lagrange(xs, ys), wh::lagrange(qx, sr));
        }
        ts.add_ef(p.data(), p.size());
        EF c0 = ts.sample_ef();
        ch.push_back(c0);
        sum = wh::peval(p, c0);
        grind(c, ts, pdev, pow_bits_sumcheck(sec, (3 + 1) * (K - 1)));
        sc0 = wh::selectors_ef(skips, c0);
        EF mm = ezero();
        for (int i = 0; i < K; i++) mm = eadd(mm, emul(sr[i], sc0[i]));
        missing = mm;
        DBuf<EF> dS = upload_ef(c, sc0);
        wk::zc_fold0(packed, n, skips, dS, U, c->stream);
    }
    lap(1);
    // ---- zerocheck rounds 1 .. n - skips ----
    int l = n - skips;
    // eq(zc[1 + j ..], .) for round j; each fold launch halves it for the next round.
    // Pipelined: the fold by round j's challenge and round j + 1 are enqueued
    // behind a device wait for that challenge before the host computes it.
    DBuf<EF> eq = std::move(eq1);
    // round ll's kernels: the fold by the previous challenge (from 2^(ll+1), when
    // cdev is given), the evaluation, and the reduction + publish that then waits
    // for this round's challenge (the returned slot)
    auto enqueue_round = [&](int ll, const EF *cdev, DBuf<EF> &parts, uint32_t &seq) {
        if (cdev) wk::fold_cols_p(U, AIR_NCOLS, M, ll + 1, cdev, eq, c->stream);
        size_t nbx = (size_t)wk::zc_round_wave_blocks(ll);
        parts = DBuf<EF>(c->pool, 3 * nbx);
        {
            KTimer kt(c, KT_ZC_ROUND, 3.0 * (double)((size_t)1 << (ll - 1)));
            wk::zc_round_wave(U, M, ll, t->air_dev, t->chain_dev, dap, eq, parts, c->stream);
        }
        if (c->host_rounds) {
            seq = reduce_async(c, parts, 3, (int)nbx);
            return wl_ctx::HostSlot{};
        }
        return reduce_async_await(c, parts, 3, (int)nbx, 8, seq);
    };
    // Rounds without grinding can run back to back with the transcript on the
    // device: each round's reduction kernel absorbs the round polynomial into the
    // device sponge and samples the challenge the next fold reads; the host reads
    // the sponge, the proof words and the challenges back once.
    const bool device_rounds = c->device_transcript && nz > 1 && nz < wk::DTS_MAX_ROUNDS &&
                               pow_bits_sumcheck(sec, 3 + 1) == 0;
    if (device_rounds) {
        wk::ZcRounds zr;
        memset(&zr, 0, sizeof zr);
        sponge_to_device(ts, zr.sp);
        zr.sum = sum;
        zr.missing = missing;
        for (int j = 1; j < nz; j++) { zr.r[j] = zc[j]; zr.rinv[j] = zc_inv[j]; }
        DBuf<wk::ZcRounds> dz(c->pool, 1);
        c->h2d_async(dz.p, &zr, 1);
        const wk::Basis4 basis = basis4();
        int ll = l;
        for (int j = 1; j < nz; j++) {
            if (j > 1) wk::fold_cols_p(U, AIR_NCOLS, M, ll + 1, &dz.p->chal[j - 1], eq, c->stream);
            const int nbx = wk::zc_round_wave_blocks(ll);
            DBuf<EF> parts(c->pool, 3 * (size_t)nbx);
            {
                KTimer kt(c, KT_ZC_ROUND, 3.0 * (double)((size_t)1 << (ll - 1)));
                wk::zc_round_wave(U, M, ll, t->air_dev, t->chain_dev, dap, eq, parts, c->stream);
            }
            wk::zc_ts_round(parts, nbx, dz.p, j, basis, pdev, c->stream);
            ll--;
        }
        wk::fold_cols_p(U, AIR_NCOLS, M, ll + 1, &dz.p->chal[nz - 1], eq, c->stream);
        c->fetch(&zr, dz.p, sizeof zr);
        sponge_from_device(zr.sp, ts);
        ts.data.insert(ts.data.end(), zr.words, zr.words + 40 * (size_t)(nz - 1));
        for (int j = 1; j < nz; j++) ch.push_back(zr.chal[j]);
        sum = zr.sum;
        missing = zr.missing;
    }
    DBuf<EF> parts_cur, parts_next;
    uint32_t seq_cur = 0;
    wl_ctx::HostSlot slot_cur{};
    if (nz > 1 && !device_rounds) slot_cur = enqueue_round(l, nullptr, parts_cur, seq_cur);
    for (int j = 1; j < nz && !device_rounds; j++) {
        Span span_round("sumcheck_round");
        EF r = zc[j];
        uint32_t seq_next = 0;
        wl_ctx::HostSlot slot_next{};
        if (!c->host_rounds) {
            if (j + 1 < nz) slot_next = enqueue_round(l - 1, (const EF *)slot_cur.dev, parts_next, seq_next);
            else wk::fold_cols_p(U, AIR_NCOLS, M, l, (const EF *)slot_cur.dev, eq, c->stream);
        }
        std::vector<EF> v3 = collect(c, seq_cur, 3);
        EF e0 = emul(v3[0], missing), e2 = emul(v3[1], missing), e3 = emul(v3[2], missing);
        EF e1 = emul(esub(sum, emul(e0, esub(eone(), r))), zc_inv[j]);
        std::vector<EF> p = wh::interp_int({e0, e1, e2, e3});
        p = wh::pmul(p, {esub(eone(), r), esub(edbl(r), eone())});
        ts.add_ef(p.data(), p.size());
        EF cj = ts.sample_ef();
        if (c->host_rounds) {
            DBuf<EF> dc = upload_ef(c, std::vector<EF>(1, cj));
            if (j + 1 < nz) slot_next = enqueue_round(l - 1, dc.p, parts_next, seq_next);
            else wk::fold_cols_p(U, AIR_NCOLS, M, l, dc.p, eq, c->stream);
        } else {
            c->release(slot_cur, &cj, sizeof(EF));
        }
        ch.push_back(cj);
        sum = wh::peval(p, cj);
        grind(c, ts, pdev, pow_bits_sumcheck(sec, 3 + 1));
        missing = emul(eadd(emul(esub(eone(), r), esub(eone(), cj)), emul(r, cj)), missing);
        seq_cur = seq_next;
        slot_cur = slot_next;
        std::swap(parts_cur, parts_next);
        l--;
    }
    eq.reset();
    lap(2);
    phase.next("inner sumchecks");
    // ---- inner sums ----
    // up (column heads of the folded table) is published first and absorbed while
    // the device computes the down sums, published behind it.
    std::vector<EF> up(AIR_NCOLS), dn(AIR_NCOLS);
    c->ensure_mapped();
    const size_t DN_OFF = (size_t)AIR_NCOLS * 8;
    const uint32_t s_up = ++c->seq;
    wk::publish_rows((const uint32_t *)U.p, AIR_NCOLS, 8, M * 8, c->mapped_dev, c->flag_dev(), s_up, c->stream);
    U.reset();
    std::vector<EF> chr(ch.begin() + 1, ch.end());
    DBuf<EF> eqr = eq_dev(c, chr, eone());
    uint32_t s_dn;
    {
        DBuf<EF> parts(c->pool, (size_t)AIR_NCOLS * K);
        wk::zc_down_parts(packed, n, skips, eqr, parts, c->stream);
        DBuf<EF> dS = upload_ef(c, sc0);
        DBuf<EF> out(c->pool, AIR_NCOLS);
        wk::zc_down_combine(parts, skips, dS, out, c->stream);
        s_dn = ++c->seq;
        wk::publish((const uint32_t *)out.p, DN_OFF, c->mapped_dev + DN_OFF, c->flag_dev(), s_dn, c->publish_counter(),
                    c->stream);
    }
    c->wait_flag(s_up);
    memcpy(up.data(), c->mapped, AIR_NCOLS * sizeof(EF));
    ts.add_ef(up.data(), up.size());
    c->wait_flag(s_dn);
    memcpy(dn.data(), c->mapped + DN_OFF, AIR_NCOLS * sizeof(EF));
    ts.add_ef(dn.data(), dn.size());
    {
        Span span_pow("pow grinding");
        grind(c, ts, pdev, pow_bits_air(sec, lc));
    }
    std::vector<EF> cbs(lc);
    for (int i = 0; i < lc; i++) cbs[i] = ts.sample_ef();
    DBuf<EF> eqc = eq_dev(c, cbs, eone());
    DBuf<EF> B(c->pool, N);
    {
        Span span_lc("multilinears_linear_combination");
        wk::lincomb_cols(packed, n, AIR_NCOLS, eqc, B, c->stream);
    }
    EF alpha2 = ts.sample_ef();
    DBuf<EF> da2 = upload_ef(c, std::vector<EF>(1, alpha2));
    DBuf<EF> mixed(c->pool, N);
    wk::mix_updown(B, n, da2, mixed, c->stream);
    std::vector<EF> sub;
    {
        int nb = wk::dot_blocks(M);
        DBuf<EF> parts(c->pool, (size_t)K * nb);
        wk::seg_dot(mixed, eqr, M, K, parts, c->stream);
        sub = reduce_many(c, parts, K, nb);
    }
    ts.add_ef(sub.data(), sub.size());
    std::vector<EF> pt;
    for (int i = 0; i < skips; i++) pt.push_back(ts.sample_ef());
    for (size_t i = 1; i < ch.size(); i++) pt.push_back(ch[i]);
    DBuf<EF> E = eq_dev(c, pt, eone());
    DBuf<EF> Mt(c->pool, N);
    {
        Span span_m("matrix_up_folded + matrix_down_folded");
        wk::build_matrix(E, n, da2, Mt, c->stream);
    }
    // the inner sum is first needed with the first inner round's values: publish
    // it aside and read it then
    const size_t ISUM_OFF = wl_ctx::ISUM_OFF;
    uint32_t s_isum;
    {
        Span span_is("inner sum evaluation");
        int nb = wk::dot_blocks(N);
        DBuf<EF> parts(c->pool, nb);
        wk::seg_dot(mixed, E, N, 1, parts, c->stream);
        s_isum = reduce_async(c, parts, 1, nb, ISUM_OFF);
    }
    mixed.reset();
    E.reset();
    lap(3);
    // ---- inner sumcheck (degree 2, coefficients sent) ----
    std::vector<EF> ich;
    EF isum;
    bool have_isum = false;
    auto read_isum = [&]() {
        if (have_isum) return;
        c->wait_flag(s_isum);
        memcpy(&isum, c->mapped + ISUM_OFF, sizeof(EF));
        have_isum = true;
    };
    // pipelined as the zerocheck: round j + 1 (fused with the fold by c_j) waits
    // on the device for c_j
    auto enqueue_inner = [&](int ll, const EF *cdev, DBuf<EF> &parts, uint32_t &seq) {
        int nb = wk::dot_blocks((size_t)1 << (ll - 1));
        parts = DBuf<EF>(c->pool, 2 * (size_t)nb);
        wk::prod_round_fused(Mt, B, ll, cdev != nullptr, cdev, parts, c->stream);
        if (c->host_rounds) {
            seq = reduce_async(c, parts, 2, nb);
            return wl_ctx::HostSlot{};
        }
        return reduce_async_await(c, parts, 2, nb, 8, seq);
    };
    // the last rounds run on the host: from tables of 2^HOST_TAIL entries a round
    // costs the host less than one device round trip
    const int HOST_TAIL = 9;
    const int gpu_rounds = n > HOST_TAIL ? n - HOST_TAIL : 0;
    auto inner_round = [&](const EF &h0, const EF &h2) {
        read_isum();
        std::vector<EF> p = wh::interp_int({h0, esub(isum, h0), h2});
        ts.add_ef(p.data(), p.size());
        EF cj = ts.sample_ef();
        ich.push_back(cj);
        isum = wh::peval(p, cj);
        return std::make_pair(cj, p);
    };
    DBuf<EF> iparts_cur, iparts_next;
    uint32_t iseq = 0;
    wl_ctx::HostSlot islot{};
    if (gpu_rounds > 0) islot = enqueue_inner(n, nullptr, iparts_cur, iseq);
    for (int j = 0; j < gpu_rounds; j++) {
        Span span_round("sumcheck_round");
        int ll = n - j;
        uint32_t iseq_next = 0;
        wl_ctx::HostSlot islot_next{};
        if (!c->host_rounds) {
            if (j + 1 < gpu_rounds) islot_next = enqueue_inner(ll - 1, (const EF *)islot.dev, iparts_next, iseq_next);
            else wk::fold_ef2_p(Mt, B, ll, (const EF *)islot.dev, c->stream);
        }
        std::vector<EF> h02 = collect(c, iseq, 2);
        EF cj = inner_round(h02[0], h02[1]).first;
        if (c->host_rounds) {
            DBuf<EF> dc = upload_ef(c, std::vector<EF>(1, cj));
            if (j + 1 < gpu_rounds) islot_next = enqueue_inner(ll - 1, dc.p, iparts_next, iseq_next);
            else wk::fold_ef2_p(Mt, B, ll, dc.p, c->stream);
        } else {
            c->release(islot, &cj, sizeof(EF));
        }
        grind(c, ts, pdev, pow_bits_sumcheck(sec, 2));
        iseq = iseq_next;
        islot = islot_next;
        std::swap(iparts_cur, iparts_next);
    }
    const size_t tail = (size_t)1 << (n - gpu_rounds);
    std::vector<EF> hM(tail), hB(tail);
    c->fetch2(hM.data(), Mt.p, tail * sizeof(EF), hB.data(), B.p, tail * sizeof(EF));
    for (size_t len = tail; len > 1; len >>= 1) {
        Span span_round("sumcheck_round");
        size_t h = len / 2;
        EF h0 = ezero(), h2 = ezero();
        for (size_t i = 0; i < h; i++) {
            h0 = eadd(h0, emul(hM[i], hB[i]));
            h2 = eadd(h2, emul(esub(edbl(hM[i + h]), hM[i]), esub(edbl(hB[i + h]), hB[i])));
        }
        EF cj = inner_round(h0, h2).first;
        grind(c, ts, pdev, pow_bits_sumcheck(sec, 2));
        for (size_t i = 0; i < h; i++) {
            hM[i] = eadd(hM[i], emul(cj, esub(hM[i + h], hM[i])));
            hB[i] = eadd(hB[i], emul(cj, esub(hB[i + h], hB[i])));
        }
    }
    EF packed_value = hB[0];
    Mt.reset();
    B.reset();
    lap(4);
    std::vector<EF> fp(cbs);
    fp.insert(fp.end(), ich.begin(), ich.end());
    phase.next("whir: prove");
    whir_prove(c, cfg, pdev, ts, w, fp, packed_value);
    lap(5);
}

// AirTable::prove (crates/air/src/prove.rs:49-226) for an AIR given as a
// constraint program: the same protocol with the zerocheck run by the generic
// sumcheck over the 2C up/down columns of preprocessed ++ witness
// (prove.rs:109-131), inner sums of the witness columns only (:135-145), column
// batching over the witness (:147-160).  packed: device, Montgomery, the nw
// witness columns zero-padded to 2^(n + log2_up(nw)).
void air_prove_generic(wl_table *t, const wl_air_settings *st, const P2Dev *pdev, const F *packed, Transcript &ts,
                       const std::vector<uint32_t> *absorb_first) {
    wl_ctx *c = t->ctx;
    const wg::AirSpec &sp = *t->gen;
    const int n = (int)t->log_length, skips = (int)t->univariate_skips, sec = (int)st->security_bits;
    const size_t N = (size_t)1 << n;
    const int C = sp.n_columns, np = sp.n_preprocessed, nw = C - np;
    const int lc = (int)wh::log2_up((size_t)nw), ncons = sp.prog.n_asserts;
    WhirCfg cfg = air_whir_config(t->log_length, (uint32_t)nw, st);
    c->ring_reset();
    Span span_prove("air: prove");
    PhaseSpan phase("whir: commit");
    WhirWitness w;
    whir_commit(c, cfg, pdev, ts, packed, (size_t)nw * N, w, absorb_first);
    phase.next("zerocheck");
    grind(c, ts, pdev, pow_bits_air(sec, (int)wh::log2_up((size_t)std::max(1, ncons))));
    const EF alpha = ts.sample_ef();
    std::vector<EF> ap(std::max(1, ncons));
    ap[0] = eone();
    for (int i = 1; i < ncons; i++) ap[i] = emul(ap[i - 1], alpha);
    grind(c, ts, pdev, pow_bits_air(sec, n));
    std::vector<EF> zc(n + 1 - skips);
    for (auto &z : zc) z = ts.sample_ef();
    // ---- zerocheck ----
    std::vector<EF> ch, fe(2 * (size_t)C);
    {
        DBuf<F> ud(c->pool, 2 * (size_t)C * N);
        wg::columns_up_down(sp.pre.p, np, packed, C, n, ud, c->stream);
        wg::Args a;
        a.skips = skips;
        a.degree = sp.degree;
        a.batching = &ap;
        a.eq_factor = &zc;
        a.is_zerofier = true;
        a.sum = ezero();
        a.grind_kind = wg::GRIND_AUTO;
        a.grind_value = sec;
        DBuf<EF> folded;
        EF fsum;
        ch = wg::prove(c, ts, pdev, ud.p, true, 2 * C, n, sp.prog, a, folded, fsum);
        c->fetch(fe.data(), folded.p, fe.size() * sizeof(EF));
    }
    // ---- inner sums of the witness columns, up then down ----
    phase.next("inner sumchecks");
    ts.add_ef(fe.data() + np, (size_t)nw);
    ts.add_ef(fe.data() + C + np, (size_t)nw);
    {
        Span span_pow("pow grinding");
        grind(c, ts, pdev, pow_bits_air(sec, lc));
    }
    std::vector<EF> cbs(lc);
    for (auto &x : cbs) x = ts.sample_ef();
    // inner-sumcheck tables [matrix | batched column], contiguous
    DBuf<EF> tabs(c->pool, 2 * N);
    EF *Mt = tabs.p, *B = tabs.p + N;
    {
        DBuf<EF> eqc = lc ? eq_dev(c, cbs, eone()) : upload_ef(c, std::vector<EF>(1, eone()));
        Span span_lc("multilinears_linear_combination");
        wk::lincomb_cols(packed, n, nw, eqc, B, c->stream);
    }
    const EF alpha2 = ts.sample_ef();
    DBuf<EF> da2 = upload_ef(c, std::vector<EF>(1, alpha2));
    DBuf<EF> mixed(c->pool, N);
    wk::mix_updown(B, n, da2, mixed, c->stream);
    const int K = 1 << skips;
    const size_t M = N >> skips;
    std::vector<EF> sub;
    {
        std::vector<EF> chr(ch.begin() + 1, ch.end());
        DBuf<EF> eqr = chr.empty() ? upload_ef(c, std::vector<EF>(1, eone())) : eq_dev(c, chr, eone());
        const int nb = wk::dot_blocks(M);
        DBuf<EF> parts(c->pool, (size_t)K * nb);
        wk::seg_dot(mixed, eqr, M, K, parts, c->stream);
        sub = reduce_many(c, parts, K, nb);
    }
    ts.add_ef(sub.data(), sub.size());
    std::vector<EF> pt;
    for (int i = 0; i < skips; i++) pt.push_back(ts.sample_ef());
    for (size_t i = 1; i < ch.size(); i++) pt.push_back(ch[i]);
    EF inner_sum;
    {
        DBuf<EF> E = eq_dev(c, pt, eone());
        {
            Span span_m("matrix_up_folded + matrix_down_folded");
            wk::build_matrix(E, n, da2, Mt, c->stream);
        }
        Span span_is("inner sum evaluation");
        const int nb = wk::dot_blocks(N);
        DBuf<EF> parts(c->pool, nb);
        wk::seg_dot(mixed, E, N, 1, parts, c->stream);
        inner_sum = reduce_one(c, parts, nb);
    }
    mixed.reset();
    // ---- inner sumcheck: InnerSumcheckCircuit = point[0] * point[1] (prove.rs:193-208) ----
    wg::Program prod;
    {
        const uint32_t code[8] = {wg::OP_MUL, 0, 1, 0, wg::OP_ASSERT, 2, 0, 0};
        wg::compile(code, 2, 2, prod);
        wg::upload(c, prod);
    }
    const std::vector<EF> one(1, eone());
    wg::Args ia;
    ia.skips = 1;
    ia.degree = 2;
    ia.batching = &one;
    ia.sum = inner_sum;
    ia.grind_kind = wg::GRIND_AUTO;
    ia.grind_value = sec;
    DBuf<EF> ifolded;
    EF isum;
    std::vector<EF> ich = wg::prove(c, ts, pdev, tabs.p, false, 2, n, prod, ia, ifolded, isum);
    EF fin[2];
    c->fetch(fin, ifolded.p, sizeof fin);
    tabs.reset();
    std::vector<EF> fp(cbs);
    fp.insert(fp.end(), ich.begin(), ich.end());
    phase.next("whir: prove");
    whir_prove(c, cfg, pdev, ts, w, fp, fin[1]);
}

}  // namespace

// Coefficients of the partial-round chain as linear maps of (init, post_sbox):
// follow the state symbolically through the 20 rounds of
// state[0] := post_sbox_r; state := (J + diag) state.
static void chain_linear_maps(wk::ChainLin &cl) {
    const int NV = 36;  // 16 entering columns + 20 post_sbox columns
    std::vector<std::vector<F>> st(16, std::vector<F>(NV, fzero()));
    for (int i = 0; i < 16; i++) st[i][i] = fone();
    for (int r = 0; r < 20; r++) {
        for (int j = 0; j < 16; j++) cl.A[r][j] = st[0][j];
        for (int q = 0; q < 20; q++) cl.B[r][q] = st[0][16 + q];
        std::fill(st[0].begin(), st[0].end(), fzero());
        st[0][16 + r] = fone();
        std::vector<F> sum(NV, fzero());
        for (int i = 0; i < 16; i++)
            for (int k = 0; k < NV; k++) sum[k] = fadd(sum[k], st[i][k]);
        for (int i = 0; i < 16; i++)
            for (int k = 0; k < NV; k++) st[i][k] = fadd(sum[k], fmul(F{DIAG.d16[i]}, st[i][k]));
    }
    for (int i = 0; i < 16; i++) {
        for (int j = 0; j < 16; j++) cl.Ei[i][j] = st[i][j];
        for (int q = 0; q < 20; q++) cl.Ep[i][q] = st[i][16 + q];
    }
}

// ===================================================================== C ABI ==
extern "C" {

static thread_local std::string g_err;

const char *wl_last_error(void) { return g_err.c_str(); }

int wl_air_settings_validate(uint32_t log_length, uint32_t n_witness_columns, const wl_air_settings *st) {
    if (!st) { g_err = "null argument"; return WL_ERR_ARG; }
    if (st->folding_factor_first < 1 || st->folding_factor < 1 || log_length < 2 || log_length > 30 ||
        n_witness_columns < 1 || st->univariate_skips < 1 || st->univariate_skips >= log_length ||
        st->univariate_skips > 16 || st->log_inv_rate < 1 || st->log_inv_rate > 16 ||
        st->security_bits < 1 || st->security_bits > 512) {
        g_err = "unsupported settings";
        return WL_ERR_ARG;
    }
    try {
        const int v = (int)log_length + (int)wh::log2_up(n_witness_columns);
        (void)wh::whir_config(v, (int)st->security_bits, 16, (int)st->folding_factor_first, (int)st->folding_factor,
                              (int)st->log_inv_rate, (int)st->rs_domain_initial_reduction, (int)st->soundness);
    } catch (const std::exception &e) {
        g_err = std::string("unsupported settings: ") + e.what();
        return WL_ERR_ARG;
    }
    return WL_OK;
}

wl_ctx *wl_create(int device) {
    try {
        HIP_CHECK(hipSetDevice(device));
        wl_ctx *c = new wl_ctx();
        c->device = device;
        // WHIRLAWAY_DEVICE_ROUNDS=1 moves the zerocheck rounds' transcript to the device
        if (const char *e = getenv("WHIRLAWAY_DEVICE_ROUNDS")) c->device_transcript = e[0] == '1';
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&c->ev0));
        HIP_CHECK(hipEventCreate(&c->ev1));
        (void)c->sync_words();  // synchronizes once, before any device wait can be pending
        return c;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void wl_destroy(wl_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->release_all();
    (void)hipStreamSynchronize(c->stream);
    // the frees below wait for the whole device: not while another thread's proof
    // has device-side waits pending
    free_guard().enter_free();
    struct Leave {
        ~Leave() { free_guard().leave_free(); }
    } leave;
    for (void *h : c->retired_host) (void)hipHostFree(h);
    for (auto &kv : c->tw) (void)hipFree(kv.second);
    for (auto &kv : c->ntt_tab) (void)hipFree(kv.second);
    c->pool.release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stage_slots) {
        (void)hipHostFree(c->stage_slots);
        for (auto &e : c->stage_ev) (void)hipEventDestroy(e);
    }
    if (c->ring) (void)hipHostFree(c->ring);
    if (c->mapped) (void)hipHostFree(c->mapped);
    if (c->mbox) (void)hipHostFree(c->mbox);
    if (c->args_dev) (void)hipFree(c->args_dev);
    if (c->sync_dev) (void)hipFree(c->sync_dev);
    if (c->p2_dev) (void)hipFree(c->p2_dev);
    for (auto &e : c->kt_free) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto &v : c->kt_pairs)
        for (auto &e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->copy_order) (void)hipEventDestroy(c->copy_order);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int wl_set_device_waits(wl_ctx *c, int enabled) {
    if (!c) { g_err = "null context"; return WL_ERR_ARG; }
    c->host_rounds = enabled == 0;
    return WL_OK;
}

int wl_set_wait_timeout(wl_ctx *c, double seconds) {
    if (!c) { g_err = "null context"; return WL_ERR_ARG; }
    if (!(seconds >= 0.0 && seconds <= 3600.0)) { g_err = "timeout must be in [0, 3600] seconds"; return WL_ERR_ARG; }
    c->set_await_limit(seconds == 0.0 ? 10.0 : seconds);
    return WL_OK;
}

int wl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

wl_table *wl_air_table_new(wl_ctx *c, const uint32_t *rc, uint32_t log_length, uint32_t univariate_skips) {
    try {
        if (!rc) throw std::runtime_error("null argument");
        if (univariate_skips >= log_length) throw std::runtime_error("univariate_skips must be < log_length");
        wl_table *t = new wl_table();
        t->ctx = c;
        t->log_length = log_length;
        t->univariate_skips = univariate_skips;
        for (int i = 0; i < wk::AIR_NRC; i++) t->air_host.rc[i] = to_mont(rc[i] % P);
        if (!c) return t;  // verification only
        HIP_CHECK(hipSetDevice(c->device));
        chain_linear_maps(t->chain_host);
        // pool memory and stream-ordered copies: no device-wide synchronization, so
        // contexts on other host threads keep running
        t->air_dev = (wk::AirConsts *)c->pool.get(sizeof(wk::AirConsts));
        t->chain_dev = (wk::ChainLin *)c->pool.get(sizeof(wk::ChainLin));
        HIP_CHECK(hipMemcpyAsync(t->air_dev, &t->air_host, sizeof(wk::AirConsts), hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(t->chain_dev, &t->chain_host, sizeof(wk::ChainLin), hipMemcpyHostToDevice,
                                 c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return t;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void wl_air_table_free(wl_table *t) {
    if (!t) return;
    if (t->ctx) {
        (void)hipStreamSynchronize(t->ctx->stream);
        t->ctx->pool.put(t->air_dev);
        t->ctx->pool.put(t->chain_dev);
    }
    delete t->gen;
    delete t;
}

wl_table *wl_air_table_new_program(wl_ctx *c, uint32_t n_columns, uint32_t n_preprocessed,
                                   const uint32_t *preprocessed, uint32_t log_length, uint32_t univariate_skips,
                                   const uint32_t *program, uint32_t n_instr, uint32_t constraint_degree) {
    wl_table *t = nullptr;
    try {
        if (!program || (n_preprocessed && !preprocessed)) throw std::runtime_error("null argument");
        if (n_columns < 1 || n_preprocessed >= n_columns || n_columns > 32767)
            throw std::runtime_error("an AIR needs at least one witness column and at most 32767 columns");
        if (log_length < 2 || log_length > 28 || univariate_skips < 1 || univariate_skips >= log_length)
            throw std::runtime_error("univariate_skips must be in [1, log_length)");
        if (constraint_degree < 1) throw std::runtime_error("constraint_degree must be at least 1");
        t = new wl_table();
        t->ctx = c;
        t->log_length = log_length;
        t->univariate_skips = univariate_skips;
        t->gen = new wg::AirSpec();
        wg::AirSpec &sp = *t->gen;
        sp.n_columns = (int)n_columns;
        sp.n_preprocessed = (int)n_preprocessed;
        sp.degree = (int)constraint_degree;
        wg::compile(program, n_instr, 2 * (int)n_columns, sp.prog);
        sp.code.assign(program, program + 4 * (size_t)n_instr);
        const size_t N = (size_t)1 << log_length;
        if (n_preprocessed) sp.pre_host.assign(preprocessed, preprocessed + (size_t)n_preprocessed * N);
        for (uint32_t &w : sp.pre_host)
            if (w >= P) throw std::runtime_error("preprocessed values must be canonical");
        if (!c) return t;  // verification only
        HIP_CHECK(hipSetDevice(c->device));
        wg::upload(c, sp.prog);
        sp.pre = DBuf<F>(c->pool, std::max<size_t>(1, (size_t)n_preprocessed * N));
        if (n_preprocessed) {
            HIP_CHECK(hipMemcpyAsync(sp.pre.p, preprocessed, (size_t)n_preprocessed * N * 4, hipMemcpyHostToDevice,
                                     c->stream));
            wk::convert_to_mont((uint32_t *)sp.pre.p, (size_t)n_preprocessed * N, c->stream);
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return t;
    } catch (const std::exception &e) {
        g_err = e.what();
        if (t) { delete t->gen; delete t; }
        return nullptr;
    }
}

// committed (witness) columns of a table
static uint32_t witness_cols(const wl_table *t) {
    return t->gen ? (uint32_t)(t->gen->n_columns - t->gen->n_preprocessed) : (uint32_t)AIR_NCOLS;
}

// Proves into `ts`: a caller's transcript (wl_transcript), or, when null, a fresh
// one over p16 that first observes this build's domain separator.
static int prove_common(wl_table *t, const wl_air_settings *st, const uint32_t *p16c,
                        const uint32_t *p24c, F *packed, uint32_t *proof_out, size_t proof_cap,
                        size_t *proof_len, wl_transcript *caller_ts = nullptr) {
    wl_ctx *c = t->ctx;
    P2Dev ph;
    load_p2(p16c, p24c, ph);
    if (caller_ts && memcmp(&caller_ts->p16, &ph.p16, sizeof ph.p16) != 0) {
        g_err = "the transcript's Poseidon2-16 differs from the Merkle compression's";
        return WL_ERR_ARG;
    }
    const P2Dev *pd = context_p2(c, ph);
    std::vector<uint32_t> proof;
    struct Guard {  // a proof with device-side waits keeps frees elsewhere out (FreeGuard)
        bool on;
        explicit Guard(bool waits) : on(waits) { if (on) free_guard().enter_proof(); }
        ~Guard() { if (on) free_guard().leave_proof(); }
    } guard(!c->host_rounds);
    auto run = [&](Transcript &ts, const std::vector<uint32_t> *domsep) {
        if (t->gen) air_prove_generic(t, st, pd, packed, ts, domsep);
        else air_prove(t, st, pd, packed, ts, domsep);
    };
    try {
        if (caller_ts) {
            run(caller_ts->ts, nullptr);
        } else {
            Transcript ts(&ph.p16);
            const std::vector<uint32_t> domsep = wh::whir_domsep(air_whir_config(t->log_length, witness_cols(t), st));
            run(ts, &domsep);
            proof = std::move(ts.data);
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));  // nothing of this proof left in flight
    } catch (...) {
        c->release_all();  // device waits for host values must not outlive the failure
        throw;
    }
    if (caller_ts) {
        if (proof_len) *proof_len = caller_ts->ts.data.size();
        return WL_OK;
    }
    if (proof_len) *proof_len = proof.size();
    if (proof_out) {
        if (proof_cap < proof.size()) { g_err = "proof buffer too small"; return WL_ERR_BUFFER; }
        memcpy(proof_out, proof.data(), proof.size() * 4);
    }
    return WL_OK;
}

static int check_settings(const wl_table *t, const wl_air_settings *st) {
    if (!t || !st) { g_err = "null argument"; return WL_ERR_ARG; }
    if (!t->ctx) { g_err = "a table created without a context only verifies"; return WL_ERR_ARG; }
    if (st->univariate_skips != t->univariate_skips) {
        g_err = "settings.univariate_skips differs from the table's";
        return WL_ERR_ARG;
    }
    return wl_air_settings_validate(t->log_length, witness_cols(t), st);
}

// packed_multilinear (crates/utils/src/multilinear.rs:169-181) of a host witness of
// ncols canonical columns, in Montgomery form, zero-padded to 2^(n + log2_up(ncols))
static size_t packed_len(uint32_t log_length, uint32_t ncols) {
    return (size_t)1 << (log_length + wh::log2_up(ncols));
}
static void upload_packed(wl_ctx *c, const uint32_t *witness, uint32_t log_length, uint32_t ncols, F *packed) {
    const size_t N = (size_t)1 << log_length, V = packed_len(log_length, ncols);
    wk::fill_zero(packed + ncols * N, (V - ncols * N) * 4, c->stream);
    c->h2d_staged(packed, witness, ncols * N * 4);
    wk::convert_to_mont((uint32_t *)packed, ncols * N, c->stream);
}

static int prove_host_witness(wl_table *t, const wl_air_settings *st, const uint32_t *p16c, const uint32_t *p24c,
                              const uint32_t *witness, uint32_t *proof_out, size_t proof_cap, size_t *proof_len,
                              wl_transcript *ts) {
    int rc = check_settings(t, st);
    if (rc) return rc;
    if (!witness) { g_err = "null witness"; return WL_ERR_ARG; }
    try {
        wl_ctx *c = t->ctx;
        HIP_CHECK(hipSetDevice(c->device));
        const uint32_t nw = witness_cols(t);
        const size_t N = (size_t)1 << t->log_length, V = packed_len(t->log_length, nw);
        DBuf<F> packed(c->pool, V);
        // The first commitment's Reed-Solomon encoding (WhirConfig of the table, as in
        // whir_commit) runs column group by column group while the rest of the
        // witness is still on its way; commit_codeword then takes the codeword.
        const WhirCfg cfg = air_whir_config(t->log_length, nw, st);
        const int f0 = cfg.f0, log_m = cfg.num_variables - f0;
        const int log_rows = cfg.num_variables + cfg.starting_log_inv_rate - f0;
        const size_t colen = (size_t)1 << log_m, rows = (size_t)1 << log_rows;
        const int n_cols = 1 << f0;
        const int nnz = (int)std::min<size_t>(n_cols, (nw * N + colen - 1) / colen);
        DBuf<F> pcw(c->pool, (size_t)nnz * rows), tmp(c->pool, (size_t)nnz * rows);
        F *tw = c->twiddle(log_rows);
        // the split the tuner chose for this length (the encoding here is not timed)
        const int nv = wt::ntt_variant(c, log_rows);
        const F *tabs = c->ntt_tabs(log_rows, nv);
        wk::fill_zero(packed + nw * N, (V - nw * N) * 4, c->stream);
        int la = 0;
        c->h2d_staged(packed, witness, nw * N * 4, colen * 4, [&](size_t u0, size_t u1) {
            const size_t e0 = u0 * colen, e1 = std::min(u1 * colen, nw * N);
            wk::convert_to_mont((uint32_t *)packed.p + e0, e1 - e0, c->stream);
            la = wk::rs_encode(packed.p + e0, (int)(u1 - u0), (int)(u1 - u0), log_m, log_rows, pcw.p + u0 * rows,
                               tmp.p + u0 * rows, tw, tabs, c->stream, false, nv);
        });
        tmp.reset();
        c->pre0.cols = packed.p;
        c->pre0.log_rows = log_rows;
        c->pre0.n_cols = n_cols;
        c->pre0.nnz = nnz;
        c->pre0.la = la;
        c->pre0.cw = std::move(pcw);
        struct Unclaimed {  // an encoding the proof did not take goes back to the pool
            wl_ctx *c;
            ~Unclaimed() { c->pre0 = wl_ctx::PreEncoded(); }
        } unclaimed{c};
        return prove_common(t, st, p16c, p24c, packed, proof_out, proof_cap, proof_len, ts);
    } catch (const std::exception &e) {
        g_err = e.what();
        return WL_ERR_DEVICE;
    }
}

int wl_air_prove(wl_table *t, const wl_air_settings *st, const uint32_t *p16c, const uint32_t *p24c,
                 const uint32_t *witness, uint32_t *proof_out, size_t proof_cap, size_t *proof_len) {
    return prove_host_witness(t, st, p16c, p24c, witness, proof_out, proof_cap, proof_len, nullptr);
}

int wl_air_prove_into(wl_table *t, const wl_air_settings *st, const uint32_t *p16c, const uint32_t *p24c,
                      const uint32_t *witness, wl_transcript *ts) {
    if (!ts || ts->verifier) { g_err = "a prover transcript is required"; return WL_ERR_ARG; }
    return prove_host_witness(t, st, p16c, p24c, witness, nullptr, 0, nullptr, ts);
}

int wl_air_prove_resident_into(wl_table *t, const wl_air_settings *st, const uint32_t *p16c,
                               const uint32_t *p24c, const wl_witness *w, wl_transcript *ts) {
    if (!ts || ts->verifier) { g_err = "a prover transcript is required"; return WL_ERR_ARG; }
    int rc = check_settings(t, st);
    if (rc) return rc;
    if (!w || w->log_length != t->log_length || t->gen) { g_err = "witness size mismatch"; return WL_ERR_ARG; }
    try {
        HIP_CHECK(hipSetDevice(t->ctx->device));
        return prove_common(t, st, p16c, p24c, w->packed, nullptr, 0, nullptr, ts);
    } catch (const std::exception &e) {
        g_err = e.what();
        return WL_ERR_DEVICE;
    }
}

int wl_air_domain_separator(uint32_t log_length, uint32_t n_witness_columns, const wl_air_settings *st,
                            uint32_t *out, size_t cap, size_t *len) {
    int rc = wl_air_settings_validate(log_length, n_witness_columns, st);
    if (rc) return rc;
    const std::vector<uint32_t> d = wh::whir_domsep(air_whir_config(log_length, n_witness_columns, st));
    if (len) *len = d.size();
    if (out) {
        if (cap < d.size()) { g_err = "buffer too small"; return WL_ERR_BUFFER; }
        memcpy(out, d.data(), d.size() * 4);
    }
    return WL_OK;
}


wl_witness *wl_witness_upload(wl_ctx *c, const uint32_t *witness, uint32_t log_length) {
    try {
        HIP_CHECK(hipSetDevice(c->device));
        wl_witness *w = new wl_witness();
        w->ctx = c;
        w->log_length = log_length;
        w->packed = (F *)c->pool.get(packed_len(log_length, AIR_NCOLS) * 4);
        upload_packed(c, witness, log_length, AIR_NCOLS, w->packed);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return w;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void wl_witness_free(wl_witness *w) {
    if (!w) return;
    (void)hipStreamSynchronize(w->ctx->stream);
    w->ctx->pool.put(w->packed);
    delete w;
}

void wl_mem_stats(wl_ctx *c, uint64_t *held, uint64_t *peak_live, int reset_peak) {
    if (!c) return;
    if (held) *held = c->pool.bytes_held;
    if (peak_live) *peak_live = c->pool.peak_live;
    if (reset_peak) c->pool.peak_live = c->pool.live_bytes;
}

int wl_air_prove_resident(wl_table *t, const wl_air_settings *st, const uint32_t *p16c,
                          const uint32_t *p24c, const wl_witness *w, uint32_t *proof_out,
                          size_t proof_cap, size_t *proof_len) {
    int rc = check_settings(t, st);
    if (rc) return rc;
    if (!w || w->log_length != t->log_length || t->gen) { g_err = "witness size mismatch"; return WL_ERR_ARG; }
    try {
        HIP_CHECK(hipSetDevice(t->ctx->device));
        return prove_common(t, st, p16c, p24c, w->packed, proof_out, proof_cap, proof_len);
    } catch (const std::exception &e) {
        g_err = e.what();
        return WL_ERR_DEVICE;
    }
}

int wl_kernel_timing(wl_ctx *c, uint32_t class_mask) {
    if (!c) { g_err = "null context"; return WL_ERR_ARG; }
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 8; k++) {
        for (auto &e : c->kt_pairs[k]) c->kt_free.push_back(e);
        c->kt_pairs[k].clear();
        c->kt_work[k] = 0;
        c->kt_launches[k] = 0;
    }
    c->kt_mask = class_mask;
    return WL_OK;
}

int wl_kernel_timing_read(wl_ctx *c, int kernel_class, long *launches, double *total_ms, double *work) {
    if (!c || kernel_class < 0 || kernel_class >= 8) { g_err = "bad arguments"; return WL_ERR_ARG; }
    try {
        HIP_CHECK(hipStreamSynchronize(c->stream));
        double ms = 0;
        for (auto &e : c->kt_pairs[kernel_class]) {
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, e.first, e.second));
            ms += t;
        }
        if (launches) *launches = c->kt_launches[kernel_class];
        if (total_ms) *total_ms = ms;
        if (work) *work = c->kt_work[kernel_class];
        return WL_OK;
    } catch (const std::exception &e) {
        g_err = e.what();
        return WL_ERR_DEVICE;
    }
}

void wl_phase_times(wl_ctx *c, double *out7, int reset) {
    for (int i = 0; i < 7; i++) out7[i] = c->phase_ms[i];
    if (reset) for (int i = 0; i < 8; i++) c->phase_ms[i] = 0;
}

}  // extern "C"

void wl_set_error(const char *msg) { g_err = msg; }
