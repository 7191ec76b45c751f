// tune.hip -- autotuner of the commitment kernels' launch shapes (see tune.h) and
// its C ABI (wl_autotune_*).
#include <algorithm>
#include <array>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <vector>
#include "../../include/whirlaway_hip.h"
#include "tune.h"

namespace wt {
namespace {

using Key = std::array<int, 4>;  // device, kind, log size, width (leaves) or 0
constexpr int SAMPLES = 2;       // in-place timings per candidate (the smaller one counts)

// Tuning state of one (kernel, shape): the candidates (default first), their
// samples, and the event pairs of calls whose timing is not read yet.
struct Pending {
    hipEvent_t e0, e1;
    int cand;
};
struct Entry {
    int variant = -1;          // chosen variant; -1 while exploring
    long long best_ns = 0;     // the chosen one's best sample; -1 when forced
    std::vector<int> cands;
    std::vector<std::vector<float>> ms;
    std::vector<Pending> pending;
    int next = 0;
};
std::mutex g_mu;
std::map<Key, Entry> g_table;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_events;  // spare event pairs
int g_mode = -1;  // 0: defaults only; 1: tune in place (WHIRLAWAY_AUTOTUNE=0 turns it off)

int mode_locked() {
    if (g_mode < 0) {
        const char *e = getenv("WHIRLAWAY_AUTOTUNE");
        g_mode = (e && e[0] == '0') ? 0 : 1;
    }
    return g_mode;
}
std::pair<hipEvent_t, hipEvent_t> take_events() {
    if (!g_events.empty()) {
        auto p = g_events.back();
        g_events.pop_back();
        return p;
    }
    std::pair<hipEvent_t, hipEvent_t> p;
    HIP_CHECK(hipEventCreate(&p.first));
    HIP_CHECK(hipEventCreate(&p.second));
    return p;
}
// Collects the finished timings of e; decides when every candidate has SAMPLES
// (the smallest sample per candidate, ties to the earlier candidate).
void settle(Entry &e) {
    for (size_t i = 0; i < e.pending.size();) {
        Pending &p = e.pending[i];
        if (hipEventQuery(p.e1) != hipSuccess) { i++; continue; }
        float t = 0;
        if (hipEventElapsedTime(&t, p.e0, p.e1) == hipSuccess) e.ms[p.cand].push_back(t);
        g_events.emplace_back(p.e0, p.e1);
        e.pending.erase(e.pending.begin() + (long)i);
    }
    if (e.variant >= 0) return;
    size_t bi = 0;
    float bt = 1e30f;
    for (size_t c = 0; c < e.cands.size(); c++) {
        if (e.ms[c].size() < (size_t)SAMPLES) return;
        const float t = *std::min_element(e.ms[c].begin(), e.ms[c].end());
        if (t < bt) { bt = t; bi = c; }
    }
    e.variant = e.cands[bi];
    e.best_ns = (long long)((double)bt * 1e6);
}

// Runs one call of a (kernel, shape) with the variant the table gives.  While the
// shape is being explored, each call runs the next candidate that still needs a
// sample, between two events on the calling stream: the timing is that of the
// call itself, in place in the caller's work (timed alone, back to back, variants
// rank differently on this GPU than they run inside a proof).  The events are read
// on later calls, once they have completed.  run(v) enqueues variant v and returns
// what the caller needs (the NTT's storage order).
template <class Run>
auto run_tuned(wl_ctx *c, const Key &k, const std::function<std::vector<int>()> &candidates, Run run)
    -> decltype(run(0)) {
    int v = -1, cand = -1;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_table.find(k);
        if (it == g_table.end() && mode_locked() == 1) {
            Entry e;
            e.cands = candidates();
            e.ms.resize(e.cands.size());
            if (e.cands.size() < 2) {
                e.variant = e.cands.empty() ? 0 : e.cands[0];
                e.best_ns = 0;
            }
            it = g_table.emplace(k, std::move(e)).first;
        }
        if (it != g_table.end()) {
            Entry &e = it->second;
            settle(e);
            v = e.variant;
            if (v < 0) {
                // the next candidate short of samples, counting calls still in flight
                const size_t nc = e.cands.size();
                for (size_t j = 0; j < nc && cand < 0; j++) {
                    const size_t i = ((size_t)e.next + j) % nc;
                    size_t inflight = 0;
                    for (const Pending &p : e.pending) inflight += p.cand == (int)i;
                    if (e.ms[i].size() + inflight < (size_t)SAMPLES) cand = (int)i;
                }
                if (cand < 0) cand = e.next % (int)nc;  // all in flight: one more sample
                e.next = (cand + 1) % (int)nc;
                v = e.cands[cand];
                ev = take_events();
            }
        }
    }
    if (cand < 0) return run(v);  // negative v: the default
    HIP_CHECK(hipEventRecord(ev.first, c->stream));
    auto r = run(v);
    HIP_CHECK(hipEventRecord(ev.second, c->stream));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(k);
    if (it != g_table.end()) it->second.pending.push_back(Pending{ev.first, ev.second, cand});
    else g_events.push_back(ev);  // the table was cleared meanwhile
    return r;
}

// split variants of 2^L rows with distinct LB, the default first
std::vector<int> ntt_candidates(int L) {
    std::vector<int> out;
    std::set<int> seen;
    for (int v = 0; v < wk::NTT_VARIANTS; v++) {
        int la, lb;
        if (!wk::ntt_split(L, &la, &lb, v)) break;
        if (seen.insert(lb).second) out.push_back(v);
    }
    return out;
}
// tree schedules with distinct launch sequences for 2^log_rows leaves, the default first
std::vector<int> tree_candidates(int log_rows) {
    std::vector<int> out;
    std::set<std::vector<long long>> seen;
    for (int v = 0; v < wk::TREE_VARIANTS; v++) {
        int W, Pn;
        wk::merkle_tree_schedule(v, &W, &Pn);
        std::vector<long long> sig;
        long long cnt = (1ll << log_rows) >> 1;
        for (; cnt >= (1ll << W); cnt >>= 1) sig.push_back(cnt < (1ll << (W + Pn)) ? -cnt : cnt);
        sig.push_back(cnt);
        if (seen.insert(sig).second) out.push_back(v);
    }
    return out;
}
std::vector<int> leaf_candidates(int log_rows) {
    std::vector<int> out{wk::merkle_leaves_default(log_rows)};
    for (int v = 0; v < wk::LEAF_VARIANTS; v++) {
        if ((v & 1) && log_rows < 5) continue;  // lane pairs need rows a multiple of 32
        if (v != out[0]) out.push_back(v);
    }
    return out;
}

}  // namespace

int ntt_variant(wl_ctx *c, int L) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(Key{c->device, TK_NTT, L, 0});
    return it == g_table.end() || it->second.variant < 0 ? 0 : it->second.variant;
}

int rs_encode(wl_ctx *c, const F *cols, int n_cols, int n_nonzero, int log_m, int L, F *cw, F *tmp) {
    const F *tw = c->twiddle(L);
    auto run = [&](int v) {
        if (v < 0) v = 0;
        return wk::rs_encode(cols, n_cols, n_nonzero, log_m, L, cw, tmp, tw, c->ntt_tabs(L, v), c->stream, false, v);
    };
    int la, lb;
    if (n_nonzero <= 0 || !wk::ntt_split(L, &la, &lb, 0)) return run(0);
    return run_tuned(c, Key{c->device, TK_NTT, L, 0}, [&] { return ntt_candidates(L); }, run);
}

void merkle_leaves(wl_ctx *c, const F *cw, int width, int nnz, int log_rows, int la, const void *pc, F *nodes) {
    auto run = [&](int v) {
        wk::merkle_leaves(cw, width, nnz, log_rows, la, pc, nodes, c->stream, v);
        return 0;
    };
    run_tuned(c, Key{c->device, TK_LEAVES, log_rows, width}, [&] { return leaf_candidates(log_rows); }, run);
}

void merkle_nodes(wl_ctx *c, int log_rows, const void *pc, F *nodes) {
    auto run = [&](int v) {
        wk::merkle_nodes(log_rows, pc, nodes, c->stream, v);
        return 0;
    };
    run_tuned(c, Key{c->device, TK_TREE, log_rows, 0}, [&] { return tree_candidates(log_rows); }, run);
}

}  // namespace wt

extern "C" {

int wl_autotune_mode(int mode) {
    std::lock_guard<std::mutex> lk(wt::g_mu);
    const int prev = wt::mode_locked();
    if (mode == 0 || mode == 1) wt::g_mode = mode;
    return prev;
}

int wl_autotune_force(int device, int kind, int log_size, int width, int variant) {
    const int nv = kind == wt::TK_NTT ? wk::NTT_VARIANTS : kind == wt::TK_LEAVES ? wk::LEAF_VARIANTS
                 : kind == wt::TK_TREE ? wk::TREE_VARIANTS : 0;
    if (nv == 0 || variant < 0 || variant >= nv || log_size < 0 || log_size > 30 || width < 0) {
        wl_set_error("autotune: bad kind, variant or shape");
        return WL_ERR_ARG;
    }
    if (kind != wt::TK_LEAVES) width = 0;
    std::lock_guard<std::mutex> lk(wt::g_mu);
    wt::Entry &e = wt::g_table[wt::Key{device, kind, log_size, width}];
    e.variant = variant;
    e.best_ns = -1;
    return WL_OK;
}

size_t wl_autotune_entries(int64_t *out, size_t cap) {
    std::lock_guard<std::mutex> lk(wt::g_mu);
    size_t i = 0;
    for (const auto &kv : wt::g_table) {
        if (out && i < cap) {
            int64_t *e = out + 6 * i;
            for (int j = 0; j < 4; j++) e[j] = kv.first[j];
            e[4] = kv.second.variant;
            e[5] = kv.second.best_ns;
        }
        i++;
    }
    return i;
}

void wl_autotune_clear(void) {
    std::lock_guard<std::mutex> lk(wt::g_mu);
    // event pairs of calls still in flight stay with their entries' memory: wait for
    // them before reusing the pairs
    for (auto &kv : wt::g_table)
        for (const wt::Pending &p : kv.second.pending) {
            (void)hipEventSynchronize(p.e1);
            wt::g_events.emplace_back(p.e0, p.e1);
        }
    wt::g_table.clear();
}

}  // extern "C"
