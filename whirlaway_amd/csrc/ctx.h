// ctx.h -- prover context (stream, device memory pool, pinned staging, twiddle
// cache) and small host<->device helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#include "common.h"
#include "host.h"
#include "kernels.h"

namespace wl {
using namespace kb;

struct P2Dev {
    P2Consts16 p16;
    P2Consts24 p24;
};

// Size-keyed caching allocator: blocks return to a free list and are reused by
// later requests of the same rounded size (no hipMalloc/hipFree in steady state).
struct Pool {
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t bytes_held = 0;
    void *get(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        auto it = free_.find(bytes);
        void *p;
        if (it != free_.end()) {
            p = it->second;
            free_.erase(it);
        } else {
            HIP_CHECK(hipMalloc(&p, bytes));
            bytes_held += bytes;
        }
        live_[p] = bytes;
        return p;
    }
    void put(void *p) {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) return;
        free_.emplace(it->second, p);
        live_.erase(it);
    }
    void release() {
        for (auto &kv : free_) (void)hipFree(kv.second);
        for (auto &kv : live_) (void)hipFree(kv.first);
        free_.clear();
        live_.clear();
        bytes_held = 0;
    }
};

template <class T>
struct DBuf {
    Pool *pool = nullptr;
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(Pool &pl, size_t count) : pool(&pl), p((T *)pl.get(count * sizeof(T))), n(count) {}
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; }
    DBuf &operator=(DBuf &&o) noexcept {
        reset();
        pool = o.pool; p = o.p; n = o.n; o.p = nullptr;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() { if (pool && p) pool->put(p); p = nullptr; }
    operator T *() const { return p; }
};

}  // namespace wl

struct wl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    wl::Pool pool;
    uint32_t *pinned = nullptr;   // host staging
    size_t pinned_words = 0;
    std::map<int, kb::F *> tw;    // twiddle tables per log size
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double phase_ms[8] = {0};

    kb::F *twiddle(int L) {
        auto it = tw.find(L);
        if (it != tw.end()) return it->second;
        kb::F *t;
        HIP_CHECK(hipMalloc(&t, ((size_t)1 << L) * sizeof(kb::F)));
        wk::twiddles(t, L, stream);
        tw[L] = t;
        return t;
    }
    uint32_t *stage(size_t words) {
        if (words > pinned_words) {
            if (pinned) (void)hipHostFree(pinned);
            size_t w = std::max(words, (size_t)1 << 20);
            HIP_CHECK(hipHostMalloc((void **)&pinned, w * 4, hipHostMallocDefault));
            pinned_words = w;
        }
        return pinned;
    }
    // device -> host, blocking on the stream
    template <class T>
    void d2h(T *dst, const T *src, size_t count) {
        if (!count) return;
        uint32_t *st = stage((count * sizeof(T) + 3) / 4);
        HIP_CHECK(hipMemcpyAsync(st, src, count * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        memcpy((void *)dst, st, count * sizeof(T));
    }
    // host -> device through the pinned buffer (returns after the copy completed)
    template <class T>
    void h2d(T *dst, const T *src, size_t count) {
        if (!count) return;
        uint32_t *st = stage((count * sizeof(T) + 3) / 4);
        memcpy(st, (const void *)src, count * sizeof(T));
        HIP_CHECK(hipMemcpyAsync(dst, st, count * sizeof(T), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
};

// AirTable for the Poseidon2 AIR: trace length, skips and round constants.
struct wl_table {
    wl_ctx *ctx;
    uint32_t log_length;
    uint32_t univariate_skips;
    wk::AirConsts air_host;
    wk::AirConsts *air_dev = nullptr;
};

// Trace resident in HBM as the commitment's packed polynomial (Montgomery form,
// 165 columns x 2^n then zeros up to 2^(n+8)).
struct wl_witness {
    wl_ctx *ctx;
    uint32_t log_length;
    kb::F *packed;
};

namespace wl {

inline EF reduce_one(wl_ctx *c, const EF *parts, int n_parts) {
    DBuf<EF> out(c->pool, 1);
    wk::reduce_parts(parts, 1, n_parts, out, c->stream);
    EF r;
    c->d2h(&r, out.p, 1);
    return r;
}
inline std::vector<EF> reduce_many(wl_ctx *c, const EF *parts, int n_items, int n_parts) {
    DBuf<EF> out(c->pool, n_items);
    wk::reduce_parts(parts, n_items, n_parts, out, c->stream);
    std::vector<EF> r(n_items);
    c->d2h(r.data(), out.p, n_items);
    return r;
}
inline DBuf<EF> upload_ef(wl_ctx *c, const std::vector<EF> &v) {
    DBuf<EF> d(c->pool, v.size() ? v.size() : 1);
    if (!v.empty()) c->h2d(d.p, v.data(), v.size());
    return d;
}
// eq(point, .) * scale on the device (2^n EF)
inline DBuf<EF> eq_dev(wl_ctx *c, const std::vector<EF> &pt, const EF &scale) {
    size_t N = (size_t)1 << pt.size();
    DBuf<EF> out(c->pool, N);
    std::vector<EF> tmp(pt);
    tmp.push_back(scale);
    DBuf<EF> d = upload_ef(c, tmp);
    wk::eq_table(d.p, (int)pt.size(), d.p + pt.size(), out, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));  // `d` is returned to the pool below
    return out;
}

inline void load_p2(const uint32_t *p16c, const uint32_t *p24c, P2Dev &host) {
    memset(&host, 0, sizeof host);
    int k = 0;
    if (p16c) {
        for (int r = 0; r < 4; r++) for (int i = 0; i < 16; i++) host.p16.ext_init[r][i] = to_mont(p16c[k++] % P);
        for (int r = 0; r < 4; r++) for (int i = 0; i < 16; i++) host.p16.ext_term[r][i] = to_mont(p16c[k++] % P);
        for (int r = 0; r < 20; r++) host.p16.internal[r] = to_mont(p16c[k++] % P);
    }
    k = 0;
    if (p24c) {
        for (int r = 0; r < 4; r++) for (int i = 0; i < 24; i++) host.p24.ext_init[r][i] = to_mont(p24c[k++] % P);
        for (int r = 0; r < 4; r++) for (int i = 0; i < 24; i++) host.p24.ext_term[r][i] = to_mont(p24c[k++] % P);
        for (int r = 0; r < 23; r++) host.p24.internal[r] = to_mont(p24c[k++] % P);
    }
}
inline DBuf<P2Dev> upload_p2(wl_ctx *c, const P2Dev &host) {
    DBuf<P2Dev> dev(c->pool, 1);
    c->h2d(dev.p, &host, 1);
    return dev;
}

inline int pow_bits_sumcheck(int sec, int degree) {  // crates/sumcheck/src/lib.rs:21-31
    int t = std::max(0, 248 - (int)wh::log2_up((size_t)degree + 1));
    return std::max(0, sec - t);
}
inline int pow_bits_air(int sec, int x) { return std::max(0, sec - std::max(0, 248 - x)); }

// PoW grinding on the device: minimal witness w with sample_bits(bits) == 0
// after observe(w) (p3-challenger grind; the minimal witness is taken).
inline void grind(wl_ctx *c, wh::Transcript &ts, const P2Dev *pdev, int bits) {
    if (bits <= 0) return;
    F st[16];
    for (int i = 0; i < 16; i++) st[i] = ts.state[i];
    for (int i = 0; i < ts.n_in; i++) st[i] = ts.in_buf[i];
    int pos = ts.n_in;  // observe(w) places w here; the sample then duplexes
    DBuf<F> dst(c->pool, 16);
    DBuf<uint32_t> found(c->pool, 1);
    c->h2d(dst.p, st, 16);
    uint32_t mask = (uint32_t)((1ull << bits) - 1);
    uint32_t w = 0xffffffffu;
    const uint32_t chunk = 1u << 22;
    for (uint64_t base = 0; base < P && w == 0xffffffffu; base += chunk) {
        uint32_t init = 0xffffffffu;
        c->h2d(found.p, &init, 1);
        uint32_t cnt = (uint32_t)std::min<uint64_t>(chunk, P - base);
        wk::grind(pdev, dst, pos, mask, (uint32_t)base, cnt, found, c->stream);
        c->d2h(&w, found.p, 1);
    }
    if (w == 0xffffffffu) throw std::runtime_error("grinding found no witness");
    ts.observe(to_mont(w));
    (void)ts.sample_bits(bits);
    ts.data.push_back(w);
}

}  // namespace wl

// per-thread message of the last failed C-ABI call
void wl_set_error(const char *msg);
