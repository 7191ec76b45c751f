// ctx.h -- prover context (stream, device memory pool, pinned staging, twiddle
// cache) and small host<->device helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <chrono>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
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
    size_t live_bytes = 0, peak_live = 0;
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
        live_bytes += bytes;
        if (live_bytes > peak_live) peak_live = live_bytes;
        return p;
    }
    void put(void *p) {
        if (!p) return;
        auto it = live_.find(p);
        if (it == live_.end()) return;
        free_.emplace(it->second, p);
        live_bytes -= it->second;
        live_.erase(it);
    }
    void release() {
        for (auto &kv : free_) (void)hipFree(kv.second);
        for (auto &kv : live_) (void)hipFree(kv.first);
        free_.clear();
        live_.clear();
        bytes_held = 0;
        live_bytes = 0;
    }
};

// Frees of device or pinned memory wait for the whole device, and a proof on a
// context with device-side waits keeps kernels spinning on the GPU until its host
// thread releases them.  A free on one thread while another thread's proof has
// such a wait pending can therefore stall both until the wait times out.  Proofs
// with device waits enter this guard for their duration; wl_destroy frees only
// once none is running, and new proofs wait while a free is pending.
struct FreeGuard {
    std::mutex mu;
    std::condition_variable cv;
    int proofs = 0, freeing = 0;
    void enter_proof() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return freeing == 0; });
        proofs++;
    }
    void leave_proof() {
        std::lock_guard<std::mutex> lk(mu);
        if (--proofs == 0) cv.notify_all();
    }
    void enter_free() {
        std::unique_lock<std::mutex> lk(mu);
        freeing++;
        cv.wait(lk, [&] { return proofs == 0; });
    }
    void leave_free() {
        std::lock_guard<std::mutex> lk(mu);
        freeing--;
        cv.notify_all();
    }
};
inline FreeGuard &free_guard() {
    static FreeGuard g;
    return g;
}

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
    // zerocheck rounds without grinding run their transcript on the device (off by
    // default: one wave's serial Poseidon2-16 duplexes take longer than the host
    // round trip they replace)
    bool device_transcript = false;
    // sumcheck rounds are enqueued once the host has their challenge, instead of
    // ahead of it behind a device wait (a kernel spinning on mapped memory holds its
    // hardware queue; streams that share the queue stall behind it)
    bool host_rounds = false;
    hipStream_t stream = nullptr;
    // a first codeword encoded while the witness uploaded (commit_codeword takes it
    // instead of encoding cols again): the source it encodes, its shape, the buffer
    struct PreEncoded {
        const kb::F *cols = nullptr;
        int log_rows = 0, n_cols = 0, nnz = 0, la = 0;
        wl::DBuf<kb::F> cw;
    } pre0;
    wl::Pool pool;
    uint32_t *pinned = nullptr;   // host staging
    size_t pinned_words = 0;
    std::map<int, kb::F *> tw;    // twiddle tables per log size
    std::map<int, kb::F *> ntt_tab;  // compact two-pass NTT tables per (log size, split)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double phase_ms[8] = {0};
    // live timing of kernel classes (HIP events on this context's stream)
    uint32_t kt_mask = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> kt_pairs[8], kt_free;
    double kt_work[8] = {0};
    long kt_launches[8] = {0};

    kb::F *twiddle(int L) {
        auto it = tw.find(L);
        if (it != tw.end()) return it->second;
        kb::F *t;
        HIP_CHECK(hipMalloc(&t, ((size_t)1 << L) * sizeof(kb::F)));
        wk::twiddles(t, L, stream);
        tw[L] = t;
        return t;
    }
    // compact twiddle tables of the two-pass RS encoding with split `variant`
    // (wk::ntt_split; nullptr when unused)
    kb::F *ntt_tabs(int L, int variant = 0) {
        int la, lb;
        if (!wk::ntt_split(L, &la, &lb, variant)) return nullptr;
        const int key = L * 32 + lb;
        auto it = ntt_tab.find(key);
        if (it != ntt_tab.end()) return it->second;
        kb::F *t;
        HIP_CHECK(hipMalloc(&t, wk::ntt_tables_words(L, variant) * sizeof(kb::F)));
        wk::ntt_tables(twiddle(L), L, t, stream, variant);
        ntt_tab[key] = t;
        return t;
    }
    // Host -> device copy of a large pageable buffer (the witness) through pinned
    // staging.  The buffer goes in chunks of 2, 4, 8, then 16 MiB (CHUNK) through
    // STAGE_SLOTS pinned slots of CHUNK bytes; STAGE_THREADS host threads copy a
    // stripe of every chunk each, and the calling thread enqueues a chunk's DMA as
    // soon as all stripes are in, in chunk order (small first chunks start the copy
    // engine early, large later ones keep the per-DMA cost low).  A slot is refilled once the DMA that read it has completed
    // (its event).  The copy engine then runs at PCIe speed one chunk behind the
    // host copies, instead of the runtime's single staging path for pageable
    // memory.  Returns after every DMA is enqueued; later work on the stream is
    // ordered behind them.
    // With `ready`, the DMAs run on a copy stream instead, and as each chunk lands
    // the work stream waits for it and ready(u0, u1) enqueues work on the units
    // [u0, u1) of `unit` bytes that are now complete (the last one may be partial),
    // so that work overlaps the rest of the copy.
    static constexpr int STAGE_THREADS = 6, STAGE_SLOTS = 4;
    static constexpr size_t CHUNK = (size_t)16 << 20, FIRST_CHUNK = (size_t)2 << 20;
    uint8_t *stage_slots = nullptr;
    hipEvent_t stage_ev[STAGE_SLOTS] = {};
    bool stage_used[STAGE_SLOTS] = {};
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_order = nullptr;
    void h2d_staged(void *dst, const void *src, size_t bytes, size_t unit = 0,
                    const std::function<void(size_t, size_t)> &ready = nullptr) {
        const size_t n_units = unit ? (bytes + unit - 1) / unit : 0;
        if (bytes < 4 * FIRST_CHUNK) {
            HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
            if (ready) ready(0, n_units);
            return;
        }
        hipStream_t dma = stream;
        if (ready) {
            if (!copy_stream) {
                HIP_CHECK(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
                HIP_CHECK(hipEventCreateWithFlags(&copy_order, hipEventDisableTiming));
            }
            // the copies start after the work already on the stream (the destination
            // may be pool memory that this work still reads)
            HIP_CHECK(hipEventRecord(copy_order, stream));
            HIP_CHECK(hipStreamWaitEvent(copy_stream, copy_order, 0));
            dma = copy_stream;
        }
        size_t units_done = 0;
        if (!stage_slots) {
            HIP_CHECK(hipHostMalloc((void **)&stage_slots, STAGE_SLOTS * CHUNK, hipHostMallocDefault));
            for (auto &e : stage_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        // slots an earlier call filled may still feed its DMAs
        for (int k = 0; k < STAGE_SLOTS; k++)
            if (stage_used[k]) HIP_CHECK(hipEventSynchronize(stage_ev[k]));
        std::vector<size_t> offs{0};
        for (size_t len = FIRST_CHUNK; offs.back() < bytes; len = std::min(CHUNK, 2 * len))
            offs.push_back(std::min(bytes, offs.back() + len));
        const size_t nch = offs.size() - 1;
        // filled[k]: stripes of chunk k in its slot; issued: chunks whose DMA is enqueued
        std::vector<std::atomic<int>> filled(nch);
        for (auto &f : filled) f.store(0, std::memory_order_relaxed);
        std::atomic<size_t> issued{0};
        std::atomic<bool> failed{false};
        std::string err;
        auto copier = [&](int t) {
            for (size_t k = 0; k < nch && !failed.load(std::memory_order_relaxed); k++) {
                if (k >= STAGE_SLOTS) {
                    // the slot's previous chunk must have gone out, and its DMA finished
                    while (issued.load(std::memory_order_acquire) <= k - STAGE_SLOTS) {
                        if (failed.load(std::memory_order_relaxed)) return;
                        std::this_thread::yield();
                    }
                    if (hipEventSynchronize(stage_ev[k % STAGE_SLOTS]) != hipSuccess) {
                        failed.store(true);
                        return;
                    }
                }
                const size_t off = offs[k], n = offs[k + 1] - off;
                const size_t stripe = (n / STAGE_THREADS + 63) & ~(size_t)63;
                const size_t a = std::min(n, (size_t)t * stripe), b = std::min(n, a + stripe);
                if (b > a) memcpy(stage_slots + (k % STAGE_SLOTS) * CHUNK + a, (const uint8_t *)src + off + a, b - a);
                filled[k].fetch_add(1, std::memory_order_release);
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < STAGE_THREADS; t++) pool.emplace_back(copier, t);
        try {
            for (size_t k = 0; k < nch; k++) {
                while (filled[k].load(std::memory_order_acquire) < STAGE_THREADS) {
                    if (failed.load(std::memory_order_relaxed)) throw std::runtime_error("staging copy failed");
                    std::this_thread::yield();
                }
                const int slot = (int)(k % STAGE_SLOTS);
                const size_t off = offs[k], n = offs[k + 1] - off;
                HIP_CHECK(hipMemcpyAsync((uint8_t *)dst + off, stage_slots + (size_t)slot * CHUNK, n,
                                         hipMemcpyHostToDevice, dma));
                HIP_CHECK(hipEventRecord(stage_ev[slot], dma));
                stage_used[slot] = true;
                issued.store(k + 1, std::memory_order_release);
                if (ready) {
                    const size_t u = k + 1 == nch ? n_units : (off + n) / unit;
                    if (u > units_done) {
                        HIP_CHECK(hipStreamWaitEvent(stream, stage_ev[slot], 0));
                        ready(units_done, u);
                        units_done = u;
                    }
                }
            }
        } catch (const std::exception &e) {
            err = e.what();
            failed.store(true);
        }
        for (auto &th : pool) th.join();
        if (!err.empty()) {
            // no copy may still be writing once the caller returns the destination
            if (ready) (void)hipStreamSynchronize(copy_stream);
            throw std::runtime_error(err);
        }
    }
    std::vector<void *> retired_host;  // outgrown pinned buffers, freed by wl_destroy
    uint32_t *stage(size_t words) {
        if (words > pinned_words) {
            // no free here: it would wait for the whole device (see wl::FreeGuard)
            if (pinned) retired_host.push_back(pinned);
            size_t w = std::max(words, (size_t)1 << 20);
            HIP_CHECK(hipHostMalloc((void **)&pinned, w * 4, hipHostMallocDefault));
            pinned_words = w;
        }
        return pinned;
    }
    // device -> host (waits for the data; see fetch)
    template <class T>
    void d2h(T *dst, const T *src, size_t count) {
        fetch((void *)dst, (const void *)src, count * sizeof(T));
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
    // Device -> host results of each protocol round: a kernel writes them into
    // host-mapped pinned memory and then stores a sequence number; the host
    // waits on that flag (far cheaper than a copy plus a stream synchronize).
    // word offsets in the mapped buffer of results read after later publishes
    static constexpr size_t ISUM_OFF = (size_t)1 << 19, GRIND_OFF = ((size_t)1 << 19) + 256;
    uint32_t *mapped = nullptr;      // host view
    uint32_t *mapped_dev = nullptr;  // device view
    size_t mapped_words = 0;
    uint32_t seq = 0;
    // device words that kernels leave as they found them: {smallest grinding
    // witness, grinding blocks done} armed as {0xffffffff, 0}, the block counter
    // of multi-block publishes (0), the witness of a grinding group's earlier
    // passes (0xffffffff)
    uint32_t *sync_dev = nullptr;
    // Poseidon2 constants of the last proof on the device (re-sent only when they change)
    wl::P2Dev *p2_dev = nullptr;
    wl::P2Dev p2_host{};
    uint32_t *sync_words() {
        if (sync_dev) return sync_dev;
        HIP_CHECK(hipMalloc((void **)&sync_dev, 4 * sizeof(uint32_t)));
        const uint32_t armed[4] = {0xffffffffu, 0u, 0u, 0xffffffffu};
        HIP_CHECK(hipMemcpyAsync(sync_dev, armed, sizeof(armed), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        return sync_dev;
    }
    uint32_t *grind_state() { return sync_words(); }
    uint32_t *publish_counter() { return sync_words() + 2; }
    void ensure_mapped() {
        if (mapped) return;
        mapped_words = (size_t)1 << 20;
        HIP_CHECK(hipHostMalloc((void **)&mapped, (mapped_words + 64) * 4,
                                hipHostMallocMapped | hipHostMallocCoherent));
        HIP_CHECK(hipHostGetDevicePointer((void **)&mapped_dev, mapped, 0));
        memset(mapped, 0, (mapped_words + 64) * 4);
    }
    volatile uint32_t *flag_dev() { return (volatile uint32_t *)(mapped_dev + mapped_words); }
    void wait_flag(uint32_t s) {
        volatile uint32_t *f = (volatile uint32_t *)(mapped + mapped_words);
        long spins = 0;
        auto t0 = std::chrono::steady_clock::now();
        // sequence numbers are published in stream order, so a later one implies s
        while ((int32_t)(*f - s) < 0) {
            if (++spins % 4096 == 0) {
                double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (el > 20.0) {
                    HIP_CHECK(hipStreamSynchronize(stream));
                    if ((int32_t)(*f - s) < 0) throw std::runtime_error("device result was not published");
                }
            }
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        if (mbox && __atomic_load_n(&mbox[1], __ATOMIC_RELAXED) != 0) {
            // the failed proof's remaining device waits are released and drained, and
            // the report cleared, so that the context serves later calls
            release_all();
            (void)hipStreamSynchronize(stream);
            __atomic_store_n(&mbox[1], 0u, __ATOMIC_RELEASE);
            throw std::runtime_error("a device wait for a host value timed out");
        }
        phase_ms[6] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    // ---- host values for pre-enqueued device work ----
    // Kernels of round j + 1 are enqueued before the host has computed round j's
    // challenge: an await_host kernel ahead of them blocks the stream until the
    // host releases the value (a store into host-mapped memory), so the host's
    // launch calls stay off the critical path.  Ids increase by one per value;
    // mbox[0] holds the last released id, mbox[1] a timeout report.
    static constexpr int MB_HEAD = 64, MB_SLOTS = 256, MB_WORDS = 64;
    uint32_t *mbox = nullptr, *mbox_dev = nullptr, *args_dev = nullptr;
    uint32_t mbox_next = 0;
    uint64_t await_ticks = 0;
    double await_limit_s = 10.0;  // bound of each device wait (wl_set_wait_timeout)
    int await_khz = 0;
    struct HostSlot { uint32_t id; uint32_t *host; uint32_t *dev; };
    void ensure_mbox() {
        if (mbox) return;
        size_t words = MB_HEAD + (size_t)MB_SLOTS * MB_WORDS;
        HIP_CHECK(hipHostMalloc((void **)&mbox, words * 4, hipHostMallocMapped | hipHostMallocCoherent));
        memset(mbox, 0, words * 4);
        HIP_CHECK(hipHostGetDevicePointer((void **)&mbox_dev, mbox, 0));
        HIP_CHECK(hipMalloc((void **)&args_dev, (size_t)MB_SLOTS * MB_WORDS * 4));
        int khz = 0;
        HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
        await_khz = khz > 0 ? khz : 100000;
        set_await_limit(await_limit_s);
    }
    void set_await_limit(double seconds) {
        await_limit_s = seconds;
        if (await_khz) await_ticks = (uint64_t)((double)await_khz * 1000.0 * seconds);
    }
    // enqueue a wait for the next host value (<= MB_WORDS words), readable by later
    // kernels at .dev
    HostSlot await_value(int words) {
        ensure_mbox();
        if (words > MB_WORDS) throw std::runtime_error("host value too large");
        uint32_t id = ++mbox_next;
        size_t off = MB_HEAD + (size_t)(id % MB_SLOTS) * MB_WORDS;
        HostSlot sl{id, mbox + off, args_dev + (off - MB_HEAD)};
        wk::await_host(mbox_dev, id, mbox_dev + off, sl.dev, words, await_ticks, mbox_dev + 1, stream);
        return sl;
    }
    HostSlot next_slot() {
        ensure_mbox();
        uint32_t id = ++mbox_next;
        size_t off = MB_HEAD + (size_t)(id % MB_SLOTS) * MB_WORDS;
        return HostSlot{id, mbox + off, args_dev + (off - MB_HEAD)};
    }
    void release(const HostSlot &sl, const void *data, size_t bytes) {
        memcpy(sl.host, data, bytes);
        __atomic_store_n(&mbox[0], sl.id, __ATOMIC_RELEASE);
    }
    // error paths: let every enqueued wait finish
    void release_all() {
        if (mbox) __atomic_store_n(&mbox[0], mbox_next, __ATOMIC_RELEASE);
    }
    // copy `bytes` of device memory to dst through the mapped buffer (waits); the
    // first n_canon words are F values delivered in canonical form
    // two device ranges in one round trip (two publishes, one wait)
    void fetch2(void *dst0, const void *src0, size_t bytes0, void *dst1, const void *src1, size_t bytes1) {
        ensure_mapped();
        const size_t w0 = (bytes0 + 3) / 4, w1 = (bytes1 + 3) / 4;
        if (!bytes0 || !bytes1 || w0 + w1 > mapped_words) {
            fetch(dst0, src0, bytes0);
            fetch(dst1, src1, bytes1);
            return;
        }
        const uint32_t s0 = ++seq;
        wk::publish((const uint32_t *)src0, w0, mapped_dev, flag_dev(), s0, publish_counter(), stream);
        const uint32_t s1 = ++seq;
        wk::publish((const uint32_t *)src1, w1, mapped_dev + w0, flag_dev(), s1, publish_counter(), stream);
        wait_flag(s1);
        memcpy(dst0, mapped, bytes0);
        memcpy(dst1, mapped + w0, bytes1);
    }
    void fetch(void *dst, const void *src, size_t bytes, size_t n_canon = 0) {
        if (!bytes) return;
        ensure_mapped();
        size_t words = (bytes + 3) / 4;
        if (words > mapped_words) {
            uint32_t *st = stage(words);
            HIP_CHECK(hipMemcpyAsync(st, src, bytes, hipMemcpyDeviceToHost, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            for (size_t i = 0; i < std::min(n_canon, words); i++) st[i] = kb::from_mont(kb::F{st[i]});
            memcpy(dst, st, bytes);
            return;
        }
        uint32_t s = ++seq;
        wk::publish((const uint32_t *)src, words, mapped_dev, flag_dev(), s, publish_counter(), stream, n_canon);
        wait_flag(s);
        memcpy(dst, mapped, bytes);
    }
    // Asynchronous host -> device: the source is copied into a pinned ring that
    // is never rewritten before the stream has drained (ring_reset), so the
    // copy needs no synchronization.
    uint32_t *ring = nullptr;
    size_t ring_words = 0, ring_off = 0;
    void ring_reset() {
        HIP_CHECK(hipStreamSynchronize(stream));
        ring_off = 0;
    }
    template <class T>
    void h2d_async(T *dst, const T *src, size_t count) {
        if (!count) return;
        size_t words = (count * sizeof(T) + 255) / 256 * 64;
        if (!ring) {
            ring_words = (size_t)1 << 22;  // 16 MiB
            HIP_CHECK(hipHostMalloc((void **)&ring, ring_words * 4, hipHostMallocDefault));
        }
        if (words > ring_words) { h2d(dst, src, count); return; }
        if (ring_off + words > ring_words) ring_reset();
        uint32_t *st = ring + ring_off;
        ring_off += words;
        memcpy(st, (const void *)src, count * sizeof(T));
        HIP_CHECK(hipMemcpyAsync(dst, st, count * sizeof(T), hipMemcpyHostToDevice, stream));
    }
};

namespace wg { struct AirSpec; }

// AirTable: trace length, skips and either the Poseidon2 AIR's round constants
// (the specialised kernels) or, when gen is set, an AIR given as a constraint
// program with its preprocessed columns.
struct wl_table {
    wl_ctx *ctx;
    uint32_t log_length;
    uint32_t univariate_skips;
    wg::AirSpec *gen = nullptr;
    wk::AirConsts air_host;
    wk::ChainLin chain_host;
    wk::AirConsts *air_dev = nullptr;   // from the context's pool
    wk::ChainLin *chain_dev = nullptr;
};

// Fiat-Shamir state of whir-p3's ProverState / VerifierState over
// DuplexChallenger<F, Poseidon2-16, 16, 8> (src/examples/poseidon2.rs:27,157-163):
// the sponge, the proof_data written (prover) or being read (verifier).  Not
// copyable: ts points at p16.
struct wl_transcript {
    kb::P2Consts16 p16;           // Montgomery form
    wh::Transcript ts;            // prover: ts.data is proof_data
    std::vector<uint32_t> proof;  // verifier: the proof_data read from pos on
    size_t pos = 0;
    bool verifier = false;
    wl_transcript() : ts(&p16) {}
    wl_transcript(const wl_transcript &) = delete;
    wl_transcript &operator=(const wl_transcript &) = delete;
};

// Trace resident in HBM as the commitment's packed polynomial (Montgomery form,
// 165 columns x 2^n then zeros up to 2^(n+8)).
struct wl_witness {
    wl_ctx *ctx;
    uint32_t log_length;
    kb::F *packed;   // from the context's pool
};

namespace wl {

// Kernel classes that can be timed live (wl_kernel_timing).
enum KernelClass { KT_NONE = 0, KT_MERKLE_LEAVES = 1, KT_ZC_ROUND = 2, KT_NTT = 3, KT_ZC_ROUND0 = 4,
                   KT_GRIND = 5, KT_MERKLE_NODES = 6 };

// Records a HIP event pair around the launches in its scope when class `id` is
// being timed; `work` is the algorithmic work of those launches.
struct KTimer {
    wl_ctx *c;
    bool on;
    std::pair<hipEvent_t, hipEvent_t> ev;
    int id;
    KTimer(wl_ctx *ctx, int cls, double work) : c(ctx), on((ctx->kt_mask >> cls) & 1), id(cls) {
        if (!on) return;
        if (!c->kt_free.empty()) {
            ev = c->kt_free.back();
            c->kt_free.pop_back();
        } else {
            HIP_CHECK(hipEventCreate(&ev.first));
            HIP_CHECK(hipEventCreate(&ev.second));
        }
        HIP_CHECK(hipEventRecord(ev.first, c->stream));
        c->kt_work[id] += work;
        c->kt_launches[id]++;
    }
    ~KTimer() {
        if (!on) return;
        (void)hipEventRecord(ev.second, c->stream);
        c->kt_pairs[id].push_back(ev);
    }
};

// Sum block partials [item][part] and bring the item sums to the host.
// `meanwhile` (optional) is host work run after the launches, before the wait.
inline std::vector<EF> reduce_many(wl_ctx *c, const EF *parts, int n_items, int n_parts,
                                   const std::function<void()> &meanwhile = nullptr) {
    std::vector<EF> r(n_items);
    c->ensure_mapped();
    const size_t words = (size_t)n_items * 8;
    DBuf<EF> out;
    if (words > c->mapped_words) {
        out = DBuf<EF>(c->pool, n_items);
        wk::reduce_parts(parts, n_items, n_parts, out, c->stream);
        if (meanwhile) meanwhile();
        c->fetch(r.data(), out.p, words * 4);
        return r;
    }
    const uint32_t s = ++c->seq;
    if (n_items <= 16) {  // one block, a wave per item
        wk::reduce_publish(parts, n_items, n_parts, (EF *)c->mapped_dev, c->flag_dev(), s, c->stream);
    } else {
        out = DBuf<EF>(c->pool, n_items);
        wk::reduce_parts(parts, n_items, n_parts, out, c->stream);
        wk::publish((const uint32_t *)out.p, words, c->mapped_dev, c->flag_dev(), s, c->publish_counter(), c->stream);
    }
    if (meanwhile) meanwhile();
    c->wait_flag(s);
    memcpy(r.data(), c->mapped, words * 4);
    return r;
}
inline EF reduce_one(wl_ctx *c, const EF *parts, int n_parts) { return reduce_many(c, parts, 1, n_parts)[0]; }
// reduce_many in two halves: enqueue the reduction and its publish (n_items <= 4),
// then, later, wait for it.  Publishes share the mapped buffer, so a second one
// may be enqueued only behind an await_value the host releases after collecting
// the first.
// `word_off` places the result away from the start of the mapped buffer, where
// the next publishes land, so that it can be read after later ones.
inline uint32_t reduce_async(wl_ctx *c, const EF *parts, int n_items, int n_parts, size_t word_off = 0) {
    c->ensure_mapped();
    uint32_t s = ++c->seq;
    wk::reduce_publish(parts, n_items, n_parts, (EF *)(c->mapped_dev + word_off), c->flag_dev(), s, c->stream);
    return s;
}
// reduce_async whose kernel then waits for the host's next value (the returned
// slot, released after collect(seq))
inline wl_ctx::HostSlot reduce_async_await(wl_ctx *c, const EF *parts, int n_items, int n_parts, int words,
                                           uint32_t &seq) {
    c->ensure_mapped();
    wl_ctx::HostSlot sl = c->next_slot();
    seq = ++c->seq;
    size_t off = (size_t)(sl.host - c->mbox);
    wk::reduce_publish_await(parts, n_items, n_parts, (EF *)c->mapped_dev, c->flag_dev(), seq, c->mbox_dev, sl.id,
                             c->mbox_dev + off, sl.dev, words, c->await_ticks, c->mbox_dev + 1, c->stream);
    return sl;
}
inline std::vector<EF> collect(wl_ctx *c, uint32_t s, int n_items) {
    c->wait_flag(s);
    std::vector<EF> r(n_items);
    memcpy(r.data(), c->mapped, (size_t)n_items * sizeof(EF));
    return r;
}
// Several small host arrays sent with one copy: add() returns each one's word
// offset (256-byte aligned) in one pool buffer, whose device pointers at() gives
// after send().
struct Uploads {
    wl_ctx *c;
    std::vector<uint32_t> h;
    DBuf<uint32_t> d;
    explicit Uploads(wl_ctx *ctx) : c(ctx) {}
    template <class T>
    size_t add(const T *p, size_t n) {
        const size_t off = h.size(), words = (n * sizeof(T) + 3) / 4;
        h.resize(off + std::max<size_t>(64, (words + 63) / 64 * 64), 0u);
        if (n) memcpy(h.data() + off, (const void *)p, n * sizeof(T));
        return off;
    }
    template <class T>
    size_t add(const std::vector<T> &v) { return add(v.data(), v.size()); }
    void send() {
        d = DBuf<uint32_t>(c->pool, std::max<size_t>(1, h.size()));
        c->h2d_async(d.p, h.data(), h.size());
    }
    template <class T>
    T *at(size_t off) { return (T *)(d.p + off); }
};
// Stream-ordered upload (no synchronization; device buffers are recycled in
// stream order, so reuse by later work is safe).
inline DBuf<EF> upload_ef(wl_ctx *c, const std::vector<EF> &v) {
    DBuf<EF> d(c->pool, v.size() ? v.size() : 1);
    if (!v.empty()) c->h2d_async(d.p, v.data(), v.size());
    return d;
}
inline wk::EqPoint eq_point(const std::vector<EF> &pt, const EF &scale) {
    if (pt.size() > 32) throw std::runtime_error("eq point longer than 32 variables");
    wk::EqPoint e;
    memset(&e, 0, sizeof e);
    for (size_t i = 0; i < pt.size(); i++) e.p[i] = pt[i];
    e.scale = scale;
    e.n = (int)pt.size();
    return e;
}
// eq(point, .) * scale on the device (2^n EF), point passed by value
inline DBuf<EF> eq_dev(wl_ctx *c, const std::vector<EF> &pt, const EF &scale) {
    size_t N = (size_t)1 << pt.size();
    DBuf<EF> out(c->pool, N);
    wk::eq_table_v(eq_point(pt, scale), out, c->stream);
    return out;
}
inline void eq_into(wl_ctx *c, const std::vector<EF> &pt, const EF &scale, EF *out) {
    wk::eq_table_v(eq_point(pt, scale), out, c->stream);
}
// several eq tables (point, scale, destination) in one launch
struct EqJob {
    std::vector<EF> pt;
    EF scale;
    EF *out;
};
inline void eq_batch(wl_ctx *c, const std::vector<EqJob> &jobs) {
    if (jobs.empty()) return;
    if (jobs.size() == 1) { eq_into(c, jobs[0].pt, jobs[0].scale, jobs[0].out); return; }
    std::vector<wk::EqPoint> pts;
    std::vector<EF *> outs;
    int max_n = 0;
    for (const auto &j : jobs) {
        pts.push_back(eq_point(j.pt, j.scale));
        outs.push_back(j.out);
        max_n = std::max(max_n, (int)j.pt.size());
    }
    Uploads up(c);
    const size_t o_pts = up.add(pts), o_outs = up.add(outs);
    up.send();
    wk::eq_tables(up.at<wk::EqPoint>(o_pts), up.at<EF *>(o_outs), (int)jobs.size(), max_n, c->stream);
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
// the context's device copy of the constants, uploaded when they differ from the
// previous call's (proofs in a row with the same permutation skip the copy)
inline const P2Dev *context_p2(wl_ctx *c, const P2Dev &host) {
    if (!c->p2_dev) {
        HIP_CHECK(hipMalloc((void **)&c->p2_dev, sizeof(P2Dev)));
    } else if (memcmp(&c->p2_host, &host, sizeof host) == 0) {
        return c->p2_dev;
    }
    c->p2_host = host;
    c->h2d(c->p2_dev, &host, 1);
    return c->p2_dev;
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
    wk::State16 st;
    for (int i = 0; i < 16; i++) st.s[i] = ts.state[i];
    for (int i = 0; i < ts.n_in; i++) st.s[i] = ts.in_buf[i];
    int pos = ts.n_in;  // observe(w) places w here; the sample then duplexes
    uint32_t mask = (uint32_t)((1ull << bits) - 1);
    uint32_t w = 0xffffffffu;
    uint32_t *gs = c->grind_state();
    // 2^(bits+2) candidates per launch below 16 bits (a miss: chance e^-4),
    // 2^(bits+1) from 16 bits on (e^-2; 2^17 candidates fill 256 CUs with two
    // waves per SIMD), between 2^14 and 2^22
    const int lg = std::min(22, std::max(14, bits < 16 ? bits + 2 : bits + 1));
    // The first pass goes alone (it usually finds the witness); later passes go
    // out in groups of four with one publish, where a pass after the one that
    // finds the witness returns at once.
    const uint64_t chunk = (uint64_t)1 << lg;
    for (uint64_t base = 0, group = 1; base < P && w == 0xffffffffu; group = 4) {
        uint32_t s = 0;
        for (uint64_t g = 0; g < group && base < P; g++, base += chunk) {
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(chunk, P - base);
            const bool publish = g == group - 1 || base + chunk >= P;
            if (publish) s = ++c->seq;
            KTimer kt(c, KT_GRIND, (double)cnt);
            wk::grind_pass(pdev, st, pos, mask, (uint32_t)base, cnt, gs, c->mapped_dev + wl_ctx::GRIND_OFF,
                           c->flag_dev(), s, publish, c->stream);
        }
        c->wait_flag(s);
        w = c->mapped[wl_ctx::GRIND_OFF];
    }
    if (w == 0xffffffffu) throw std::runtime_error("grinding found no witness");
    ts.observe(to_mont(w));
    (void)ts.sample_bits(bits);
    ts.data.push_back(w);
}

}  // namespace wl

// per-thread message of the last failed C-ABI call
void wl_set_error(const char *msg);
