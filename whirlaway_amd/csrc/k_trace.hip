// k_trace.hip -- synthetic benchmark inputs and Poseidon2 AIR trace generation.
//
// Inputs follow src/examples/poseidon2.rs:104-151: StdRng::seed_from_u64(seed)
// (rand 0.9 ChaCha12, PCG32 seed expansion), RoundConstants::from_rng, 2^n x 16
// inputs, then Poseidon2KoalaBear<16> and <24> new_from_rng_128 constants.  F is
// sampled as next_u32() >> 1 read as Montgomery form, rejected when >= p.
// The trace is generate_trace_rows (poseidon2.rs:122-131) computed on the GPU,
// one row per thread, written column-major straight into the packed
// polynomial of the commitment.
#include <hip/hip_runtime.h>
#include "../../include/whirlaway_hip.h"
#include "ctx.h"

namespace {
using namespace kb;

struct ChaCha12 {
    uint32_t key[8];
    uint64_t counter = 0;
    uint32_t buf[64];
    int idx = 64;

    static uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
    static void qr(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d) {
        a += b; d ^= a; d = rotl(d, 16);
        c += d; b ^= c; b = rotl(b, 12);
        a += b; d ^= a; d = rotl(d, 8);
        c += d; b ^= c; b = rotl(b, 7);
    }
    void block(uint64_t ctr, uint32_t *out) const {
        uint32_t in[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
        for (int i = 0; i < 8; i++) in[4 + i] = key[i];
        in[12] = (uint32_t)ctr; in[13] = (uint32_t)(ctr >> 32); in[14] = 0; in[15] = 0;
        uint32_t x[16];
        memcpy(x, in, sizeof x);
        for (int r = 0; r < 12; r += 2) {
            qr(x[0], x[4], x[8], x[12]); qr(x[1], x[5], x[9], x[13]);
            qr(x[2], x[6], x[10], x[14]); qr(x[3], x[7], x[11], x[15]);
            qr(x[0], x[5], x[10], x[15]); qr(x[1], x[6], x[11], x[12]);
            qr(x[2], x[7], x[8], x[13]); qr(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
    }
    explicit ChaCha12(uint64_t seed) {
        // rand_core seed_from_u64: PCG32 output words fill the 32-byte seed
        for (int i = 0; i < 8; i++) {
            seed = seed * 6364136223846793005ull + 11634580027462260723ull;
            uint32_t xs = (uint32_t)(((seed >> 18) ^ seed) >> 27);
            uint32_t rot = (uint32_t)(seed >> 59);
            key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
        }
    }
    uint32_t next_u32() {
        if (idx >= 64) {
            for (int b = 0; b < 4; b++) block(counter + b, buf + 16 * b);
            counter += 4;
            idx = 0;
        }
        return buf[idx++];
    }
    // canonical value of a uniformly sampled MontyField31
    uint32_t sample_f() {
        for (;;) {
            uint32_t u = next_u32() >> 1;
            if (u < P) return from_mont(F{u});
        }
    }
};

__global__ void __launch_bounds__(256) k_trace(const uint32_t *inputs, size_t N, const wk::AirConsts *air,
                                              F *cols) {
    size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N) return;
    F s[16];
    int col = 0;
    auto put = [&](F v) { cols[(size_t)(col++) * N + row] = v; };
    put(fone());
#pragma unroll
    for (int i = 0; i < 16; i++) { s[i] = to_mont(inputs[row * 16 + i]); put(s[i]); }
    mds_light<16>(s);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = fcube(fadd(s[i], air->rc[r * 16 + i]));
        mds_light<16>(s);
#pragma unroll
        for (int i = 0; i < 16; i++) put(s[i]);
    }
#pragma unroll 1
    for (int r = 0; r < 20; r++) {
        s[0] = fcube(fadd(s[0], air->rc[64 + r]));
        put(s[0]);
        internal_layer_fast<16>(s);
    }
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = fcube(fadd(s[i], air->rc[84 + r * 16 + i]));
        mds_light<16>(s);
#pragma unroll
        for (int i = 0; i < 16; i++) put(s[i]);
    }
}

}  // namespace

extern "C" {

int wl_p2_bench_inputs(uint64_t seed, uint32_t log_n, uint32_t *air_rc, uint32_t *inputs, uint32_t *p16,
                       uint32_t *p24) {
    if (log_n > 30) { wl_set_error("log_n too large"); return WL_ERR_ARG; }
    ChaCha12 rng(seed);
    uint32_t rc[148];
    for (int i = 0; i < 148; i++) rc[i] = rng.sample_f();
    if (air_rc) memcpy(air_rc, rc, sizeof rc);
    size_t n = (size_t)16 << log_n;
    for (size_t i = 0; i < n; i++) {
        uint32_t v = rng.sample_f();
        if (inputs) inputs[i] = v;
    }
    for (int i = 0; i < 148; i++) { uint32_t v = rng.sample_f(); if (p16) p16[i] = v; }
    for (int i = 0; i < 215; i++) { uint32_t v = rng.sample_f(); if (p24) p24[i] = v; }
    return WL_OK;
}

wl_witness *wl_air_generate_witness(wl_table *t, const uint32_t *inputs) {
    if (!t || !inputs) { wl_set_error("null argument"); return nullptr; }
    if (t->gen) { wl_set_error("trace generation is for the Poseidon2 AIR table"); return nullptr; }
    wl_ctx *c = t->ctx;
    try {
        HIP_CHECK(hipSetDevice(c->device));
        size_t N = (size_t)1 << t->log_length, V = N << 8;
        wl_witness *w = new wl_witness();
        w->ctx = c;
        w->log_length = t->log_length;
        w->packed = (kb::F *)c->pool.get(V * 4);
        wk::fill_zero(w->packed + 165 * N, (V - 165 * N) * 4, c->stream);
        wl::DBuf<uint32_t> din(c->pool, N * 16);
        HIP_CHECK(hipMemcpyAsync(din.p, inputs, N * 16 * 4, hipMemcpyHostToDevice, c->stream));
        k_trace<<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(din, N, t->air_dev, w->packed);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        return w;
    } catch (const std::exception &e) {
        wl_set_error(e.what());
        return nullptr;
    }
}

int wl_witness_download(const wl_witness *w, uint32_t *out) {
    if (!w || !out) { wl_set_error("null argument"); return WL_ERR_ARG; }
    wl_ctx *c = w->ctx;
    try {
        HIP_CHECK(hipSetDevice(c->device));
        size_t n = (size_t)165 << w->log_length;
        wl::DBuf<uint32_t> tmp(c->pool, n);
        HIP_CHECK(hipMemcpyAsync(tmp.p, w->packed, n * 4, hipMemcpyDeviceToDevice, c->stream));
        wk::convert_from_mont(tmp, n, c->stream);
        c->d2h(out, tmp.p, n);
        return WL_OK;
    } catch (const std::exception &e) {
        wl_set_error(e.what());
        return WL_ERR_DEVICE;
    }
}

}  // extern "C"
