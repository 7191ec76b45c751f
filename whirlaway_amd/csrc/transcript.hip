// transcript.hip -- the Fiat-Shamir transcript handle of the C ABI: whir-p3's
// ProverState / VerifierState over DuplexChallenger<F, Poseidon2-16, 16, 8>
// (src/examples/poseidon2.rs:27,157-163, passed by the caller to
// AirTable::prove at crates/air/src/prove.rs:54 and to sumcheck::prove at
// crates/sumcheck/src/prove.rs:29).
//
// Prover: add_* appends canonical words to proof_data and observes them; hint
// appends without observing; observe only observes; grinding finds the smallest
// witness w with sample_bits(bits) == 0 after observe(w) and appends it.
// Verifier: next_* reads proof_data in the same order (values >= p are refused),
// check_pow reads a witness and checks it.  All host code; grinding may run on a
// context's device.
#include <hip/hip_runtime.h>
#include "../../include/whirlaway_hip.h"
#include "ctx.h"

using namespace wl;

namespace {

constexpr size_t STATE_WORDS = 16 + 8 + 1 + 8 + 1;

void load_p16(const uint32_t *p16c, P2Consts16 &out) {
    P2Dev tmp;
    load_p2(p16c, nullptr, tmp);
    out = tmp.p16;
}

bool read_words(wl_transcript *t, uint32_t *out, size_t n, bool observe) {
    if (t->pos + n > t->proof.size()) { wl_set_error("proof data exhausted"); return false; }
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = t->proof[t->pos + i];
        if (v >= P) { wl_set_error("proof word is not a canonical field element"); return false; }
        if (out) out[i] = v;
    }
    if (observe)
        for (size_t i = 0; i < n; i++) t->ts.observe(to_mont(t->proof[t->pos + i]));
    t->pos += n;
    return true;
}

// smallest w with sample_bits(bits) == 0 after observe(w), on the host
uint32_t grind_host(const wh::Transcript &ts, int bits) {
    wh::Transcript t(ts.p16);
    for (uint32_t w = 0; w < P; w++) {
        memcpy(t.state, ts.state, sizeof t.state);
        memcpy(t.in_buf, ts.in_buf, sizeof t.in_buf);
        memcpy(t.out_buf, ts.out_buf, sizeof t.out_buf);
        t.n_in = ts.n_in;
        t.n_out = ts.n_out;
        t.observe(to_mont(w));
        if (t.sample_bits(bits) == 0) return w;
    }
    throw std::runtime_error("grinding found no witness");
}

template <class Fn>
int guard(Fn fn) {
    try {
        return fn();
    } catch (const std::exception &e) {
        wl_set_error(e.what());
        return WL_ERR_DEVICE;
    }
}

}  // namespace

extern "C" {

wl_transcript *wl_transcript_new(const uint32_t *p16, const uint32_t *domsep, size_t domsep_len) {
    if (!p16 || (domsep_len && !domsep)) { wl_set_error("null argument"); return nullptr; }
    auto *t = new wl_transcript();
    load_p16(p16, t->p16);
    for (size_t i = 0; i < domsep_len; i++) t->ts.observe(to_mont(domsep[i] % P));
    return t;
}

wl_transcript *wl_transcript_new_verifier(const uint32_t *p16, const uint32_t *domsep, size_t domsep_len,
                                          const uint32_t *proof, size_t proof_len) {
    if (proof_len && !proof) { wl_set_error("null proof"); return nullptr; }
    wl_transcript *t = wl_transcript_new(p16, domsep, domsep_len);
    if (!t) return nullptr;
    t->verifier = true;
    t->proof.assign(proof, proof + proof_len);
    return t;
}

void wl_transcript_free(wl_transcript *t) { delete t; }

int wl_transcript_observe(wl_transcript *t, const uint32_t *v, size_t n) {
    if (!t || (n && !v)) { wl_set_error("null argument"); return WL_ERR_ARG; }
    for (size_t i = 0; i < n; i++) t->ts.observe(to_mont(v[i] % P));
    return WL_OK;
}

int wl_transcript_add_base(wl_transcript *t, const uint32_t *v, size_t n) {
    if (!t || t->verifier || (n && !v)) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    for (size_t i = 0; i < n; i++) {
        if (v[i] >= P) { wl_set_error("not a canonical field element"); return WL_ERR_ARG; }
        F x = to_mont(v[i]);
        t->ts.add_base(&x, 1);
    }
    return WL_OK;
}

int wl_transcript_add_ext(wl_transcript *t, const uint32_t *v, size_t n_ef) {
    return wl_transcript_add_base(t, v, 8 * n_ef);
}

int wl_transcript_hint(wl_transcript *t, const uint32_t *v, size_t n) {
    if (!t || t->verifier || (n && !v)) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    t->ts.hint_canonical(v, n);
    return WL_OK;
}

int wl_transcript_sample_ext(wl_transcript *t, uint32_t *out, size_t n_ef) {
    if (!t || (n_ef && !out)) { wl_set_error("null argument"); return WL_ERR_ARG; }
    for (size_t i = 0; i < n_ef; i++) wh::from_mont_ef(t->ts.sample_ef(), out + 8 * i);
    return WL_OK;
}

int wl_transcript_sample_bits(wl_transcript *t, uint32_t bits, uint32_t *out) {
    if (!t || !out || bits > 30) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    *out = t->ts.sample_bits((int)bits);
    return WL_OK;
}

int wl_transcript_grind(wl_transcript *t, wl_ctx *ctx, uint32_t bits, uint32_t *witness) {
    if (!t || t->verifier || bits > 30) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    if (bits == 0) return WL_OK;  // pow_grinding(0) sends nothing
    return guard([&]() {
        if (ctx) {
            HIP_CHECK(hipSetDevice(ctx->device));
            P2Dev ph;
            memset(&ph, 0, sizeof ph);
            ph.p16 = t->p16;
            grind(ctx, t->ts, context_p2(ctx, ph), (int)bits);
        } else {
            const uint32_t w = grind_host(t->ts, (int)bits);
            t->ts.observe(to_mont(w));
            (void)t->ts.sample_bits((int)bits);
            t->ts.data.push_back(w);
        }
        if (witness) *witness = t->ts.data.back();
        return WL_OK;
    });
}

size_t wl_transcript_proof_len(const wl_transcript *t) {
    if (!t) return 0;
    return t->verifier ? t->proof.size() : t->ts.data.size();
}

int wl_transcript_proof_data(const wl_transcript *t, uint32_t *out, size_t cap) {
    if (!t || !out) { wl_set_error("null argument"); return WL_ERR_ARG; }
    const std::vector<uint32_t> &d = t->verifier ? t->proof : t->ts.data;
    if (cap < d.size()) { wl_set_error("buffer too small"); return WL_ERR_BUFFER; }
    memcpy(out, d.data(), d.size() * 4);
    return WL_OK;
}

int wl_transcript_export(const wl_transcript *t, uint32_t *state) {
    if (!t || !state) { wl_set_error("null argument"); return WL_ERR_ARG; }
    size_t k = 0;
    for (int i = 0; i < 16; i++) state[k++] = from_mont(t->ts.state[i]);
    for (int i = 0; i < 8; i++) state[k++] = from_mont(t->ts.in_buf[i]);
    state[k++] = (uint32_t)t->ts.n_in;
    for (int i = 0; i < 8; i++) state[k++] = from_mont(t->ts.out_buf[i]);
    state[k++] = (uint32_t)t->ts.n_out;
    return k == STATE_WORDS ? WL_OK : WL_ERR_ARG;
}

int wl_transcript_import(wl_transcript *t, const uint32_t *state) {
    if (!t || !state) { wl_set_error("null argument"); return WL_ERR_ARG; }
    if (state[24] > 7 || state[33] > 8) { wl_set_error("invalid sponge state"); return WL_ERR_ARG; }
    for (size_t i = 0; i < STATE_WORDS; i++)
        if (i != 24 && i != 33 && state[i] >= P) { wl_set_error("invalid sponge state"); return WL_ERR_ARG; }
    size_t k = 0;
    for (int i = 0; i < 16; i++) t->ts.state[i] = to_mont(state[k++]);
    for (int i = 0; i < 8; i++) t->ts.in_buf[i] = to_mont(state[k++]);
    t->ts.n_in = (int)state[k++];
    for (int i = 0; i < 8; i++) t->ts.out_buf[i] = to_mont(state[k++]);
    t->ts.n_out = (int)state[k++];
    return WL_OK;
}

int wl_transcript_next_base(wl_transcript *t, uint32_t *out, size_t n) {
    if (!t || !t->verifier) { wl_set_error("a verifier transcript is required"); return WL_ERR_ARG; }
    return read_words(t, out, n, true) ? WL_OK : WL_ERR_VERIFY;
}

int wl_transcript_next_ext(wl_transcript *t, uint32_t *out, size_t n_ef) {
    return wl_transcript_next_base(t, out, 8 * n_ef);
}

int wl_transcript_next_hint(wl_transcript *t, uint32_t *out, size_t n) {
    if (!t || !t->verifier) { wl_set_error("a verifier transcript is required"); return WL_ERR_ARG; }
    return read_words(t, out, n, false) ? WL_OK : WL_ERR_VERIFY;
}

int wl_transcript_check_pow(wl_transcript *t, uint32_t bits) {
    if (!t || !t->verifier || bits > 30) { wl_set_error("bad arguments"); return WL_ERR_ARG; }
    if (bits == 0) return WL_OK;
    uint32_t w;
    if (!read_words(t, &w, 1, true)) return WL_ERR_VERIFY;
    if (t->ts.sample_bits((int)bits) != 0) { wl_set_error("invalid proof-of-work witness"); return WL_ERR_VERIFY; }
    return WL_OK;
}

size_t wl_transcript_remaining(const wl_transcript *t) {
    return t && t->verifier ? t->proof.size() - t->pos : 0;
}

}  // extern "C"
