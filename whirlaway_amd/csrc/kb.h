// kb.h -- KoalaBear field, degree-8 extension and Poseidon2, shared by host and
// device code of the MI355X prover.
//
// F: p = 2^31 - 2^24 + 1 held in Montgomery form (R = 2^32) in one 32-bit VALU
// lane; products are 64-bit (v_mad_u64_u32) and reduced by REDC, which for this
// p needs no division.  Sums of up to four products are accumulated in 64 bits
// before one reduction (redc64).
// EF: F[X]/(X^8 - 3), eight consecutive u32 (coefficient 0 first), matching
// BinomialExtensionField<KoalaBear, 8> named at src/examples/poseidon2.rs:31.
// Poseidon2: Plonky3 layer structure, S-box x^3, width 16 (20 partial rounds)
// and 24 (23 partial rounds); constants live in a P2Consts block (Montgomery
// form) read with uniform (scalar) loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KB_HD __host__ __device__ __forceinline__

namespace kb {

constexpr uint32_t P = 0x7f000001u;
constexpr uint32_t MU = 0x81000001u;          // P^-1 mod 2^32
constexpr uint32_t R2 = 0x17f7efe4u;          // 2^64 mod P (to Montgomery)
constexpr uint32_t ONE_M = 0x01fffffeu;       // 2^32 mod P (Montgomery 1)
constexpr uint32_t W_EXT = 3u;                // X^8 = W
constexpr int TWO_ADICITY = 24;               // 2^24 divides P - 1 exactly

struct F { uint32_t v; };

// Conditional corrections are written as unsigned minima: for s in [0, 2P),
// min(s, s - P) (wrapping) is s mod P; for d in (-P, P) stored wrapped,
// min(d, d + P) is d mod P.  One v_min_u32 replaces a compare and a select.
KB_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
constexpr uint32_t NEG_PINV = 0x7effffffu;    // -P^-1 mod 2^32

// Montgomery reduction in its additive form: m = x (-P^-1) mod 2^32 makes
// x + m P divisible by 2^32, and for x < 2^32 P the sum stays below
// 2^33 P < 2^64, so (x + m P) / 2^32 is the high word of one v_mad_u64_u32 and
// lies in [0, 2P).  Three instructions with the product; the 32-bit multiplies
// issue at the full VALU rate on gfx950.
KB_HD uint32_t redc_lazy(uint64_t x) {  // x < 2^32 P  ->  x / 2^32 mod P, in [0, 2P)
    uint32_t m = (uint32_t)x * NEG_PINV;
    return (uint32_t)(((uint64_t)m * P + x) >> 32);
}
KB_HD uint32_t redc(uint64_t x) {  // x < 2^32 P  ->  x / 2^32 mod P
    uint32_t r = redc_lazy(x);
    return umin(r, r - P);
}

// s mod P-preserving shrink of a 64-bit sum: hi * (2^32 mod P) + lo < 2^57.1, so
// four more products below P^2 can be added without overflow (one v_mad_u64_u32).
constexpr uint32_t C32 = 0x01fffffeu;  // 2^32 mod P
KB_HD uint64_t fold64(uint64_t s) { return (uint64_t)(uint32_t)(s >> 32) * C32 + (uint32_t)s; }
KB_HD uint32_t redc64(uint64_t x) { return redc(fold64(x)); }  // any x < 2^64 (fold64(x) < 2^57.1)

KB_HD F fadd(F a, F b) { uint32_t s = a.v + b.v; return F{umin(s, s - P)}; }
KB_HD F fsub(F a, F b) { uint32_t d = a.v - b.v; return F{umin(d, d + P)}; }
KB_HD F fneg(F a) { return F{a.v ? P - a.v : 0u}; }
KB_HD F fdbl(F a) { return fadd(a, a); }
KB_HD F fmul(F a, F b) { return F{redc((uint64_t)a.v * b.v)}; }
KB_HD F fzero() { return F{0}; }
KB_HD F fone() { return F{ONE_M}; }
KB_HD F to_mont(uint32_t canon) { return F{redc((uint64_t)canon * R2)}; }
KB_HD uint32_t from_mont(F a) { return redc((uint64_t)a.v); }
// x^3 with the square left in [0, 2P): x^2 / 2^32 < 1.5 P, so x^2 x < 1.5 P^2 < 2^32 P
KB_HD F fcube(F x) {
    uint32_t x2 = redc_lazy((uint64_t)x.v * x.v);
    uint32_t r = redc_lazy((uint64_t)x2 * x.v);
    return F{umin(r, r - P)};
}
KB_HD bool feq(F a, F b) { return a.v == b.v; }

inline F fpow(F a, uint64_t e) {
    F r = fone();
    while (e) { if (e & 1) r = fmul(r, a); a = fmul(a, a); e >>= 1; }
    return r;
}
inline F finv(F a) { return fpow(a, P - 2); }
inline F fsmall(uint32_t c) { return to_mont(c % P); }

struct EF { F c[8]; };

KB_HD EF ezero() { EF r; for (int i = 0; i < 8; i++) r.c[i] = F{0}; return r; }
KB_HD EF eone() { EF r = ezero(); r.c[0] = fone(); return r; }
KB_HD EF efrom(F a) { EF r = ezero(); r.c[0] = a; return r; }
KB_HD EF eadd(const EF &a, const EF &b) { EF r; for (int i = 0; i < 8; i++) r.c[i] = fadd(a.c[i], b.c[i]); return r; }
KB_HD EF esub(const EF &a, const EF &b) { EF r; for (int i = 0; i < 8; i++) r.c[i] = fsub(a.c[i], b.c[i]); return r; }
KB_HD EF eneg(const EF &a) { EF r; for (int i = 0; i < 8; i++) r.c[i] = fneg(a.c[i]); return r; }
KB_HD EF edbl(const EF &a) { return eadd(a, a); }
KB_HD EF emulf(const EF &a, F b) { EF r; for (int i = 0; i < 8; i++) r.c[i] = fmul(a.c[i], b); return r; }
KB_HD bool eeq(const EF &a, const EF &b) { for (int i = 0; i < 8; i++) if (a.c[i].v != b.c[i].v) return false; return true; }

KB_HD F ftimes3(F a) { return fadd(fdbl(a), a); }

// c_k = sum_{i<=k} a_i b_{k-i} + sum_{i>k} a_i (3 b_{k+8-i}); eight products
// per coefficient in one 64-bit sum (shrunk by fold64 after four), one reduction.
// b3 = 3 b, coefficient-wise (a caller multiplying by the same b twice passes it).
KB_HD EF emul_b3(const EF &a, const EF &b, const uint32_t *b3) {
    EF r;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            uint32_t bb = i <= k ? b.c[k - i].v : b3[k + 8 - i];
            s += (uint64_t)a.c[i].v * bb;
            if (i == 3) s = fold64(s);
        }
        r.c[k] = F{redc64(s)};
    }
    return r;
}
// an extension element that multiplies many others (a fold's challenge): 3c once
struct EFm { EF c; uint32_t c3[8]; };
KB_HD EFm emul_prep(const EF &c) {
    EFm m;
    m.c = c;
#pragma unroll
    for (int j = 0; j < 8; j++) m.c3[j] = ftimes3(c.c[j]).v;
    return m;
}
KB_HD EF emulm(const EF &a, const EFm &m) { return emul_b3(a, m.c, m.c3); }
KB_HD EF emul(const EF &a, const EF &b) {
    uint32_t b3[8];
#pragma unroll
    for (int j = 0; j < 8; j++) b3[j] = ftimes3(b.c[j]).v;
    EF r;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            uint32_t bb = i <= k ? b.c[k - i].v : b3[k + 8 - i];
            s += (uint64_t)a.c[i].v * bb;
            if (i == 3) s = fold64(s);
        }
        r.c[k] = F{redc64(s)};
    }
    return r;
}

KB_HD EF esqr(const EF &a) { return emul(a, a); }
// a^3 with 3a computed once for both products
KB_HD EF ecube(const EF &a) {
    uint32_t a3[8];
#pragma unroll
    for (int j = 0; j < 8; j++) a3[j] = ftimes3(a.c[j]).v;
    return emul_b3(emul_b3(a, a, a3), a, a3);
}

// ---------------------------------------------------------------- Poseidon2 --
// Constants block (Montgomery form): ext_init[4][W], ext_term[4][W], internal[RP].
struct P2Consts16 { F ext_init[4][16]; F ext_term[4][16]; F internal[20]; };
struct P2Consts24 { F ext_init[4][24]; F ext_term[4][24]; F internal[23]; };

// Internal diagonal (canonical): 16: [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8,
// 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24]; 24: [-2, 1, 2, 1/2, 3, 4, -1/2,
// -3, -4, 1/2^8, 1/4, 1/8, 1/16, 1/32, 1/64, 1/2^24, -1/2^8, -1/8, -1/16,
// -1/32, -1/64, -1/2^7, -1/2^9, -1/2^24].  Montgomery images precomputed.
struct Diag { uint32_t d16[16]; uint32_t d24[24]; };

KB_HD void mat4(F *x) {
    F t01 = fadd(x[0], x[1]), t23 = fadd(x[2], x[3]);
    F t0123 = fadd(t01, t23);
    F t01123 = fadd(t0123, x[1]);
    F t01233 = fadd(t0123, x[3]);
    F y3 = fadd(t01233, fdbl(x[0]));
    F y1 = fadd(t01123, fdbl(x[2]));
    F y0 = fadd(t01123, t01);
    F y2 = fadd(t01233, t23);
    x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
}

template <int W>
KB_HD void mds_light(F *s) {
#pragma unroll
    for (int c = 0; c < W; c += 4) mat4(s + c);
    F sums[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        F acc = s[k];
#pragma unroll
        for (int j = k + 4; j < W; j += 4) acc = fadd(acc, s[j]);
        sums[k] = acc;
    }
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = fadd(s[i], sums[i & 3]);
}

// x * 2^-k (2 <= k <= 24) without a multiplication: with x = hi 2^k + lo and
// 2^-k = -127 * 2^(24-k) mod P (since 127 * 2^24 = P - 1),
// x 2^-k = hi + P - lo * 127 * 2^(24-k), a value in [1, 2P).
// Works on Montgomery residues because scaling by a constant commutes with R.
KB_HD F fdiv2k(F x, int k) {
    uint32_t lo = x.v & ((1u << k) - 1), hi = x.v >> k;
    uint32_t t = ((lo << 7) - lo) << (24 - k);
    uint32_t r = hi + (P - t);
    return F{umin(r, r - P)};
}
KB_HD F fhalf(F x) { return F{(x.v & 1) ? (x.v + P) >> 1 : x.v >> 1}; }

// Internal diagonal entries as (sign, kind, value): kind 0 = small integer
// multiplier, kind 1 = 2^-value.  16: [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4,
// 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24]; 24: [-2, 1, 2, 1/2, 3, 4,
// -1/2, -3, -4, 1/2^8, 1/4, 1/8, 1/16, 1/32, 1/64, 1/2^24, -1/2^8, -1/8,
// -1/16, -1/32, -1/64, -1/2^7, -1/2^9, -1/2^24].
struct DiagEntry { int neg, kind, val; };
constexpr DiagEntry DIAG16_SPEC[16] = {{1, 0, 2}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 0, 3}, {0, 0, 4},
                                       {1, 1, 1}, {1, 0, 3}, {1, 0, 4}, {0, 1, 8}, {0, 1, 3}, {0, 1, 24},
                                       {1, 1, 8}, {1, 1, 3}, {1, 1, 4}, {1, 1, 24}};
constexpr DiagEntry DIAG24_SPEC[24] = {{1, 0, 2}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 0, 3}, {0, 0, 4},
                                       {1, 1, 1}, {1, 0, 3}, {1, 0, 4}, {0, 1, 8}, {0, 1, 2}, {0, 1, 3},
                                       {0, 1, 4}, {0, 1, 5}, {0, 1, 6}, {0, 1, 24}, {1, 1, 8}, {1, 1, 3},
                                       {1, 1, 4}, {1, 1, 5}, {1, 1, 6}, {1, 1, 7}, {1, 1, 9}, {1, 1, 24}};

// Montgomery image (v 2^32 mod P) of a small constant
constexpr uint32_t cmont_small(uint32_t v) { return (uint32_t)(((uint64_t)v << 32) % P); }

template <int W, int I>
KB_HD F diag_scale(F x) {  // |d_I| * x
    constexpr DiagEntry e = W == 16 ? DIAG16_SPEC[I] : DIAG24_SPEC[I];
    if constexpr (e.kind == 0) {
        if constexpr (e.val == 1) return x;
        else if constexpr (e.val == 2) return fdbl(x);
        else return fmul(x, F{cmont_small(e.val)});  // 3, 4: one product (5 instructions)
    } else if constexpr (e.val == 1) {
        return fhalf(x);
    } else {
        return fdiv2k(x, e.val);
    }
}
template <int W, int I>
KB_HD F diag_apply(F sum, F x) {  // sum + d_I * x
    constexpr DiagEntry e = W == 16 ? DIAG16_SPEC[I] : DIAG24_SPEC[I];
    F t = diag_scale<W, I>(x);
    if constexpr (e.neg) return fsub(sum, t);
    else return fadd(sum, t);
}
template <int W, int I = 0>
KB_HD void internal_apply_all(F *s, F sum) {
    if constexpr (I < W) {
        s[I] = diag_apply<W, I>(sum, s[I]);
        internal_apply_all<W, I + 1>(s, sum);
    }
}
// sum of W reduced values: pairs add without overflow (each < P < 2^31), the
// W/2 pair sums accumulate in a 32-bit word plus a carry count (an add with
// carry-out and an add of the carry per pair), then hi 2^32 + lo folds as
// lo + hi (2^32 mod P) with the carry of that addition folded once more, and two
// conditional subtractions bring the total below P.
template <int W>
KB_HD F fsum_lazy(const F *s) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < W; i += 2) {
        const uint32_t pr = s[i].v + s[i + 1].v;
        lo += pr;
        hi += lo < pr;
    }
    const uint32_t m = hi * C32;  // hi < W/2: m < 2^29
    uint32_t t = lo + m;
    t += t < m ? C32 : 0u;        // a carry: the wrapped sum is below 2^29
    t = umin(t, t - P);
    return F{umin(t, t - P)};
}
// s = (J + diag(d)) s with shifts and adds only
template <int W>
KB_HD void internal_layer_fast(F *s) {
    internal_apply_all<W>(s, fsum_lazy<W>(s));
}

// ---- internal layer of the permutation kernels (p2_permute) ----
// v mod P for v < 2^37: with q = v >> 31 (one v_alignbit_b32) and
// q P = q 2^31 - q (2^24 - 1), r = (v mod 2^31) + q (2^24 - 1) (one v_mad_u32_u24,
// q < 2^6) is congruent to v and below 2^31 + 2^30 < 2P: one subtraction.
KB_HD uint32_t fred37(uint64_t v) {
    const uint32_t q = (uint32_t)(v >> 31);
    const uint32_t r = ((uint32_t)v & 0x7fffffffu) + q * 0x00ffffffu;
    return umin(r, r - P);
}
// acc + x as x * 1 + acc: one v_mad_u64_u32 instead of an add with carry pair
KB_HD uint64_t acc64(uint64_t acc, uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t carry;
    asm("v_mad_u64_u32 %0, %1, %2, 1, %0" : "+v"(acc), "=s"(carry) : "v"(x));
    return acc;
#else
    return acc + x;
#endif
}
constexpr uint32_t HALF_P1 = (P + 1) / 2;  // 1/2 mod P
template <int W, int I>
constexpr DiagEntry diag_entry() { return W == 16 ? DIAG16_SPEC[I] : DIAG24_SPEC[I]; }
// entries 2^-k and -2^-k run lazily in [0, 2P), except 2^-k with k >= 8 (below)
template <int W, int I>
constexpr bool diag_lazy() {
    constexpr DiagEntry e = diag_entry<W, I>();
    return e.kind == 1 && e.val >= 2 && (e.neg || e.val < 8);
}
// S + d_I x, with S the canonical state sum and SP = S + P.  Every entry is fed its
// own previous output.  Lazy entries (x < 2P): fdiv2k is canonical for any
// x < 2^32, and S + t or SP - t lies in [0, 2P) with no subtraction.  Entries 2^-k
// with k >= 8 stay in [0, P + 2^23.01): x < 2^31 gives hi = x >> k < 2^23 and
// SP + hi - lo 127 2^(24-k) in (S, 2P + 2^23), one subtraction from that bound.
// The other entries take and return canonical values.
template <int W, int I>
KB_HD uint32_t diag_mix(uint32_t S, uint32_t SP, uint32_t x) {
    constexpr DiagEntry e = diag_entry<W, I>();
    if constexpr (e.kind == 0 && e.val == 1) {
        const uint32_t t = S + x;
        return umin(t, t - P);
    } else if constexpr (e.kind == 0 && e.val == 2) {
        return e.neg ? fsub(F{S}, fdbl(F{x})).v : fadd(F{S}, fdbl(F{x})).v;
    } else if constexpr (e.kind == 0) {  // 3, 4: one 64-bit multiply-add, then fred37
        const uint32_t xx = e.neg ? P - x : x;
        return fred37((uint64_t)xx * e.val + S);
    } else if constexpr (e.val == 1) {  // 1/2: x/2 or (x + P)/2 = (x >> 1) + (P + 1)/2
        const uint32_t h = (x >> 1) + ((x & 1u) ? HALF_P1 : 0u);
        const uint32_t t = e.neg ? SP - h : S + h;
        return umin(t, t - P);
    } else if constexpr (!diag_lazy<W, I>()) {
        constexpr uint32_t c = 127u << (24 - e.val);  // x 2^-k = hi - lo c mod P
        const uint32_t lo = x & ((1u << e.val) - 1u), hi = x >> e.val;
        const uint32_t t = SP + hi - lo * c;
        return umin(t, t - P);
    } else {
        const uint32_t t = fdiv2k(F{x}, e.val).v;
        return e.neg ? SP - t : S + t;
    }
}
template <int W, int I = 0>
KB_HD void diag_mix_all(F *s, uint32_t S, uint32_t SP) {
    if constexpr (I < W) {
        s[I].v = diag_mix<W, I>(S, SP, s[I].v);
        diag_mix_all<W, I + 1>(s, S, SP);
    }
}
template <int W, int I>
constexpr bool pair_ok() {
    if constexpr (I + 1 < W) return !diag_lazy<W, I>() && !diag_lazy<W, I + 1>();
    else return false;
}
// sum of the state: neighbours that are not lazy (each below 2^31) added in pairs,
// lazy entries (below 2P) alone, all accumulated in 64 bits (below 37 * 2^31 <
// 2^37).  The terms alternate between two accumulators, so the serial chain of
// 64-bit multiply-adds is half as long (one wave per SIMD on the small commits
// waits on it).
template <int W, int I = 0>
KB_HD void state_sum64_2(const F *s, uint64_t &a, uint64_t &b) {
    if constexpr (I < W) {
        if constexpr (pair_ok<W, I>()) { a = acc64(a, s[I].v + s[I + 1].v); state_sum64_2<W, I + 2>(s, b, a); }
        else { a = acc64(a, s[I].v); state_sum64_2<W, I + 1>(s, b, a); }
    }
}
// s[0], the S-box output, joins the sum last: the other terms do not wait for the cube
template <int W>
KB_HD void internal_layer_mix(F *s) {
    uint64_t a = 0, b = 0;
    state_sum64_2<W, 1>(s, a, b);
    const uint32_t S = fred37(a + b + s[0].v);
    diag_mix_all<W>(s, S, S + P);
}
// lazy entries back below P (the state leaving the partial rounds)
template <int W, int I = 0>
KB_HD void canon_lazy(F *s) {
    if constexpr (I < W) {
        if constexpr (diag_lazy<W, I>()) s[I].v = umin(s[I].v, s[I].v - P);
        canon_lazy<W, I + 1>(s);
    }
}

template <int W>
KB_HD void internal_layer(F *s, const uint32_t *diag) {
    F sum = s[0];
#pragma unroll
    for (int i = 1; i < W; i++) sum = fadd(sum, s[i]);
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = fadd(sum, fmul(s[i], F{diag[i]}));
}

// ---- full rounds with a 64-bit external layer ----
// The S-box output is left in [0, 1.78P) (no final correction) and the external
// layer adds in 64 bits without reductions (rows of M4 weigh 7, the block sums
// six blocks: outputs < 88 P < 2^38); the next S-box input folds the 64-bit value
// (plus its round constant) once with fold38, which leaves it below 1.02 P, what
// the lazy cube accepts (x^2 + m P and x2 x + m P stay below 2^64).

// x in [0, 1.02P): x^3 in Montgomery form, in [0, 1.78P)
KB_HD uint32_t fcube_lazy(uint32_t x) {
    uint32_t x2 = redc_lazy((uint64_t)x * x);
    return redc_lazy((uint64_t)x2 * x);
}
// v < 2^38, reduced to [0, 2^32 - P) (below 1.016 P): as in fred37, q = v >> 31 <
// 2^7 and r = (v mod 2^31) + q (2^24 - 1) < 2^32, then one subtraction
KB_HD uint32_t fold38(uint64_t v) {
    const uint32_t q = (uint32_t)(v >> 31);
    const uint32_t r = ((uint32_t)v & 0x7fffffffu) + q * 0x00ffffffu;
    return umin(r, r - P);
}
// the external layer on y_i < 2^32 into 64-bit sums
template <int W>
KB_HD void mds_light64(const uint32_t *y, uint64_t *o) {
#pragma unroll
    for (int c = 0; c < W; c += 4) {
        const uint64_t x0 = y[c], x1 = y[c + 1], x2 = y[c + 2], x3 = y[c + 3];
        const uint64_t t01 = x0 + x1, t23 = x2 + x3, t0123 = t01 + t23;
        const uint64_t t01123 = t0123 + x1, t01233 = t0123 + x3;
        o[c] = t01123 + t01;
        o[c + 1] = t01123 + (x2 << 1);
        o[c + 2] = t01233 + t23;
        o[c + 3] = t01233 + (x0 << 1);
    }
    uint64_t sums[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint64_t acc = o[k];
#pragma unroll
        for (int j = k + 4; j < W; j += 4) acc += o[j];
        sums[k] = acc;
    }
#pragma unroll
    for (int i = 0; i < W; i++) o[i] += sums[i & 3];
}

// Round loops are kept rolled so that each round loads only its own constants
// (uniform scalar loads); unrolling them hoists every constant of the
// permutation into scalar registers and spills.
// The initial linear layer and the four beginning full rounds; s leaves canonical.
template <int W, typename C>
KB_HD void p2_begin(F *s, const C &k) {
    uint32_t y[W];
    uint64_t o[W];
    mds_light64<W>(&s[0].v, o);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) y[i] = fcube_lazy(fold38(o[i] + k.ext_init[r][i].v));
        mds_light64<W>(y, o);
    }
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
}
// The four ending full rounds on a canonical state.
template <int W, typename C>
KB_HD void p2_end(F *s, const C &k) {
    uint32_t y[W];
    uint64_t o[W];
#pragma unroll
    for (int i = 0; i < W; i++) y[i] = fcube_lazy(fadd(s[i], k.ext_term[0][i]).v);
    mds_light64<W>(y, o);
#pragma unroll 1
    for (int r = 1; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) y[i] = fcube_lazy(fold38(o[i] + k.ext_term[r][i].v));
        mds_light64<W>(y, o);
    }
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
}

// ---- partial rounds as a convolution ----
// In the partial rounds only s_0 passes an S-box.  With S(t) the state sum after
// round t's S-box, every other element follows s_i(t+1) = S(t) + d_i s_i(t), so
//   s_i(t) = d_i^t s_i(0) + sum_{u<t} d_i^(t-1-u) S(u)            (i >= 1).
// The rounds then only carry the scalar chain
//   y = (s_0 + c_t)^3,   S(t) = y + R(t),   s_0 <- S(t) - 2y,
//   R(t+1) = sum_{i>=1} s_i(t+1) = (W-1) S(t) + A(t+1) + sum_{u<t} C(t-u) S(u),
// with A(k) = sum_{i>=1} d_i^k s_i(0) and C(k) = sum_{i>=1} d_i^k, and the final
// s_i(RP) comes from the closed form.  Every sum is a dot product of canonical
// values with Montgomery constants: 64-bit multiply-adds, a fold64 after every
// four, one reduction.
//
// Most entries d_i (i >= 1) come in pairs d, -d (DiagIdx).  For a pair (a, b) =
// (s_+(0), s_-(0)) the terms d^k a + (-d)^k b = d^k (a + (-1)^k b) take one
// product with P = a + b (k even) or M = a - b (k odd).  In the closed form, with
// E and O the sums of d^e S(u) over the even and the odd exponents e = RP-1-u,
//   s_+(RP) = E + O + h (P + M) + S(RP-1),
//   s_-(RP) = E - O + (-1)^RP h (P - M) + S(RP-1),      h = d^RP / 2,
// i.e. s_+/- = E' +/- O' + S(RP-1) with E' = E + h M, O' = O + h P for RP odd
// (E' = E + h P, O' = O + h M for RP even): both elements of a pair cost the
// products of one.
template <int W>
struct DiagIdx;
// width 16: pairs (3,6) 1/2, (4,7) 3, (5,8) 4, (9,12) 2^-8, (10,13) 1/8, (11,15) 2^-24;
// singles 1, 2, -1/16
template <>
struct DiagIdx<16> {
    static constexpr int NP = 6, NS = 3;
    KB_HD static constexpr int plus(int j) { constexpr int t[6] = {3, 4, 5, 9, 10, 11}; return t[j]; }
    KB_HD static constexpr int minus(int j) { constexpr int t[6] = {6, 7, 8, 12, 13, 15}; return t[j]; }
    KB_HD static constexpr int single(int j) { constexpr int t[3] = {1, 2, 14}; return t[j]; }
};
// width 24: pairs (3,6) 1/2, (4,7) 3, (5,8) 4, (9,16) 2^-8, (11,17) 1/8, (12,18) 1/16,
// (13,19) 1/32, (14,20) 1/64, (15,23) 2^-24; singles 1, 2, 1/4, -2^-7, -2^-9
template <>
struct DiagIdx<24> {
    static constexpr int NP = 9, NS = 5;
    KB_HD static constexpr int plus(int j) { constexpr int t[9] = {3, 4, 5, 9, 11, 12, 13, 14, 15}; return t[j]; }
    KB_HD static constexpr int minus(int j) { constexpr int t[9] = {6, 7, 8, 16, 17, 18, 19, 20, 23}; return t[j]; }
    KB_HD static constexpr int single(int j) { constexpr int t[5] = {1, 2, 10, 21, 22}; return t[j]; }
};
template <int W, int RP>
struct PartialConv {
    uint32_t dpow[RP][W - 1];  // d_i^k (Montgomery) at [k - 1][i - 1], k = 1..RP
    uint32_t csum[RP];         // C(k) (Montgomery) at [k - 1], k = 1..RP - 1
    uint32_t hpow[W];          // d^RP / 2 (Montgomery) of pair j (d = the entry of DiagIdx::plus(j))
    uint32_t wm1;              // W - 1 (Montgomery)
};
// sum_j x_j c_j for canonical x_j and Montgomery c_j: products below P^2, four of
// them below 2^64, and after a fold64 (< 2^57.1) four more still fit
struct Lazy64 {
    uint64_t s = 0;
    int n = 0;
    KB_HD void add(uint32_t x, uint32_t c) {
        if (n == 4) { s = fold64(s); n = 0; }
        s += (uint64_t)x * c;
        n++;
    }
    KB_HD F get() const { return F{redc64(s)}; }
};
template <int W, int RP, typename C, typename PC>
KB_HD void p2_partial_conv(F *s, const C &k, const PC &pc) {
    using D = DiagIdx<W>;
    uint64_t r64 = 0;
#pragma unroll
    for (int i = 1; i < W; i++) r64 += s[i].v;
    F R{fred37(r64)};
    F pa[D::NP], ma[D::NP];  // P = s_+ + s_-, M = s_+ - s_- of each pair, canonical
#pragma unroll
    for (int j = 0; j < D::NP; j++) {
        pa[j] = fadd(s[D::plus(j)], s[D::minus(j)]);
        ma[j] = fsub(s[D::plus(j)], s[D::minus(j)]);
    }
    F S[RP];
    F x0 = s[0];
#pragma unroll
    for (int t = 0; t < RP; t++) {
        const F y = fcube(fadd(x0, k.internal[t]));
        S[t] = fadd(y, R);
        x0 = fsub(S[t], fdbl(y));
        if (t + 1 < RP) {
            // A(t+1) from the entering state and the convolution of the earlier
            // sums, in one dot product
            Lazy64 a;
#pragma unroll
            for (int j = 0; j < D::NP; j++) a.add(((t + 1) & 1) ? ma[j].v : pa[j].v, pc.dpow[t][D::plus(j) - 1]);
#pragma unroll
            for (int j = 0; j < D::NS; j++) a.add(s[D::single(j)].v, pc.dpow[t][D::single(j) - 1]);
            a.add(S[t].v, pc.wm1);
#pragma unroll
            for (int u = 0; u < t; u++) a.add(S[u].v, pc.csum[t - u - 1]);
            R = a.get();
        }
    }
    s[0] = x0;
#pragma unroll
    for (int j = 0; j < D::NS; j++) {
        const int i = D::single(j);
        Lazy64 a;
        a.add(s[i].v, pc.dpow[RP - 1][i - 1]);
#pragma unroll
        for (int u = 0; u + 1 < RP; u++) a.add(S[u].v, pc.dpow[RP - 2 - u][i - 1]);
        s[i] = fadd(a.get(), S[RP - 1]);
    }
#pragma unroll
    for (int j = 0; j < D::NP; j++) {
        const int ip = D::plus(j);
        Lazy64 e, o;
#pragma unroll
        for (int u = 0; u + 1 < RP; u++) {
            const int ex = RP - 1 - u;
            if (ex & 1) o.add(S[u].v, pc.dpow[ex - 1][ip - 1]);
            else e.add(S[u].v, pc.dpow[ex - 1][ip - 1]);
        }
        e.add((RP & 1) ? ma[j].v : pa[j].v, pc.hpow[j]);
        o.add((RP & 1) ? pa[j].v : ma[j].v, pc.hpow[j]);
        const F ev = e.get(), ov = o.get();
        s[ip] = fadd(fadd(ev, ov), S[RP - 1]);
        s[D::minus(j)] = fadd(fsub(ev, ov), S[RP - 1]);
    }
}
// Poseidon2 with the partial rounds by convolution (same permutation as p2_permute)
template <int W, int RP, typename C, typename PC>
KB_HD void p2_permute_conv(F *s, const C &k, const PC &pc) {
    p2_begin<W>(s, k);
    p2_partial_conv<W, RP>(s, k, pc);
    p2_end<W>(s, k);
}

template <int W, int RP, typename C>
KB_HD void p2_permute(F *s, const C &k, const uint32_t *diag) {
    (void)diag;
    uint32_t y[W];
    uint64_t o[W];
    mds_light64<W>(&s[0].v, o);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) y[i] = fcube_lazy(fold38(o[i] + k.ext_init[r][i].v));
        mds_light64<W>(y, o);
    }
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
#pragma unroll 1
    for (int r = 0; r < RP; r++) {
        s[0] = fcube(fadd(s[0], k.internal[r]));
        internal_layer_mix<W>(s);
    }
    canon_lazy<W>(s);
#pragma unroll
    for (int i = 0; i < W; i++) y[i] = fcube_lazy(fadd(s[i], k.ext_term[0][i]).v);
    mds_light64<W>(y, o);
#pragma unroll 1
    for (int r = 1; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < W; i++) y[i] = fcube_lazy(fold38(o[i] + k.ext_term[r][i].v));
        mds_light64<W>(y, o);
    }
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
}

}  // namespace kb
