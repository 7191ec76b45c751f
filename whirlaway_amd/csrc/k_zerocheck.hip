// k_zerocheck.hip -- zerocheck rounds >= 1 (crates/sumcheck/src/prove.rs:138-188,
// skips = 1, z in {0, 2, 3}) with 64 or 256 lanes per (point, z).
//
// The Poseidon2 AIR constraints (p3-poseidon2-air, order fixed by
// crates/utils/src/constraints_folder.rs:51-56) split into independent tasks:
//   - S-boxes of the 7 full rounds whose input state is a column vector
//     (beginning rounds 0..3: ext(inputs) / previous post; ending rounds 1..3:
//     previous post),
//   - the 20 partial rounds and the input of ending round 0, whose states are
//     fixed linear combinations of columns (ChainLin), evaluated as dot products,
//   - the external MDS row and the alpha-weighted constraint of each of the
//     128 full-round outputs; the 4x4 blocks of every external layer are
//     computed once per state (32 block tasks, then rows = own block output +
//     the four blocks' sum of its residue class), not once per row.
// The folded column values of each point live in LDS, the constants (alpha
// powers, round constants, chain coefficients) are read through the caches.  The
// largest rounds (2^11 points and more) give each point 16 lanes, one state
// element per lane (k_zc_round_r16); the 2^10-point round gives each point one
// wave (4 points per block); smaller rounds, which are latency-bound, give each
// point the whole 256-lane block.
#include "common.h"
#include "kernels.h"
#include "p16_lanes.h"

namespace wk {
using namespace kb;

namespace {

constexpr int NC = 165;

// the 4x4 block of the external layer on v[0..4) in place: rows
// [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]]
__device__ __forceinline__ void m4_block(EF *v) {
    const EF t01 = eadd(v[0], v[1]), t23 = eadd(v[2], v[3]);
    const EF t0123 = eadd(t01, t23);
    const EF t01123 = eadd(t0123, v[1]), t01233 = eadd(t0123, v[3]);
    const EF y3 = eadd(t01233, edbl(v[0])), y1 = eadd(t01123, edbl(v[2]));
    v[0] = eadd(t01123, t01);
    v[1] = y1;
    v[2] = eadd(t01233, t23);
    v[3] = y3;
}
// row i of the external layer once every block of y holds its m4_block output:
// y_i plus the sum over the four blocks of residue class i mod 4
__device__ __forceinline__ EF mds_out(const EF *y, int i) {
    const int r = i & 3;
    return eadd(y[i], eadd(eadd(y[r], y[4 + r]), eadd(y[8 + r], y[12 + r])));
}

__device__ __forceinline__ EF add_rc(EF t, F rc) {
    t.c[0] = fadd(t.c[0], rc);
    return t;
}

}  // namespace

// PTS points per 256-thread block; LANES = 256 / PTS lanes per point.
template <int PTS>
__global__ void __launch_bounds__(256) k_zc_round_w(const EF *U, size_t stride, int l,
                                                   const AirConsts *air, const ChainLin *cl,
                                                   const EF *ap, const EF *eq, EF *parts) {
    constexpr int LANES = 256 / PTS;
    __shared__ EF cols[PTS][NC];
    __shared__ EF sb[PTS][128];
    __shared__ EF red[4];
    // constants (alpha powers, round constants, chain coefficients) are read through
    // the caches: keeping them out of LDS leaves room for one more block per CU
    struct { const EF *ap; const F *rc; const ChainLin &cl; } K{ap, air->rc, *cl};
    const int pt = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const size_t h = (size_t)1 << (l - 1);
    const int zi = blockIdx.y;
    const int z = zi ? zi + 1 : 0;  // 0, 2, 3
    const size_t x = (size_t)blockIdx.x * PTS + pt;
    const bool act = x < h;
    EF *cw = cols[pt];
    if (act) {
        for (int c = lane; c < NC; c += LANES) {
            const EF *col = U + (size_t)c * stride;
            EF a = col[x];
            if (z) {
                EF d = esub(col[x + h], a);
                a = eadd(a, edbl(d));
                if (z == 3) a = eadd(a, d);
            }
            cw[c] = a;
        }
    }
    __syncthreads();
    // the external layer on the 16 input columns (the state entering beginning
    // round 0): blocks first, then the rows, shared by the point's lanes; kept in
    // sb[pt][0..16), which the S-box of the same element overwrites after reading
    EF *ext_in = sb[pt];
    if (act && lane < 4) {
        EF v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = cw[1 + 4 * lane + q];
        m4_block(v);
#pragma unroll
        for (int q = 0; q < 4; q++) ext_in[4 * lane + q] = v[q];
    }
    __syncthreads();
    EF ext_row = ezero();
    if (act && lane < 16) ext_row = mds_out(ext_in, lane);
    __syncthreads();
    if (act && lane < 16) ext_in[lane] = ext_row;
    __syncthreads();
    EF acc = ezero();
    if (act) {
        // phase 1: S-boxes (112 full-round tasks) and the linear chain (20 + 16 tasks)
        auto sbox = [&](int t) {  // t < 112
            int r8 = t >> 4, i = t & 15;  // 0..3 beginning rounds, 4..6 ending rounds 1..3
            EF prev;
            F rc;
            int slot;
            if (r8 < 4) {
                prev = r8 == 0 ? ext_in[i] : cw[17 + 16 * (r8 - 1) + i];
                rc = K.rc[r8 * 16 + i];
                slot = r8;
            } else {
                int er = r8 - 3;  // 1..3
                prev = cw[101 + 16 * (er - 1) + i];
                rc = K.rc[84 + er * 16 + i];
                slot = 4 + er;
            }
            sb[pt][slot * 16 + i] = ecube(add_rc(prev, rc));
        };
        // chain task h < 36: h < 20 is partial round h (its S-box input and constraint),
        // h >= 20 the input of ending round 0, element h - 20.  One code path: 16
        // chain-input terms and 20 post_sbox terms (B[r][q] = 0 for q >= r).
        auto chain = [&](int h) {
            const F *ca = h < 20 ? K.cl.A[h] : K.cl.Ei[h - 20];
            const F *cb = h < 20 ? K.cl.B[h] : K.cl.Ep[h - 20];
            AccF d;
#pragma unroll 4
            for (int j = 0; j < 16; j++) d.add(cw[65 + j], ca[j]);
#pragma unroll 4
            for (int q = 0; q < 20; q++) d.add(cw[81 + q], cb[q]);
            EF tt = ecube(add_rc(d.get(), K.rc[h < 20 ? 64 + h : 84 + h - 20]));
            if (h < 20) acc = eadd(acc, emul(K.ap[65 + h], esub(tt, cw[81 + h])));
            else sb[pt][64 + h - 20] = tt;
        };
        if (LANES == 64) {
            // every pass runs one code path across the wave: 64 S-boxes, the 36
            // chain tasks, the remaining 48 S-boxes
            sbox(lane);
            if (lane < 36) chain(lane);
            if (lane < 48) sbox(64 + lane);
        } else {
            // 256 lanes: wave 0 the first 64 S-boxes, wave 1 the chain, wave 2 the rest
            if (lane < 64) sbox(lane);
            else if (lane < 100) chain(lane - 64);
            else if (lane >= 128 && lane < 176) sbox(lane - 64);
        }
    }
    __syncthreads();
    // phase 2: the external layer of the 8 full rounds (32 block tasks in place,
    // then 128 rows) and the full-round constraints
    if (act) {
        for (int t = lane; t < 32; t += LANES) m4_block(&sb[pt][4 * t]);
    }
    __syncthreads();
    if (act) {
        for (int t = lane; t < 128; t += LANES) {
            int r8 = t >> 4, i = t & 15;  // 0..3 beginning, 4..7 ending rounds 0..3
            EF out = mds_out(&sb[pt][r8 * 16], i);
            int post = r8 < 4 ? 17 + 16 * r8 + i : 101 + 16 * (r8 - 4) + i;
            int k = r8 < 4 ? 1 + 16 * r8 + i : 85 + 16 * (r8 - 4) + i;
            acc = eadd(acc, emul(K.ap[k], esub(out, cw[post])));
        }
        if (lane == 0) {
            EF e = cw[0];
            acc = eadd(acc, emul(K.ap[0], emul(e, esub(e, eone()))));
        }
    }
    acc = wave_sum(acc);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        EF s = ezero();
        for (int p = 0; p < PTS; p++) {
            size_t xp = (size_t)blockIdx.x * PTS + p;
            if (xp >= h) continue;
            EF v = ezero();
            for (int w = p * LANES / 64; w < (p + 1) * LANES / 64; w++) v = eadd(v, red[w]);
            s = eadd(s, emul(v, eq[xp]));
        }
        parts[(size_t)zi * gridDim.x + blockIdx.x] = s;
    }
}

// 16 lanes per (point, z), 16 points per block, for the large rounds.  Lane i of a
// point's 16-lane DPP row holds state element i: the external layers are
// ext_row16 on each of the 8 coefficients, every S-box and constraint of a full
// round is one per lane, and the 36 chain dot products (20 partial rounds, the 16
// inputs of ending round 0) are 2 or 3 per lane, reading the point's 36 chain
// inputs from LDS.  Same sum as k_zc_round_w (the constraints only add in
// another order).
__device__ __forceinline__ EF ext_row16_ef(EF v) {
#pragma unroll
    for (int c = 0; c < 8; c++) v.c[c] = ext_row16(v.c[c]);
    return v;
}
__global__ void __launch_bounds__(256) k_zc_round_r16(const EF *__restrict__ U, size_t stride, int l,
                                                      const AirConsts *__restrict__ air,
                                                      const ChainLin *__restrict__ cl, const EF *__restrict__ ap,
                                                      const EF *__restrict__ eq, EF *__restrict__ parts) {
    __shared__ EF chin[16][36];
    __shared__ EF vitem[16];
    const int item = threadIdx.x >> 4, lane = threadIdx.x & 15;
    const size_t h = (size_t)1 << (l - 1);  // a multiple of 16: every item is a point
    const int zi = blockIdx.y;
    const int z = zi ? zi + 1 : 0;  // 0, 2, 3
    const size_t x = (size_t)blockIdx.x * 16 + item;
    auto col = [&](int c) -> EF {
        const EF *p = U + (size_t)c * stride;
        EF a = p[x];
        if (z) {
            const EF d = esub(p[x + h], a);
            a = eadd(a, edbl(d));
            if (z == 3) a = eadd(a, d);
        }
        return a;
    };
    for (int t = lane; t < 36; t += 16) chin[item][t] = col(65 + t);
    // the state entering beginning round 0: the external layer of the inputs
    EF s = ext_row16_ef(col(1 + lane));
    EF acc = ezero();
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
        const EF y = ext_row16_ef(ecube(add_rc(s, air->rc[r * 16 + lane])));
        const EF post = col(17 + 16 * r + lane);
        acc = eadd(acc, emul(ap[1 + 16 * r + lane], esub(y, post)));
        s = post;
    }
    __syncthreads();
    const EF *in = chin[item];
    auto dot = [&](const F *ca, const F *cb) {
        AccF d;
#pragma unroll 4
        for (int j = 0; j < 16; j++) d.add(in[j], ca[j]);
#pragma unroll 4
        for (int q = 0; q < 20; q++) d.add(in[16 + q], cb[q]);
        return d.get();
    };
    auto partial = [&](int hh) {
        const EF tt = ecube(add_rc(dot(cl->A[hh], cl->B[hh]), air->rc[64 + hh]));
        acc = eadd(acc, emul(ap[65 + hh], esub(tt, in[16 + hh])));
    };
    partial(lane);
    if (lane < 4) partial(16 + lane);
    // ending round 0: its S-box on the state the chain leaves
    s = ecube(add_rc(dot(cl->Ei[lane], cl->Ep[lane]), air->rc[84 + lane]));
#pragma unroll 1
    for (int er = 0; er < 4; er++) {
        if (er) s = ecube(add_rc(s, air->rc[84 + er * 16 + lane]));
        const EF y = ext_row16_ef(s);
        const EF post = col(101 + 16 * er + lane);
        acc = eadd(acc, emul(ap[85 + 16 * er + lane], esub(y, post)));
        s = post;
    }
    if (lane == 0) {
        const EF e = col(0);
        acc = eadd(acc, emul(ap[0], emul(e, esub(e, eone()))));
    }
#pragma unroll
    for (int c = 0; c < 8; c++) acc.c[c] = row_sum16(acc.c[c]);
    if (lane == 0) vitem[item] = emul(acc, eq[x]);
    __syncthreads();
    if (threadIdx.x == 0) {
        EF t = ezero();
        for (int i = 0; i < 16; i++) t = eadd(t, vitem[i]);
        parts[(size_t)zi * gridDim.x + blockIdx.x] = t;
    }
}

// points per block for a round with 2^(l-1) points per z
static int pts_per_block(int l) {
    const size_t h = (size_t)1 << (l - 1);
    return h >= 2048 ? 16 : h >= 1024 ? 4 : 1;
}

int zc_round_wave_blocks(int l) {
    size_t h = (size_t)1 << (l - 1);
    int p = pts_per_block(l);
    return (int)((h + p - 1) / p);
}

int zc_round_wave(const EF *U, size_t stride, int l, const AirConsts *air, const ChainLin *cl,
                  const EF *ap, const EF *eq, EF *parts, hipStream_t s) {
    int nb = zc_round_wave_blocks(l);
    if (pts_per_block(l) == 16)
        k_zc_round_r16<<<dim3(nb, 3), 256, 0, s>>>(U, stride, l, air, cl, ap, eq, parts);
    else if (pts_per_block(l) == 4)
        k_zc_round_w<4><<<dim3(nb, 3), 256, 0, s>>>(U, stride, l, air, cl, ap, eq, parts);
    else
        k_zc_round_w<1><<<dim3(nb, 3), 256, 0, s>>>(U, stride, l, air, cl, ap, eq, parts);
    return nb;
}

}  // namespace wk
