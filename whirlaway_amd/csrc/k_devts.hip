// k_devts.hip -- the duplex-sponge transcript on the device (one wave).
//
// The sponge lives in LDS while a round kernel runs: lane 0 writes observed
// values, every lane keeps the same input/output counts (the control flow is
// wave-uniform), and a duplex permutes the state held by lanes 0..15 with the
// 16-lane Poseidon2-16 (p16_lanes.h; all 64 lanes execute it, as its DPP
// operations need).  observe / sample follow DuplexChallenger: observe clears the
// output and duplexes when 8 inputs are buffered; sample duplexes when inputs are
// pending or the output is used up and pops the last output element.
#include "common.h"
#include "devts.h"
#include "p16_lanes.h"

namespace wk {
using namespace kb;

namespace {

struct P2DevT {
    P2Consts16 p16;
    P2Consts24 p24;
};

struct WaveSponge {
    F *st, *inb, *outb;  // LDS
    int n_in, n_out;     // wave-uniform
    Row16 rc;
    const P2Consts16 *k;

    __device__ void load(const DevSponge &g) {
        const int t = threadIdx.x;
        if (t < 16) st[t] = g.st[t];
        if (t < 8) { inb[t] = g.inb[t]; outb[t] = g.outb[t]; }
        n_in = (int)g.n_in;
        n_out = (int)g.n_out;
        __syncthreads();
    }
    __device__ void store(DevSponge &g) {
        const int t = threadIdx.x;
        __syncthreads();
        if (t < 16) g.st[t] = st[t];
        if (t < 8) { g.inb[t] = inb[t]; g.outb[t] = outb[t]; }
        if (t == 0) { g.n_in = (uint32_t)n_in; g.n_out = (uint32_t)n_out; }
    }
    __device__ void duplex() {
        const int t = threadIdx.x;
        __syncthreads();
        if (t < n_in) st[t] = inb[t];
        __syncthreads();
        F x = st[t & 15];
        x = p16_row(x, rc, *k);
        __syncthreads();
        if (t < 16) st[t] = x;
        if (t < 8) outb[t] = x;
        n_in = 0;
        n_out = 8;
        __syncthreads();
    }
    __device__ void observe(F v) {
        n_out = 0;
        if (threadIdx.x == 0) inb[n_in] = v;
        n_in++;
        if (n_in == 8) duplex();
    }
    __device__ F sample() {
        if (n_in > 0 || n_out == 0) duplex();
        n_out--;
        return outb[n_out];
    }
};

// the sum over nb block partials of item `it`, in every lane
__device__ __forceinline__ EF wave_total(const EF *parts, int nb, int it) {
    EF acc = ezero();
    for (int i = threadIdx.x; i < nb; i += 64) acc = eadd(acc, parts[(size_t)it * nb + i]);
    return wave_sum(acc);
}

__global__ void __launch_bounds__(64) k_zc_ts_round(const EF *parts, int nb, ZcRounds *z, int j, Basis4 basis,
                                                    const P2DevT *pc) {
    __shared__ F st[16], inb[8], outb[8];
    const int t = threadIdx.x;
    WaveSponge sp{st, inb, outb, 0, 0, row16_consts(t & 15, pc->p16), &pc->p16};
    sp.load(z->sp);
    const EF missing = z->missing, sum = z->sum, r = z->r[j], rinv = z->rinv[j];
    EF e[4];
    e[0] = emul(wave_total(parts, nb, 0), missing);
    e[2] = emul(wave_total(parts, nb, 1), missing);
    e[3] = emul(wave_total(parts, nb, 2), missing);
    const EF one = eone(), omr = esub(one, r);
    e[1] = emul(esub(sum, emul(e[0], omr)), rinv);
    // coefficients of the cubic through (i, e_i), then the product with (1 - r) + (2r - 1) X
    EF co[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        EF acc = ezero();
#pragma unroll
        for (int i = 0; i < 4; i++) acc = eadd(acc, emulf(e[i], basis.b[4 * i + d]));
        co[d] = acc;
    }
    const EF b = esub(edbl(r), one);
    EF p[5];
    p[0] = emul(co[0], omr);
#pragma unroll
    for (int d = 1; d < 4; d++) p[d] = eadd(emul(co[d], omr), emul(co[d - 1], b));
    p[4] = emul(co[3], b);
    uint32_t *words = z->words + 40 * (j - 1);
#pragma unroll
    for (int d = 0; d < 5; d++)
        for (int c = 0; c < 8; c++) {
            if (t == 0) words[8 * d + c] = from_mont(p[d].c[c]);
            sp.observe(p[d].c[c]);
        }
    EF ch;
    for (int c = 0; c < 8; c++) ch.c[c] = sp.sample();
    EF s = ezero();  // p(ch) by Horner
    for (int d = 4; d >= 0; d--) s = eadd(emul(s, ch), p[d]);
    const EF nm = emul(eadd(emul(omr, esub(one, ch)), emul(r, ch)), missing);
    sp.store(z->sp);
    if (t == 0) {
        z->sum = s;
        z->missing = nm;
        z->chal[j] = ch;
    }
}

}  // namespace

void zc_ts_round(const EF *parts, int nb, ZcRounds *z, int j, const Basis4 &basis, const void *pc, hipStream_t s) {
    k_zc_ts_round<<<1, 64, 0, s>>>(parts, nb, z, j, basis, (const P2DevT *)pc);
}

}  // namespace wk
