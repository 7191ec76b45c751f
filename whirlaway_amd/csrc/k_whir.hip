// k_whir.hip -- HIP kernels for the WHIR commitment and opening: Reed-Solomon
// encoding (Mobius transform + radix-2 NTT, LDS-staged four-step), Poseidon2
// Merkle trees, proof-of-work grinding, and the WHIR sumcheck reductions.
//
// Behaviour follows the reference's WHIR call sites (crates/air/src/prove.rs:82-91,
// 216-225) with the conventions stated in oracle/whir.c.  The NTT uses
// n = N2*n1 + n2, k = k1 + N1*k2:
//   step A (kernel k_ntt_a): Mobius over the evaluation bits of n1, DFT along n1
//   step B (kernel k_ntt_b): Mobius over the evaluation bits of n2, twiddle
//                            w_N^(n2*k1), DFT along n2, store X[k1 + N1*k2]
// Mobius along one axis commutes with the DFT along the other, so the fused
// pipeline equals Mobius-then-NTT on the whole column.
#include <stdexcept>
#include "common.h"
#include "kernels.h"
#include "p16_lanes.h"

namespace wk {
using namespace kb;

struct P2DevT {
    P2Consts16 p16;
    P2Consts24 p24;
};

static inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
__device__ __forceinline__ uint32_t brev(uint32_t x, int bits) { return bits ? __brev(x) >> (32 - bits) : 0; }

// ------------------------------------------------------------------ NTT ----
__global__ void k_twiddles(F *out, int L, F w) {
    // out[i] = w^i for i < 2^L; each thread computes a chunk by repeated multiplication
    const size_t N = (size_t)1 << L;
    const size_t chunk = 64;
    size_t start = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * chunk;
    if (start >= N) return;
    // w^start by square-and-multiply
    F acc = fone(), b = w;
    size_t e = start;
    while (e) { if (e & 1) acc = fmul(acc, b); b = fmul(b, b); e >>= 1; }
    for (size_t i = start; i < start + chunk && i < N; i++) { out[i] = acc; acc = fmul(acc, w); }
}
void twiddles(F *out, int L, hipStream_t s) {
    // generator of order 2^L: 3^((p-1)/2^L); KoalaBear's two-adicity is 24, so no
    // larger power-of-two domain exists
    if (L < 0 || L > TWO_ADICITY) throw std::runtime_error("NTT domain beyond the field's two-adicity (2^24)");
    F g = fpow(fsmall(3), (uint64_t)(P - 1) >> L);
    size_t N = (size_t)1 << L;
    k_twiddles<<<nblk((N + 63) / 64), 256, 0, s>>>(out, L, g);
}

// In-LDS helpers on a tile of T rows, row stride `ld`, row length 2^lg, data held
// in bit-reversed positions.
__device__ __forceinline__ void lds_mobius(F *tile, int T, int ld, int lg, uint32_t natural_mask) {
    for (int b = 0; b < lg; b++) {
        if (!((natural_mask >> b) & 1)) continue;
        int pb = lg - 1 - b;
        int half = 1 << (lg - 1);
        for (int i = threadIdx.x; i < T * half; i += blockDim.x) {
            int t = i >> (lg - 1), j = i & (half - 1);
            int lo = ((j >> pb) << (pb + 1)) | (j & ((1 << pb) - 1));
            int hi = lo | (1 << pb);
            F *row = tile + t * ld;
            row[hi] = fsub(row[hi], row[lo]);
        }
        __syncthreads();
    }
}
__device__ __forceinline__ void lds_dit(F *tile, int T, int ld, int lg, const F *tw /* w_{2^lg}^k, k < 2^(lg-1) */) {
    int halfN = 1 << (lg - 1);
    for (int s = 1; s <= lg; s++) {
        int half = 1 << (s - 1);
        int tstride = 1 << (lg - s);
        for (int i = threadIdx.x; i < T * halfN; i += blockDim.x) {
            int t = i >> (lg - 1), j = i & (halfN - 1);
            int k = j & (half - 1);
            int i0 = ((j >> (s - 1)) << s) | k;
            int i1 = i0 + half;
            F *row = tile + t * ld;
            F u = row[i0], v = fmul(row[i1], tw[k * tstride]);
            row[i0] = fadd(u, v);
            row[i1] = fsub(u, v);
        }
        __syncthreads();
    }
}

// Whole column in LDS (2^L <= 4096): Mobius over the log_m evaluation bits, NTT.
__global__ void k_ntt_small(const F *cols, int log_m, int L, const F *tw_full, F *cw) {
    extern __shared__ F sm[];
    const int N = 1 << L;
    F *tile = sm;
    F *tw = sm + N;
    const F *col = cols + ((size_t)blockIdx.x << log_m);
    F *out = cw + ((size_t)blockIdx.x << L);
    const size_t m = (size_t)1 << log_m;
    for (int i = threadIdx.x; i < N; i += blockDim.x) tile[brev(i, L)] = (size_t)i < m ? col[i] : F{0};
    for (int i = threadIdx.x; i < N / 2; i += blockDim.x) tw[i] = tw_full[i];
    __syncthreads();
    lds_mobius(tile, 1, N, L, (uint32_t)(m - 1));
    lds_dit(tile, 1, N, L, tw);
    for (int i = threadIdx.x; i < N; i += blockDim.x) out[i] = tile[i];
}

__global__ void k_ntt_a(const F *cols, int log_m, int L1, int L2, int T, const F *tw_full, F *Y) {
    extern __shared__ F sm[];
    const int N1 = 1 << L1, N2 = 1 << L2, ld = N1 + 1;
    const int L = L1 + L2;
    F *tile = sm;
    F *tw = sm + T * ld;
    const size_t m = (size_t)1 << log_m;
    const F *col = cols + ((size_t)blockIdx.y << log_m);
    F *y = Y + ((size_t)blockIdx.y << L);
    const int n2_0 = blockIdx.x * T;
    for (int i = threadIdx.x; i < T * N1; i += blockDim.x) {
        int n1 = i / T, t = i % T;
        size_t n = (size_t)N2 * n1 + n2_0 + t;
        tile[t * ld + brev(n1, L1)] = n < m ? col[n] : F{0};
    }
    for (int i = threadIdx.x; i < N1 / 2; i += blockDim.x) tw[i] = tw_full[(size_t)i << L2];
    __syncthreads();
    // evaluation bits of n1: natural bit b of n1 is bit L2 + b of n
    uint32_t mask = 0;
    for (int b = 0; b < L1; b++) if (L2 + b < log_m) mask |= 1u << b;
    lds_mobius(tile, T, ld, L1, mask);
    lds_dit(tile, T, ld, L1, tw);
    for (int i = threadIdx.x; i < T * N1; i += blockDim.x) {
        int k1 = i / T, t = i % T;
        y[(size_t)k1 * N2 + n2_0 + t] = tile[t * ld + k1];
    }
}

__global__ void k_ntt_b(const F *Y, int log_m, int L1, int L2, int T, const F *tw_full, F *cw) {
    extern __shared__ F sm[];
    const int N1 = 1 << L1, N2 = 1 << L2, ld = N2 + 1;
    const int L = L1 + L2;
    const size_t Nmask = ((size_t)1 << L) - 1;
    F *tile = sm;
    F *tw = sm + T * ld;
    const F *y = Y + ((size_t)blockIdx.y << L);
    F *out = cw + ((size_t)blockIdx.y << L);
    const int k1_0 = blockIdx.x * T;
    for (int i = threadIdx.x; i < T * N2; i += blockDim.x) {
        int t = i / N2, n2 = i % N2;
        tile[t * ld + brev(n2, L2)] = y[(size_t)(k1_0 + t) * N2 + n2];
    }
    for (int i = threadIdx.x; i < N2 / 2; i += blockDim.x) tw[i] = tw_full[(size_t)i << L1];
    __syncthreads();
    uint32_t mask = 0;
    for (int b = 0; b < L2; b++) if (b < log_m) mask |= 1u << b;
    lds_mobius(tile, T, ld, L2, mask);
    for (int i = threadIdx.x; i < T * N2; i += blockDim.x) {
        int t = i / N2, p = i % N2;
        int n2 = brev(p, L2);
        size_t e = ((size_t)n2 * (size_t)(k1_0 + t)) & Nmask;
        tile[t * ld + p] = fmul(tile[t * ld + p], tw_full[e]);
    }
    __syncthreads();
    lds_dit(tile, T, ld, L2, tw);
    for (int i = threadIdx.x; i < T * N2; i += blockDim.x) {
        int k2 = i / T, t = i % T;
        out[(size_t)(k1_0 + t) + (size_t)N1 * k2] = tile[t * ld + k2];
    }
}

// ------------------------------------------------------- two-pass RS encoding --
// Columns of length N = 2^L, L = LA + LB, n = NB*nA + nB, k = kA + NA*kB:
//   pass A: for each nB, Mobius over the evaluation bits of nA, DFT of length NA
//           along nA -> Y[kA*NB + nB];
//   pass B: for each kA, Mobius over nB, twiddle w_N^(nB*kA), DFT of length NB
//           along nB -> X[kA + NA*kB], stored at position kA*NB + kB of the column.
// Each thread keeps 2^R elements of one row in registers and runs R radix-2
// stages on them (decimation in time on bit-reversed storage positions); rows
// move between phases through LDS.  Mobius along natural bits of a later phase
// commutes with the DIT stages of earlier phases (those act on higher natural
// bits with coefficients independent of the lower ones), so each phase first
// applies the Mobius steps of its own bits.
constexpr int NTT_TILE_A = 16384;  // words of LDS per pass-A block (64 KiB)
constexpr int NTT_TILE_B = 8192;   // words of LDS per pass-B block

__device__ __forceinline__ uint32_t brevc(uint32_t x, int bits) { return bits ? __brev(x) >> (32 - bits) : 0; }

// Radix-2 DIT stages P0+1 .. P0+R of a length-2^LG transform on a[s] = storage
// position base + (s << P0); base_lo = base mod 2^P0.  tw[i] = w_{2^LG}^i.
template <int R, int LG, int P0>
__device__ __forceinline__ void reg_dit(F *a, uint32_t base_lo, const F *tw) {
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int st = P0 + q + 1;
        const int half = 1 << q;
#pragma unroll
        for (int s = 0; s < (1 << R); s++) {
            if (s & half) continue;
            const uint32_t klo = (uint32_t)(s & (half - 1)) << P0;
            F v = a[s + half];
            if (P0 != 0 || klo != 0) v = fmul(v, tw[(base_lo | klo) << (LG - st)]);
            const F u = a[s];
            a[s] = fadd(u, v);
            a[s + half] = fsub(u, v);
        }
    }
}
// hi -= lo along every register bit q with bit q of mask set
template <int R>
__device__ __forceinline__ void reg_mobius(F *a, uint32_t mask) {
#pragma unroll
    for (int q = 0; q < R; q++) {
        if (!((mask >> q) & 1)) continue;
#pragma unroll
        for (int s = 0; s < (1 << R); s++)
            if (s & (1 << q)) a[s] = fsub(a[s], a[s ^ (1 << q)]);
    }
}

// Pass A.  Block: T = NTT_TILE_A / NA consecutive nB of one column; grid (NB / T, cols).
template <int LA>
__global__ void __launch_bounds__(512) k_ntt2_a(const F *cols, int log_m, int LB, const F *twA, F *Y) {
    constexpr int NA = 1 << LA, T = NTT_TILE_A / NA;
    constexpr int R1 = (LA + 1) / 2, R2 = LA - R1;
    __shared__ F tile[NTT_TILE_A];
    const size_t NB = (size_t)1 << LB;
    const size_t m = (size_t)1 << log_m;
    const F *col = cols + ((size_t)blockIdx.y << log_m);
    F *y = Y + ((size_t)blockIdx.y << (LA + LB));
    const size_t nB0 = (size_t)blockIdx.x * T;
    // phase 1: natural high R1 bits of nA in registers (register s <-> j = brev(s))
    uint32_t mask1 = 0;
#pragma unroll
    for (int b = 0; b < R1; b++)
        if (LB + R2 + b < log_m) mask1 |= 1u << (R1 - 1 - b);
    for (int gid = threadIdx.x; gid < (T << R2); gid += blockDim.x) {
        const int t = gid % T, g = gid / T;
        F a[1 << R1];
#pragma unroll
        for (int s = 0; s < (1 << R1); s++) {
            const size_t nA = ((size_t)brevc(s, R1) << R2) | g;
            const size_t n = NB * nA + nB0 + t;
            a[s] = n < m ? col[n] : F{0};
        }
        reg_mobius<R1>(a, mask1);
        reg_dit<R1, LA, 0>(a, 0, twA);
        const uint32_t base = brevc(g, R2) << R1;
        F *row = tile + t * NA;
#pragma unroll
        for (int s = 0; s < (1 << R1); s++) row[(base + s) ^ (t & (NA - 1))] = a[s];
    }
    __syncthreads();
    // phase 2: storage bits [R1, LA) in registers, storage low bits c fixed
    uint32_t mask2 = 0;
#pragma unroll
    for (int q = 0; q < R2; q++)
        if (LB + R2 - 1 - q < log_m) mask2 |= 1u << q;
    for (int gid = threadIdx.x; gid < (T << R1); gid += blockDim.x) {
        const int t = gid % T, c = gid / T;
        const F *row = tile + t * NA;
        F a[1 << R2];
#pragma unroll
        for (int s = 0; s < (1 << R2); s++) a[s] = row[((s << R1) | c) ^ (t & (NA - 1))];
        reg_mobius<R2>(a, mask2);
        reg_dit<R2, LA, R1>(a, (uint32_t)c, twA);
#pragma unroll
        for (int s = 0; s < (1 << R2); s++) y[(size_t)((s << R1) | c) * NB + nB0 + t] = a[s];
    }
}

// Pass B.  Block: TB = NTT_TILE_B / NB rows kA of one column; grid (NA / TB, cols).
// tl[i] = w_N^i (i < 1024), th[i] = w_N^(1024 i) (i < N / 1024), twB[i] = w_NB^i.
template <int LB>
__global__ void __launch_bounds__(256) k_ntt2_b(const F *Y, int log_m, int LA, const F *tl, const F *th,
                                               const F *twB, F *cw) {
    constexpr int NB = 1 << LB, TB = NTT_TILE_B / NB;
    constexpr int RB1 = (LB + 1) / 2, RB2 = LB - RB1;  // high / low natural bits of nB
    constexpr int LOW = (1 << RB2) - 1;
    __shared__ F tile[NTT_TILE_B];
    __shared__ F tws[NB / 2];
    const int L = LA + LB;
    const uint32_t Nmask = (1u << L) - 1;
    const size_t colo = (size_t)blockIdx.y << L;
    const F *y = Y + colo;
    F *out = cw + colo;
    const uint32_t kA0 = blockIdx.x * TB;
    for (int i = threadIdx.x; i < NB / 2; i += blockDim.x) tws[i] = twB[i];
    // LDS position of natural index nB in a row (xor swizzle against bank conflicts)
    auto pos = [](uint32_t nB) { return nB ^ ((nB >> RB2) & LOW); };
    // B1: natural low RB2 bits in registers, contiguous in global memory
    uint32_t maskl = 0;
#pragma unroll
    for (int q = 0; q < RB2; q++)
        if (q < log_m) maskl |= 1u << q;
    for (int gid = threadIdx.x; gid < (TB << RB1); gid += blockDim.x) {
        const int r = gid >> RB1, h = gid & ((1 << RB1) - 1);
        const F *src = y + (size_t)(kA0 + r) * NB + ((size_t)h << RB2);
        F a[1 << RB2];
#pragma unroll
        for (int j = 0; j < (1 << RB2); j++) a[j] = src[j];
        reg_mobius<RB2>(a, maskl);
        F *row = tile + r * NB;
#pragma unroll
        for (int j = 0; j < (1 << RB2); j++) row[pos(((uint32_t)h << RB2) | j)] = a[j];
    }
    __syncthreads();
    // B2: natural high RB1 bits in registers; Mobius, twiddle, DIT stages 1..RB1
    uint32_t maskh = 0;
#pragma unroll
    for (int b = 0; b < RB1; b++)
        if (RB2 + b < log_m) maskh |= 1u << (RB1 - 1 - b);
    for (int gid = threadIdx.x; gid < (TB << RB2); gid += blockDim.x) {
        const int r = gid >> RB2, l = gid & LOW;
        const uint32_t kA = kA0 + r;
        F *row = tile + r * NB;
        F a[1 << RB1];
#pragma unroll
        for (int s = 0; s < (1 << RB1); s++) a[s] = row[pos((brevc(s, RB1) << RB2) | l)];
        reg_mobius<RB1>(a, maskh);
        // w_N^(kA * nB) for nB = j * 2^RB2 + l: base w^(kA l), step w^(kA 2^RB2)
        const uint32_t e0 = (kA * (uint32_t)l) & Nmask, e1 = (kA << RB2) & Nmask;
        const F step = fmul(tl[e1 & 1023], th[e1 >> 10]);
        F cur = fmul(tl[e0 & 1023], th[e0 >> 10]);
#pragma unroll
        for (int j = 0; j < (1 << RB1); j++) {
            const int s = brevc(j, RB1);
            a[s] = fmul(a[s], cur);
            cur = fmul(cur, step);
        }
        reg_dit<RB1, LB, 0>(a, 0, tws);
#pragma unroll
        for (int s = 0; s < (1 << RB1); s++) row[pos((brevc(s, RB1) << RB2) | l)] = a[s];
    }
    __syncthreads();
    // B3: storage bits [RB1, LB) in registers; DIT stages RB1+1..LB; output k = storage position
    for (int gid = threadIdx.x; gid < (TB << RB1); gid += blockDim.x) {
        const int r = gid >> RB1, c = gid & ((1 << RB1) - 1);
        const F *row = tile + r * NB;
        const uint32_t hi = brevc(c, RB1) << RB2;
        F a[1 << RB2];
#pragma unroll
        for (int s = 0; s < (1 << RB2); s++) a[s] = row[pos(hi | brevc(s, RB2))];
        reg_dit<RB2, LB, RB1>(a, (uint32_t)c, tws);
        F *o = out + (size_t)(kA0 + r) * NB;
#pragma unroll
        for (int s = 0; s < (1 << RB2); s++) o[(s << RB1) | c] = a[s];
    }
}

// Middle pass of the three-pass encoding of length 2^L, L = LA1 + LM + 10 (lengths
// beyond the two-pass range).  Pass A of the two-pass scheme, a DFT of length
// NA = 2^(LA1+LM) along nA, splits into two strided passes with nA = NM*nA1 + nM
// and kA = kA1 + NA1*kM:
//   A1 (k_ntt2_a<LA1> with stride NM*NB): Mobius over nA1, DFT along nA1
//        -> Z[(kA1*NM + nM)*NB + nB];
//   M  (this kernel): Mobius over nM, twiddle w_NA^(nM*kA1), DFT of length NM
//        along nM -> Y[(kA1 + NA1*kM)*NB + nB], the layout pass B reads;
// then pass B (k_ntt2_b<10>) runs in place on Y.  The twiddle couples nM with kA1,
// so the whole Mobius over nM precedes it; the phases are those of pass B with
// rows strided by NB instead of contiguous.  Block: T = NTT_TILE_A / NM
// consecutive nB at one kA1 (consecutive threads take consecutive nB, so global
// accesses are coalesced and LDS rows, xor-swizzled by t, hit distinct banks);
// grid (NA1 * NB / T, cols).  th[i] = w_NA^i (i < NA), twM[i] = w_NM^i.
template <int LM>
__global__ void __launch_bounds__(512) k_ntt3_m(const F *Z, int log_m, int LA1, const F *twM, const F *th, F *Y) {
    constexpr int LB = 10, NB = 1 << LB;
    constexpr int NM = 1 << LM, T = NTT_TILE_A / NM, BPK = NB / T;  // blocks per kA1
    constexpr int R1 = (LM + 1) / 2, R2 = LM - R1;  // high / low natural bits of nM
    static_assert(T <= NB && NM >= 64, "tile shape");
    __shared__ F tile[NTT_TILE_A];
    const int L = LA1 + LM + LB;
    const uint32_t NAm = (1u << (LA1 + LM)) - 1;
    const uint32_t kA1 = blockIdx.x / BPK;
    const int nB0 = (blockIdx.x % BPK) * T;
    const F *z = Z + ((size_t)blockIdx.y << L) + (size_t)kA1 * NM * NB + nB0;
    F *y = Y + ((size_t)blockIdx.y << L) + nB0;
    auto at = [](int t, uint32_t x) { return t * NM + (x ^ (uint32_t)(t & (NM - 1))); };
    // M1: natural low R2 bits of nM in registers (bit q of nM is bit LB + q of n)
    uint32_t maskl = 0;
#pragma unroll
    for (int q = 0; q < R2; q++)
        if (LB + q < log_m) maskl |= 1u << q;
    for (int gid = threadIdx.x; gid < (T << R1); gid += blockDim.x) {
        const int t = gid % T, h = gid / T;
        F a[1 << R2];
#pragma unroll
        for (int j = 0; j < (1 << R2); j++) a[j] = z[(size_t)((h << R2) | j) * NB + t];
        reg_mobius<R2>(a, maskl);
#pragma unroll
        for (int j = 0; j < (1 << R2); j++) tile[at(t, (h << R2) | j)] = a[j];
    }
    __syncthreads();
    // M2: natural high R1 bits in registers; Mobius, twiddle, DIT stages 1..R1
    uint32_t maskh = 0;
#pragma unroll
    for (int b = 0; b < R1; b++)
        if (LB + R2 + b < log_m) maskh |= 1u << (R1 - 1 - b);
    const F step = th[(kA1 << R2) & NAm];  // w_NA^(kA1 * 2^R2)
    for (int gid = threadIdx.x; gid < (T << R2); gid += blockDim.x) {
        const int t = gid % T, l = gid / T;
        F a[1 << R1];
#pragma unroll
        for (int s = 0; s < (1 << R1); s++) a[s] = tile[at(t, (brevc(s, R1) << R2) | l)];
        reg_mobius<R1>(a, maskh);
        // nM = j * 2^R2 + l: w_NA^(kA1 nM) from base w_NA^(kA1 l) and the step above
        F cur = th[(kA1 * (uint32_t)l) & NAm];
#pragma unroll
        for (int j = 0; j < (1 << R1); j++) {
            const int s = brevc(j, R1);
            a[s] = fmul(a[s], cur);
            cur = fmul(cur, step);
        }
        reg_dit<R1, LM, 0>(a, 0, twM);
#pragma unroll
        for (int s = 0; s < (1 << R1); s++) tile[at(t, (brevc(s, R1) << R2) | l)] = a[s];
    }
    __syncthreads();
    // M3: storage bits [R1, LM) in registers; DIT stages R1+1..LM; kM = storage position
    for (int gid = threadIdx.x; gid < (T << R1); gid += blockDim.x) {
        const int t = gid % T, c = gid / T;
        const uint32_t hi = brevc(c, R1) << R2;
        F a[1 << R2];
#pragma unroll
        for (int s = 0; s < (1 << R2); s++) a[s] = tile[at(t, hi | brevc(s, R2))];
        reg_dit<R2, LM, R1>(a, (uint32_t)c, twM);
#pragma unroll
        for (int s = 0; s < (1 << R2); s++) {
            const uint32_t kM = ((uint32_t)s << R1) | (uint32_t)c;
            y[(size_t)(kA1 + (kM << LA1)) * NB + t] = a[s];
        }
    }
}
// the split NA = NA1 * NM of the three-pass encoding's pass A
static inline void ntt3_split(int LA, int *LA1, int *LM) {
    *LA1 = LA / 2;
    *LM = LA - *LA1;
}

// Compact twiddle tables for the two- and three-pass encodings of length 2^L:
// [twA: w_NA^i, NA/2 | twB: w_NB^i, NB/2 | tl: w_N^i, 1024 | th: w_N^(1024 i), N/1024].
// With three passes (LA1 < LA) the twA region holds twA1: w_NA1^i (NA1/2 words),
// then twM: w_NM^i (NM/2 words), then zeros.
__global__ void k_ntt_tables(const F *tw_full, int L, int LA, int LA1, F *out) {
    const int LB = L - LA, LM = LA - LA1;
    const int na = 1 << (LA - 1), nb = 1 << (LB - 1), nh = 1 << (L - 10);
    const size_t N = (size_t)1 << L;
    auto w = [&](size_t e) -> F { return tw_full[e & (N - 1)]; };
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb + 1024 + nh; i += gridDim.x * blockDim.x) {
        F v;
        if (i < na) {
            if (LA1 == LA) v = w((size_t)i << LB);
            else if (i < (1 << (LA1 - 1))) v = w((size_t)i << (L - LA1));
            else if (i < (1 << (LA1 - 1)) + (1 << (LM - 1))) v = w((size_t)(i - (1 << (LA1 - 1))) << (L - LM));
            else v = F{0};
        } else if (i < na + nb) v = w((size_t)(i - na) << LA);
        else if (i < na + nb + 1024) v = w((size_t)(i - na - nb));
        else v = w((size_t)(i - na - nb - 1024) << 10);
        out[i] = v;
    }
}
size_t ntt_tables_words(int L, int variant) {
    int LA, LB;
    ntt_split(L, &LA, &LB, variant);
    return ((size_t)1 << (LA - 1)) + ((size_t)1 << (LB - 1)) + 1024 + ((size_t)1 << (L - 10));
}
void ntt_tables(const F *tw_full, int L, F *out, hipStream_t s, int variant) {
    int LA, LB, LA1 = 0, LM = 0;
    ntt_split(L, &LA, &LB, variant);
    if (L > NTT2_MAX_L) ntt3_split(LA, &LA1, &LM);
    else LA1 = LA;
    k_ntt_tables<<<L > NTT2_MAX_L ? 64 : 8, 256, 0, s>>>(tw_full, L, LA, LA1, out);
}
// LB = min(L - 4, 10) by default (variant 0 or negative); variants 1 and 2 take
// LB = 9 and 8 where that leaves 4 <= LA <= 10.  Beyond 2^20 rows (three passes)
// LB = 10 and LA = L - 10 for every variant.
bool ntt_split(int L, int *LA, int *LB, int variant) {
    if (L < NTT2_MIN_L || L > NTT3_MAX_L) { *LA = 0; *LB = L; return false; }
    *LB = L - 4 < 10 ? L - 4 : 10;
    if (variant > 0 && variant < NTT_VARIANTS && L <= NTT2_MAX_L) {
        const int lb = 10 - variant, la = L - lb;
        if (la >= 4 && la <= 10) *LB = lb;
    }
    *LA = L - *LB;
    return true;
}

template <int LA>
static void launch_a(const F *cols, int n, int log_m, int LB, const F *twA, F *Y, hipStream_t s) {
    constexpr int T = NTT_TILE_A / (1 << LA);
    k_ntt2_a<LA><<<dim3((unsigned)(((size_t)1 << LB) / T), n), 512, 0, s>>>(cols, log_m, LB, twA, Y);
}
template <int LB>
static void launch_b(const F *Y, int n, int log_m, int LA, const F *tabs, F *cw, hipStream_t s) {
    constexpr int TB = NTT_TILE_B / (1 << LB);
    const int L = LA + LB;
    const F *twB = tabs + ((size_t)1 << (LA - 1));
    const F *tl = twB + ((size_t)1 << (LB - 1));
    const F *th = tl + 1024;
    k_ntt2_b<LB><<<dim3((unsigned)(((size_t)1 << LA) / TB), n), 256, 0, s>>>(Y, log_m, LA, tl, th, twB, cw);
    (void)L;
}

template <int LM>
static void launch_m(const F *Z, int n, int log_m, int LA1, const F *twM, const F *th, F *Y, hipStream_t s) {
    constexpr int T = NTT_TILE_A / (1 << LM);
    k_ntt3_m<LM><<<dim3((unsigned)((((size_t)1 << LA1) << 10) / T), n), 512, 0, s>>>(Z, log_m, LA1, twM, th, Y);
}
static void launch_a_any(int LA, const F *cols, int n, int log_m, int LB, const F *twA, F *Y, hipStream_t s) {
    switch (LA) {
        case 4: launch_a<4>(cols, n, log_m, LB, twA, Y, s); break;
        case 5: launch_a<5>(cols, n, log_m, LB, twA, Y, s); break;
        case 6: launch_a<6>(cols, n, log_m, LB, twA, Y, s); break;
        case 7: launch_a<7>(cols, n, log_m, LB, twA, Y, s); break;
        case 8: launch_a<8>(cols, n, log_m, LB, twA, Y, s); break;
        case 9: launch_a<9>(cols, n, log_m, LB, twA, Y, s); break;
        case 10: launch_a<10>(cols, n, log_m, LB, twA, Y, s); break;
        default: throw std::runtime_error("rs_encode: pass A length out of range");
    }
}

int rs_encode(const F *cols, int n_cols, int n_nonzero, int log_m, int L, F *cw, F *tmp,
              const F *tw_full, const F *tabs, hipStream_t s, bool zero_fill, int variant) {
    size_t rows = (size_t)1 << L;
    if (zero_fill && n_nonzero < n_cols)
        fill_zero(cw + (size_t)n_nonzero * rows, (size_t)(n_cols - n_nonzero) * rows * 4, s);
    int LA, LB;
    const bool two = ntt_split(L, &LA, &LB, variant);
    if (n_nonzero <= 0) return two ? LA : 0;
    if (L <= 12) {
        // whole column in LDS: 2^L + 2^(L-1) words (dynamic LDS stays below 48 KiB)
        size_t sh = ((size_t)1 << L) * 4 + ((size_t)1 << L) / 2 * 4 + 16;
        k_ntt_small<<<n_nonzero, 256, sh, s>>>(cols, log_m, L, tw_full, cw);
        return 0;
    }
    if (two && L > NTT2_MAX_L) {
        // three passes: A1 cols -> tmp, M tmp -> cw, B in place on cw (each block of
        // pass B reads its rows whole before it writes them)
        int LA1, LM;
        ntt3_split(LA, &LA1, &LM);
        const F *twA1 = tabs, *twM = tabs + ((size_t)1 << (LA1 - 1));
        const F *th = tabs + ((size_t)1 << (LA - 1)) + ((size_t)1 << (LB - 1)) + 1024;
        launch_a_any(LA1, cols, n_nonzero, log_m, LM + LB, twA1, tmp, s);
        switch (LM) {
            case 6: launch_m<6>(tmp, n_nonzero, log_m, LA1, twM, th, cw, s); break;
            case 7: launch_m<7>(tmp, n_nonzero, log_m, LA1, twM, th, cw, s); break;
            default: throw std::runtime_error("rs_encode: middle pass length out of range");
        }
        launch_b<10>(cw, n_nonzero, log_m, LA, tabs, cw, s);
        return LA;
    }
    if (two) {
        launch_a_any(LA, cols, n_nonzero, log_m, LB, tabs, tmp, s);
        if (LB == 8) launch_b<8>(tmp, n_nonzero, log_m, LA, tabs, cw, s);
        else if (LB == 9) launch_b<9>(tmp, n_nonzero, log_m, LA, tabs, cw, s);
        else launch_b<10>(tmp, n_nonzero, log_m, LA, tabs, cw, s);
        return LA;
    }
    int L1 = L / 2, L2 = L - L1;
    // tiles of T transforms: T * (len + 1) + len / 2 words must fit 64 KiB of LDS
    auto tile_count = [](int lg, int cap) {
        int len = 1 << lg, T = 64;
        while (T > 1 && (size_t)T * (len + 1) + len / 2 > 16384) T >>= 1;
        return T < cap ? T : cap;
    };
    int Ta = tile_count(L1, 1 << L2);
    int Tb = tile_count(L2, 1 << L1);
    size_t sha = ((size_t)Ta * ((1 << L1) + 1) + (1 << L1) / 2) * 4;
    size_t shb = ((size_t)Tb * ((1 << L2) + 1) + (1 << L2) / 2) * 4;
    k_ntt_a<<<dim3((1 << L2) / Ta, n_nonzero), 256, sha, s>>>(cols, log_m, L1, L2, Ta, tw_full, tmp);
    k_ntt_b<<<dim3((1 << L1) / Tb, n_nonzero), 256, shb, s>>>(tmp, log_m, L1, L2, Tb, tw_full, cw);
    return 0;  // natural order
}

// ---------------------------------------------------------------- Merkle ---
// Thread = storage position q of the codeword columns (coalesced reads); the row it
// holds is r = (q >> lb) | ((q mod 2^lb) << la) (natural order when la = 0).
// columns from nnz on are zero and are not read
__global__ void __launch_bounds__(256) k_merkle_leaves(const F *cw, int width, int nnz, size_t rows, int la,
                                                      int lb, const P2DevT *pc, F *nodes) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const size_t r = (q >> lb) | ((q & (((size_t)1 << lb) - 1)) << la);
    F s[24];
#pragma unroll
    for (int i = 0; i < 24; i++) s[i] = F{0};
    // the next block's 16 values are loaded before the current permutation, so
    // their memory latency hides behind it (a commit with few rows has one wave
    // per SIMD and nothing else to hide it)
    F nxt[16];
#pragma unroll
    for (int j = 0; j < 16; j++) nxt[j] = j < width && j < nnz ? cw[(size_t)j * rows + q] : F{0};
    for (int off = 0; off < width; off += 16) {
        int n = width - off < 16 ? width - off : 16;
        // compile-time state indices (a runtime index would place the state in scratch)
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < n) s[j] = nxt[j];
        const int nof = off + 16;
#pragma unroll
        for (int j = 0; j < 16; j++) nxt[j] = nof + j < width && nof + j < nnz ? cw[(size_t)(nof + j) * rows + q] : F{0};
        p24_permute(s, pc->p24);
    }
    F *o = nodes + (rows + r) * 8;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = s[i];
}

// Poseidon2-24 on a lane pair (lanes 2p and 2p + 1 of a wave), for commits with
// too few rows to give every SIMD two waves at one row per lane.  Lane h holds
// state elements 12h .. 12h + 11, the external layer's blocks 3h .. 3h + 2: the
// 4x4 blocks stay in the lane and only the four residue-class sums over the blocks
// cross lanes (one DPP swap of the 64-bit partial sums).  A partial round sums the
// lane's 12 elements, swaps the partial sum, and gives every element S + d x with
// the lane's own diagonal entry (one Montgomery product); the S-box result is kept
// by lane 0 only.  Same permutation as p2_permute<24, 23>.
__device__ __forceinline__ uint32_t pair_swap(uint32_t x) {  // DPP quad_perm [1,0,3,2]
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
}
__device__ __forceinline__ uint64_t pair_swap64(uint64_t x) {
    return (uint64_t)pair_swap((uint32_t)x) | ((uint64_t)pair_swap((uint32_t)(x >> 32)) << 32);
}
// the external layer on the lane's 12 values y < 2^32 into 64-bit sums
__device__ __forceinline__ void mds_pair(const uint32_t *y, uint64_t *o) {
#pragma unroll
    for (int c = 0; c < 12; c += 4) {
        const uint64_t x0 = y[c], x1 = y[c + 1], x2 = y[c + 2], x3 = y[c + 3];
        const uint64_t t01 = x0 + x1, t23 = x2 + x3, t0123 = t01 + t23;
        const uint64_t t01123 = t0123 + x1, t01233 = t0123 + x3;
        o[c] = t01123 + t01;
        o[c + 1] = t01123 + (x2 << 1);
        o[c + 2] = t01233 + t23;
        o[c + 3] = t01233 + (x0 << 1);
    }
    uint64_t ps[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ps[k] = o[k] + o[4 + k] + o[8 + k];
        ps[k] += pair_swap64(ps[k]);
    }
#pragma unroll
    for (int i = 0; i < 12; i++) o[i] += ps[i & 3];
}
// every lane of the wave must call this (DPP reads the partner lane)
__device__ __forceinline__ void p24_pair(F *s, bool hi, const P2Consts24 &k, const F *d) {
    uint32_t y[12];
    uint64_t o[12];
    mds_pair(&s[0].v, o);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const uint32_t c = hi ? k.ext_init[r][12 + i].v : k.ext_init[r][i].v;
            y[i] = fcube_lazy(fold38(o[i] + c));
        }
        mds_pair(y, o);
    }
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
#pragma unroll 1
    for (int r = 0; r < 23; r++) {
        const F y0 = fcube(fadd(s[0], k.internal[r]));
        if (!hi) s[0] = y0;
        uint64_t acc = (uint32_t)(s[0].v + s[1].v);
#pragma unroll
        for (int i = 2; i < 12; i += 2) acc = acc64(acc, s[i].v + s[i + 1].v);
        const uint32_t part = fred37(acc);
        const F S = fadd(F{part}, F{pair_swap(part)});
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = fadd(S, fmul(s[i], d[i]));
    }
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const uint32_t c = hi ? k.ext_term[0][12 + i].v : k.ext_term[0][i].v;
        y[i] = fcube_lazy(fadd(s[i], F{c}).v);
    }
    mds_pair(y, o);
#pragma unroll 1
    for (int r = 1; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const uint32_t c = hi ? k.ext_term[r][12 + i].v : k.ext_term[r][i].v;
            y[i] = fcube_lazy(fold38(o[i] + c));
        }
        mds_pair(y, o);
    }
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
}
// k_merkle_leaves with a lane pair per row: 2 rows threads (rows a multiple of
// 32, so every wave is full and every lane has its partner)
__global__ void __launch_bounds__(256) k_merkle_leaves_pair(const F *cw, int width, int nnz, size_t rows, int la,
                                                           int lb, const P2DevT *pc, F *nodes) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t q = g >> 1;
    const bool hi = (g & 1) != 0;
    const size_t r = (q >> lb) | ((q & (((size_t)1 << lb) - 1)) << la);
    F d[12], s[12];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        d[i] = F{hi ? DDIAG.d24[12 + i] : DDIAG.d24[i]};
        s[i] = F{0};
    }
    const int base = hi ? 12 : 0;  // absorbed positions of the lane's slots
    for (int off = 0; off < width; off += 16) {
        const int n = width - off < 16 ? width - off : 16;
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const int j = base + i;
            if (j < n) s[i] = off + j < nnz ? cw[(size_t)(off + j) * rows + q] : F{0};
        }
        p24_pair(s, hi, pc->p24, d);
    }
    if (!hi) {
        F *o = nodes + (rows + r) * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = s[i];
    }
}

// Poseidon2-16 on a lane pair, as p24_pair: lane h holds state elements
// 8h .. 8h + 7 (blocks 2h, 2h + 1); one DPP swap for the residue-class sums of the
// external layer and one for the partial rounds' state sum.  Same permutation as
// p2_permute<16, 20>; every lane of the wave must call it.
__device__ __forceinline__ void mds_pair16(const uint32_t *y, uint64_t *o) {
#pragma unroll
    for (int c = 0; c < 8; c += 4) {
        const uint64_t x0 = y[c], x1 = y[c + 1], x2 = y[c + 2], x3 = y[c + 3];
        const uint64_t t01 = x0 + x1, t23 = x2 + x3, t0123 = t01 + t23;
        const uint64_t t01123 = t0123 + x1, t01233 = t0123 + x3;
        o[c] = t01123 + t01;
        o[c + 1] = t01123 + (x2 << 1);
        o[c + 2] = t01233 + t23;
        o[c + 3] = t01233 + (x0 << 1);
    }
    uint64_t ps[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ps[k] = o[k] + o[4 + k];
        ps[k] += pair_swap64(ps[k]);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] += ps[i & 3];
}
__device__ __forceinline__ void p16_pair(F *s, bool hi, const P2Consts16 &k, const F *d) {
    uint32_t y[8];
    uint64_t o[8];
    mds_pair16(&s[0].v, o);
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t c = hi ? k.ext_init[r][8 + i].v : k.ext_init[r][i].v;
            y[i] = fcube_lazy(fold38(o[i] + c));
        }
        mds_pair16(y, o);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
#pragma unroll 1
    for (int r = 0; r < 20; r++) {
        const F y0 = fcube(fadd(s[0], k.internal[r]));
        if (!hi) s[0] = y0;
        uint64_t acc = (uint32_t)(s[0].v + s[1].v);
#pragma unroll
        for (int i = 2; i < 8; i += 2) acc = acc64(acc, s[i].v + s[i + 1].v);
        const uint32_t part = fred37(acc);
        const F S = fadd(F{part}, F{pair_swap(part)});
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = fadd(S, fmul(s[i], d[i]));
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t c = hi ? k.ext_term[0][8 + i].v : k.ext_term[0][i].v;
        y[i] = fcube_lazy(fadd(s[i], F{c}).v);
    }
    mds_pair16(y, o);
#pragma unroll 1
    for (int r = 1; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t c = hi ? k.ext_term[r][8 + i].v : k.ext_term[r][i].v;
            y[i] = fcube_lazy(fold38(o[i] + c));
        }
        mds_pair16(y, o);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t t = fold38(o[i]);
        s[i] = F{umin(t, t - P)};
    }
}
__device__ __forceinline__ void diag16_pair(bool hi, F *d) {
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = F{hi ? DDIAG.d16[8 + i] : DDIAG.d16[i]};
}
// one tree level on lane pairs (2 cnt threads, cnt a multiple of 32): lane h of
// pair i loads child 2(cnt + i) + h, lane 0 writes the parent
__global__ void __launch_bounds__(256) k_merkle_level_pair(F *nodes, size_t cnt, const P2DevT *pc) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i = cnt + (g >> 1);
    const bool hi = (g & 1) != 0;
    F d[8], s[8];
    diag16_pair(hi, d);
    const F *child = nodes + (2 * i + (hi ? 1 : 0)) * 8;
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] = child[j];
    p16_pair(s, hi, pc->p16, d);
    if (!hi) {
#pragma unroll
        for (int j = 0; j < 8; j++) nodes[i * 8 + j] = s[j];
    }
}

__device__ __forceinline__ void compress_node(F *nodes, size_t i, const P2DevT *pc) {
    F s[16];
    const F *l = nodes + 2 * i * 8;
#pragma unroll
    for (int j = 0; j < 16; j++) s[j] = l[j];
    p16_permute(s, pc->p16);
    F *o = nodes + i * 8;
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = s[j];
}

__global__ void __launch_bounds__(256) k_merkle_level(F *nodes, size_t cnt, const P2DevT *pc) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    compress_node(nodes, cnt + i, pc);
}

// Poseidon2-16 spread over one 16-lane DPP row: lane i of the row holds s[i].
// The external layer needs the quad sum and the next lane of the quad (M4 row
// r is s + x_r + 2 x_(r+1 mod 4)) and the sum over the row's four quads of each
// residue class (two row rotations); the internal layer needs the row sum
// (quad sums, then half mirror and mirror).  Cuts the dependent chain of one
// compression about 3x against one lane per permutation; used for the tree
// levels that have too few nodes to fill the GPU.
// Tree levels with parents [cnt, 2cnt), [cnt/2, cnt), ... (`levels` of them): block
// b owns the 16 parents [cnt + 16b, cnt + 16b + 16) of the first level and the
// subtree above them; group g (16 lanes) compresses parent g of each level.  The
// children of the first level come from nodes, the digests of later levels stay
// in LDS (and are written to nodes).
__global__ void __launch_bounds__(256) k_merkle_sub(F *nodes, size_t cnt, int levels, const P2DevT *pc) {
    __shared__ F kid[32 * 8], par[16 * 8];
    const int g = threadIdx.x >> 4, i = threadIdx.x & 15;
    const size_t pb = (size_t)1 << (levels - 1);  // parents per block at the first level
    const Row16 rc = row16_consts(i, pc->p16);
    for (int w = threadIdx.x; w < (int)(2 * pb * 8); w += blockDim.x)
        kid[w] = nodes[(2 * (cnt + blockIdx.x * pb)) * 8 + w];
    __syncthreads();
    for (int l = 0; l < levels; l++) {
        const int np = (int)(pb >> l);
        const size_t first = (cnt >> l) + (size_t)blockIdx.x * np;
        // DPP reads stay inside a 16-lane row, so a wave (4 groups) with no parent
        // at this level skips the permutation; in a wave that has one, every group
        // runs it and only g < np stores
        if ((g & ~3) < np) {
            const int gp = g < np ? g : 0;
            F x = kid[gp * 16 + i];
            x = p16_row(x, rc, pc->p16);
            if (g < np && i < 8) {
                par[g * 8 + i] = x;
                nodes[(first + g) * 8 + i] = x;
            }
        }
        __syncthreads();
        for (int w = threadIdx.x; w < np * 8; w += blockDim.x) kid[w] = par[w];
        __syncthreads();
    }
}

// Leaf-hashing launch variants (identical digests): 0 a lane per row in blocks of
// 256, 1 a lane pair per row in blocks of 256, 2 a lane per row in blocks of 64,
// 3 a lane pair per row in blocks of 64.  The default takes lane pairs for commits
// of 2^5 .. 2^15 rows, which at a lane per row leave SIMDs idle.
int merkle_leaves_default(int log_rows) { return log_rows >= 5 && log_rows <= 15 ? 1 : 0; }
void merkle_leaves(const F *cw, int width, int nnz, int log_rows, int la, const void *pc, F *nodes, hipStream_t s,
                   int variant) {
    const size_t rows = (size_t)1 << log_rows;
    if (variant < 0 || variant >= LEAF_VARIANTS) variant = merkle_leaves_default(log_rows);
    if ((variant & 1) && log_rows < 5) variant = 0;  // lane pairs need rows a multiple of 32
    const unsigned cap = variant >= 2 ? 64u : 256u;
    if (variant & 1) {
        const size_t threads = 2 * rows;  // a multiple of 64
        const unsigned bs = threads < cap ? (unsigned)threads : cap;
        k_merkle_leaves_pair<<<(unsigned)(threads / bs), bs, 0, s>>>(cw, width, nnz, rows, la, log_rows - la,
                                                                     (const P2DevT *)pc, nodes);
        return;
    }
    k_merkle_leaves<<<nblk(rows, (int)cap), cap, 0, s>>>(cw, width, nnz, rows, la, log_rows - la,
                                                         (const P2DevT *)pc, nodes);
}
void merkle_build(const F *cw, int width, int log_rows, int la, const void *pc, F *nodes, hipStream_t s) {
    merkle_leaves(cw, width, width, log_rows, la, pc, nodes, s);
    merkle_nodes(log_rows, pc, nodes, s);
}
// Tree schedules (identical nodes): levels of at least 2^W parents take one thread
// per node, the lowest P of them (down to 2^W) a lane pair per node instead, and
// the levels below 2^W go to 16-lane subtrees of up to five levels per block.
// (W, P) by variant: 0 (15, 1) the default, 1 (15, 0), 2 (14, 1), 3 (14, 2),
// 4 (16, 1), 5 (13, 2).
void merkle_tree_schedule(int variant, int *W, int *Pn) {
    static const int tab[TREE_VARIANTS][2] = {{15, 1}, {15, 0}, {14, 1}, {14, 2}, {16, 1}, {13, 2}};
    if (variant < 0 || variant >= TREE_VARIANTS) variant = 0;
    *W = tab[variant][0];
    *Pn = tab[variant][1];
}
void merkle_nodes(int log_rows, const void *pc, F *nodes, hipStream_t s, int variant) {
    const P2DevT *p = (const P2DevT *)pc;
    int W, Pn;
    merkle_tree_schedule(variant, &W, &Pn);
    size_t rows = (size_t)1 << log_rows;
    size_t cnt = rows >> 1;
    while (cnt >= ((size_t)1 << W)) {
        if (cnt < ((size_t)1 << (W + Pn)))
            k_merkle_level_pair<<<(unsigned)(2 * cnt / 256), 256, 0, s>>>(nodes, cnt, p);
        else
            k_merkle_level<<<nblk(cnt), 256, 0, s>>>(nodes, cnt, p);
        cnt >>= 1;
    }
    // narrow levels: subtrees of up to 5 levels per block, 16 lanes per compression
    while (cnt >= 1) {
        int lg = 0;
        while (((size_t)1 << (lg + 1)) <= cnt) lg++;
        const int levels = lg + 1 < 5 ? lg + 1 : 5;
        const size_t pb = (size_t)1 << (levels - 1);
        k_merkle_sub<<<(unsigned)(cnt / pb), 256, 0, s>>>(nodes, cnt, levels, p);
        cnt >>= levels;
    }
}

__global__ void k_merkle_open(const F *cw, const F *nodes, int width, int nnz, int log_rows, int la,
                              const uint32_t *idx, F *rows_out, F *paths_out) {
    int q = blockIdx.x;
    size_t rows = (size_t)1 << log_rows;
    uint32_t r = idx[q];
    const int lb = log_rows - la;
    const size_t at = ((size_t)(r & ((1u << la) - 1)) << lb) | (r >> la);  // storage position of row r
    for (int c = threadIdx.x; c < width; c += blockDim.x)
        rows_out[(size_t)q * width + c] = c < nnz ? cw[(size_t)c * rows + at] : F{0};
    for (int t = threadIdx.x; t < log_rows * 8; t += blockDim.x) {
        int l = t / 8, j = t % 8;
        size_t node = ((rows + r) >> l) ^ 1;
        paths_out[((size_t)q * log_rows + l) * 8 + j] = nodes[node * 8 + j];
    }
}
__global__ void k_fold_rows(const F *rows, int fold, bool base, const EF *eqr, EF *out) {
    __shared__ EF lds[16];
    const int q = blockIdx.x, K = 1 << fold;
    const int width = base ? K : 8 * K;
    EF acc = ezero();
    for (int b = threadIdx.x; b < K; b += blockDim.x) {
        const F *row = rows + (size_t)q * width;
        if (base) {
            acc = eadd(acc, emulf(eqr[b], row[b]));
        } else {
            EF v;
#pragma unroll
            for (int c = 0; c < 8; c++) v.c[c] = row[8 * b + c];
            acc = eadd(acc, emul(eqr[b], v));
        }
    }
    EF t = block_sum(acc, lds);
    if (threadIdx.x == 0) out[q] = t;
}
void fold_rows(const F *rows, int nq, int fold, bool base, const EF *eqr, EF *out, hipStream_t s) {
    if (nq > 0) k_fold_rows<<<nq, 128, 0, s>>>(rows, fold, base, eqr, out);
}

void merkle_open(const F *cw, const F *nodes, int width, int nnz, int log_rows, int la, const uint32_t *idx,
                 int nq, F *rows_out, F *paths_out, hipStream_t s) {
    if (nq > 0) k_merkle_open<<<nq, 128, 0, s>>>(cw, nodes, width, nnz, log_rows, la, idx, rows_out, paths_out);
}

// ----------------------------------------------------------------- grinding --
// One launch tests the candidates base + i, i < count (one per thread), keeping
// the smallest witness in state[0] (atomicMin); the last block to finish writes
// it (0xffffffff: none) to host memory, re-arms state = {0xffffffff, 0} and
// stores seq into *flag, so a grind costs one launch and no upload.  All the
// candidates of a launch finish together, so the launch size (not an early
// exit) sets the work: the host takes 2 to 4 x 2^bits and launches again on a
// miss.
// Grinding passes over consecutive candidate ranges run back to back, with one
// publish at the end of a group.  state = {smallest witness of the group, blocks
// done in this pass, -, smallest witness of the group's earlier passes}, armed as
// {0xffffffff, 0, -, 0xffffffff}: a pass whose group already holds a witness skips
// its permutations (every earlier range is below its own, so the minimum stands),
// and the pass's last block carries the minimum forward or, ending the group,
// publishes it and re-arms the words.
__device__ __forceinline__ bool grind_pass_live(const uint32_t *state) {
    return __hip_atomic_load(state + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0xffffffffu;
}
__device__ __forceinline__ void grind_pass_end(uint32_t *state, uint32_t *host_dst, volatile uint32_t *flag,
                                               uint32_t seq, bool publish) {
    __shared__ uint32_t last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(state + 1, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        const uint32_t m = __hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(state + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (publish) {
            host_dst[0] = m;
            __hip_atomic_store(state, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(state + 3, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence_system();
            *flag = seq;
        } else {
            __hip_atomic_store(state + 3, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
        }
    }
}
__global__ void __launch_bounds__(256) k_grind_pass(const P2DevT *pc, State16 st, int pos, uint32_t mask,
                                                   uint32_t base, uint32_t count, uint32_t *state,
                                                   uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq,
                                                   bool publish) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t w = base + i;
    if (i < count && w < P && grind_pass_live(state)) {
        F s[16];
#pragma unroll
        for (int k = 0; k < 16; k++) s[k] = st.s[k];
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (k == pos) s[k] = to_mont(w);
        p16_permute(s, pc->p16);
        if ((from_mont(s[7]) & mask) == 0) atomicMin(state, w);
    }
    grind_pass_end(state, host_dst, flag, seq, publish);
}
// k_grind_pass on lane pairs (for passes of up to 2^16 candidates, which at one
// candidate per lane leave SIMDs idle): pair i tries w = base + i; the witness
// position (below 8) and the sampled element 7 are in lane 0.  Lanes past the
// count run the permutation for their partner's sake and record nothing.
__global__ void __launch_bounds__(256) k_grind_pass_pair(const P2DevT *pc, State16 st, int pos, uint32_t mask,
                                                        uint32_t base, uint32_t count, uint32_t *state,
                                                        uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq,
                                                        bool publish) {
    if (!grind_pass_live(state)) {  // the same for every lane of the grid
        grind_pass_end(state, host_dst, flag, seq, publish);
        return;
    }
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = g >> 1, w = base + i;
    const bool hi = (g & 1) != 0;
    const bool valid = i < count && w < P;
    F d[8], s[8];
    diag16_pair(hi, d);
#pragma unroll
    for (int k = 0; k < 8; k++) s[k] = hi ? st.s[8 + k] : st.s[k];
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (!hi && k == pos) s[k] = to_mont(valid ? w : 0u);
    p16_pair(s, hi, pc->p16, d);
    if (!hi && valid && (from_mont(s[7]) & mask) == 0) atomicMin(state, w);
    grind_pass_end(state, host_dst, flag, seq, publish);
}
void grind_pass(const void *pc, State16 st, int pos, uint32_t mask, uint32_t base, uint32_t count,
                uint32_t *state, uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq, bool publish,
                hipStream_t s) {
    if (count <= (1u << 16) && pos < 8) {
        const unsigned blocks = (unsigned)((2ull * count + 255) / 256);
        k_grind_pass_pair<<<blocks, 256, 0, s>>>((const P2DevT *)pc, st, pos, mask, base, count, state, host_dst,
                                                 flag, seq, publish);
        return;
    }
    k_grind_pass<<<nblk(count), 256, 0, s>>>((const P2DevT *)pc, st, pos, mask, base, count, state, host_dst, flag,
                                            seq, publish);
}
__global__ void k_p2_batch(const P2DevT *pc, int width, F *st, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (width == 16) {
        F s[16];
        for (int j = 0; j < 16; j++) s[j] = st[(size_t)i * 16 + j];
        p16_permute(s, pc->p16);
        for (int j = 0; j < 16; j++) st[(size_t)i * 16 + j] = s[j];
    } else {
        F s[24];
        for (int j = 0; j < 24; j++) s[j] = st[(size_t)i * 24 + j];
        p24_permute(s, pc->p24);
        for (int j = 0; j < 24; j++) st[(size_t)i * 24 + j] = s[j];
    }
}
void p2_permute_batch(const void *pc, int width, F *states, int n, hipStream_t s) {
    k_p2_batch<<<nblk(n), 256, 0, s>>>((const P2DevT *)pc, width, states, n);
}

// ------------------------------------------------------- WHIR reductions ----
using AccF2 = AccF;

// T_UG rows u per block share each E[j][x] load (E is re-read once per row
// group instead of once per row); one constraint j per blockIdx.z.
constexpr int T_UG = 4;
__global__ void __launch_bounds__(256) k_whir_T_f(const F *f, int m, int k, const EF *E, int nE, EF *parts,
                                                  size_t nz_len) {
    const size_t X = (size_t)1 << (m - k);
    const int U = 1 << k;
    const int u0 = blockIdx.y * T_UG;
    const int j = blockIdx.z;
    AccF2 acc[T_UG];
    const size_t st = (size_t)gridDim.x * blockDim.x;
    const EF *Ej = E + (size_t)j * X;
    // f is zero from nz_len on: rows (and row tails) past it contribute zero sums
    size_t xend[T_UG], xmax = 0;
#pragma unroll
    for (int g = 0; g < T_UG; g++) {
        const size_t base = (size_t)(u0 + g) * X;
        xend[g] = u0 + g >= U || base >= nz_len ? 0 : (nz_len - base < X ? nz_len - base : X);
        xmax = xend[g] > xmax ? xend[g] : xmax;
    }
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < xmax; x += st) {
        const EF e = Ej[x];
        F v[T_UG];
#pragma unroll
        for (int g = 0; g < T_UG; g++) v[g] = x < xend[g] ? f[(size_t)(u0 + g) * X + x] : fzero();
#pragma unroll
        for (int g = 0; g < T_UG; g++) acc[g].add(e, v[g]);
    }
    __shared__ EF lds[16];
#pragma unroll
    for (int g = 0; g < T_UG; g++) {
        EF t = block_sum(acc[g].get(), lds);
        if (threadIdx.x == 0 && u0 + g < U) parts[((size_t)j * U + u0 + g) * gridDim.x + blockIdx.x] = t;
    }
}
int whir_T_chunks(int m, int k) {
    size_t X = (size_t)1 << (m - k);
    unsigned nc = nblk(X, 256 * 16);
    return nc < 1 ? 1 : (int)nc;
}
int whir_T_f(const F *f, int m, int k, const EF *E, int nE, EF *parts, size_t nz_len, hipStream_t s) {
    unsigned nc = (unsigned)whir_T_chunks(m, k);
    k_whir_T_f<<<dim3(nc, ((1 << k) + T_UG - 1) / T_UG, nE), 256, 0, s>>>(f, m, k, E, nE, parts, nz_len);
    return (int)nc;
}

__global__ void k_whir_fold_f(const F *f, int m, int k, const EF *eqr, EF *out, size_t nz_len) {
    const size_t X = (size_t)1 << (m - k);
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= X) return;
    AccF2 acc;
    // rows u with u X + x < nz_len, eight loads in flight per thread
    const size_t urem = x < nz_len ? (nz_len - x + X - 1) / X : 0;
    const int un = (int)(urem < ((size_t)1 << k) ? urem : ((size_t)1 << k));
    int u = 0;
    for (; u + 8 <= un; u += 8) {
        F v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = f[(size_t)(u + q) * X + x];
#pragma unroll
        for (int q = 0; q < 8; q++) acc.add(eqr[u + q], v[q]);
    }
    for (; u < un; u++) acc.add(eqr[u], f[(size_t)u * X + x]);
    out[x] = acc.get();
}
void whir_fold_f(const F *f, int m, int k, const EF *eqr, EF *out, size_t nz_len, hipStream_t s) {
    size_t X = (size_t)1 << (m - k);
    k_whir_fold_f<<<nblk(X), 256, 0, s>>>(f, m, k, eqr, out, nz_len);
}

// the hiE tables already carry their constraint's scalar
__global__ void k_build_w(EF *W, size_t Nw, int lo, int nF, const EF *scalF, const F *hiF,
                          const F *loF, int nE, const EF *hiE, const EF *loE, bool accumulate) {
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= Nw) return;
    const size_t LO = (size_t)1 << lo, HI = Nw >> lo;
    size_t xh = x >> lo, xl = x & (LO - 1);
    AccF2 acc;
    int j = 0;
    for (; j + 8 <= nF; j += 8) {  // eight load pairs in flight per thread
        F hv[8], lv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            hv[q] = hiF[(size_t)(j + q) * HI + xh];
            lv[q] = loF[(size_t)(j + q) * LO + xl];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc.add(scalF[j + q], fmul(hv[q], lv[q]));
    }
    for (; j < nF; j++) acc.add(scalF[j], fmul(hiF[(size_t)j * HI + xh], loF[(size_t)j * LO + xl]));
    EF r = acc.get();
    for (int e = 0; e < nE; e++) r = eadd(r, emul(hiE[(size_t)e * HI + xh], loE[(size_t)e * LO + xl]));
    W[x] = accumulate ? eadd(W[x], r) : r;
}
void whir_build_w(EF *W, int m, int lo, int nF, const EF *scalF, const F *hiF, const F *loF, int nE,
                  const EF *hiE, const EF *loE, bool accumulate, hipStream_t s) {
    size_t Nw = (size_t)1 << m;
    k_build_w<<<nblk(Nw), 256, 0, s>>>(W, Nw, lo, nF, scalF, hiF, loF, nE, hiE, loE, accumulate);
}

__global__ void k_eq_univ_f(const F *ys, int nY, int n, F *out) {
    size_t N = (size_t)1 << n;
    size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (x >= N || j >= nY) return;
    F y = ys[j];
    F acc = fone();
    for (int b = 0; b < n; b++) {
        acc = fmul(acc, ((x >> b) & 1) ? y : fsub(fone(), y));
        y = fmul(y, y);
    }
    out[(size_t)j * N + x] = acc;
}
void eq_univ_f_tables(const F *ys, int nY, int n, F *out, hipStream_t s) {
    if (nY <= 0) return;
    size_t N = (size_t)1 << n;
    k_eq_univ_f<<<dim3(nblk(N), nY), 256, 0, s>>>(ys, nY, n, out);
}

__global__ void k_ef_to_cols(const EF *in, size_t X, int K, F *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    if (i >= X) return;
    EF v = in[(size_t)b * X + i];
#pragma unroll
    for (int c = 0; c < 8; c++) out[((size_t)b * 8 + c) * X + i] = v.c[c];
}
void ef_to_cols(const EF *in, int m, int k, F *out, hipStream_t s) {
    size_t X = (size_t)1 << (m - k);
    k_ef_to_cols<<<dim3(nblk(X), 1 << k), 256, 0, s>>>(in, X, 1 << k, out);
}

}  // namespace wk
