// kernels.h -- host-side launchers for the HIP kernels of the prover.
// All launchers enqueue on the given stream and never synchronize.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "kb.h"

namespace wk {
using kb::EF;
using kb::F;

constexpr int AIR_NCOLS = 165;
constexpr int AIR_NCONS = 149;
constexpr int AIR_NRC = 4 * 16 + 20 + 4 * 16;  // beginning, partial, ending

struct AirConsts { F rc[AIR_NRC]; };  // Montgomery form, beginning[4][16], partial[20], ending[4][16]

// The 20 partial rounds overwrite state[0] with the post_sbox column and then
// apply the internal layer, so every state they produce is a fixed linear
// combination of the 16 columns entering the chain (post of beginning round 3)
// and the post_sbox columns: s0 before partial round r = A[r] . init + B[r] . ps,
// state after the chain = Ei . init + Ep . ps (Montgomery form constants).
struct ChainLin { F A[20][16]; F B[20][20]; F Ei[16][16]; F Ep[16][20]; };

// -------------------------------------------------------------- basic --
void convert_to_mont(uint32_t *d, size_t n, hipStream_t s);
void convert_from_mont(uint32_t *d, size_t n, hipStream_t s);
void fill_zero(void *d, size_t bytes, hipStream_t s);
// out[i] = sum_p in[i * n_parts + p]  (n_items outputs)
void reduce_parts(const EF *in, int n_items, int n_parts, EF *out, hipStream_t s);
// Copy n words to host-mapped memory, then (after a system-scope fence) store
// seq into *flag; the host waits on the flag instead of synchronizing the stream.
// counter: a device word at 0, used (and left at 0) when several blocks copy.
// Words [0, n_canon) are F values converted from Montgomery to canonical form.
void publish(const uint32_t *src, size_t n, uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq,
             uint32_t *counter, hipStream_t s, size_t n_canon = 0);
// rows x row_words words, stride_words apart, packed into host_dst, then flag = seq
void publish_rows(const uint32_t *src, int rows, int row_words, size_t stride_words, uint32_t *host_dst,
                  volatile uint32_t *flag, uint32_t seq, hipStream_t s);
// reduce_parts whose outputs go straight to host-mapped memory, then the flag.
void reduce_publish(const EF *in, int n_items, int n_parts, EF *host_dst, volatile uint32_t *flag,
                    uint32_t seq, hipStream_t s);

// reduce_publish, then (same launch) wait for the host's value id and copy it to
// dst, as await_host
void reduce_publish_await(const EF *in, int n_items, int n_parts, EF *host_dst, volatile uint32_t *flag,
                          uint32_t seq, const uint32_t *aflag, uint32_t id, const uint32_t *src, uint32_t *dst,
                          int words, uint64_t timeout_ticks, uint32_t *err, hipStream_t s);
// eq(point, .) over n <= 32 variables, big-endian, times scale (point by value).
struct EqPoint { EF p[32]; EF scale; int n; };
void eq_table_v(const EqPoint &pt, EF *out, hipStream_t s);
// count tables eq(pts[j], .) * scale_j into outs[j] in one launch (pts and outs in
// device memory); max_n = the largest pts[j].n
void eq_tables(const EqPoint *pts, EF *const *outs, int count, int max_n, hipStream_t s);
struct State16 { F s[16]; };

// ------------------------------------------------------------ zerocheck --
// Round 0 (univariate skip, 2^skips nodes): for each of nz evaluation points z
// (selector rows sel[nz][2^skips], F), partial sums over blocks of
//   sum_x eq[x] * AIR(fold_z(up columns))(x)
// parts: [nz][n_blocks] EF; returns n_blocks.
int zc_round0(const F *trace, int log_n, int skips, const F *sel, int nz, const AirConsts *air,
              const EF *alpha_pows, const EF *eq, EF *parts, hipStream_t s);
// The same sums for z = 2^skips .. 2^skips + nz - 1 in two passes (skips <= 4):
// the extended rows U0 (nz x 165 x M F scratch) from the selector rows sel, then
// the AIR per (x, z).  Returns n_blocks, or -1 when skips is not supported.
int zc_round0_lde(const F *trace, int log_n, int skips, int nz, const F *sel, const AirConsts *air,
                  const EF *alpha_pows, const EF *eq, F *U0, EF *parts, hipStream_t s);
// U[c][x] = sum_j S_j * up_c[j*M + x]  (S: 2^skips EF on device)
void zc_fold0(const F *trace, int log_n, int skips, const EF *S, EF *U, hipStream_t s);
// Rounds >= 1: U holds 165 EF columns (column c at U + c * stride) of current
// length 2^l.  For z in {0,2,3}: parts[zi][block] partial sums of
// eq[x] * AIR(U_z)(x), x < 2^(l-1).
// One 64-lane wave per (point, z); returns the block count (parts[zi][block]).
int zc_round_wave(const EF *U, size_t stride, int l, const AirConsts *air, const ChainLin *cl,
                  const EF *alpha_pows, const EF *eq, EF *parts, hipStream_t s);
int zc_round_wave_blocks(int l);
// In-place fold of n_cols EF columns (stride apart) from 2^l to 2^(l-1) by
// (1 - c, c), c in device memory; also halves the eq table E (length 2^(l-1))
// when E != nullptr.
void fold_cols_p(EF *U, int n_cols, size_t stride, int l, const EF *c, EF *E, hipStream_t s);
// Block the stream until the host sets *flag >= id (host-mapped, written after
// enqueue), then copy `words` words from the host-mapped src to device dst.
// Bounded: after timeout_ticks of the constant-rate clock it writes id to *err
// (host-mapped) and returns without copying.
void await_host(const uint32_t *flag, uint32_t id, const uint32_t *src, uint32_t *dst, int words,
                uint64_t timeout_ticks, uint32_t *err, hipStream_t s);
// down inner sums: parts[c][j] = sum_{x'} eqr[x'] * down_c(j*M + x'), M = 2^(log_n-skips)
void zc_down_parts(const F *trace, int log_n, int skips, const EF *eqr, EF *parts, hipStream_t s);
// out[c] = sum_j S[j] * parts[c][j]
void zc_down_combine(const EF *parts, int skips, const EF *S, EF *out, hipStream_t s);

// ------------------------------------------------------------- batching --
// B[x] = sum_{c<165} eqc[c] * col_c[x]
void lincomb_cols(const F *trace, int log_n, int n_cols, const EF *eqc, EF *B, hipStream_t s);
// mixed[x] = up(B)[x] + a * down(B)[x]
void mix_updown(const EF *B, int log_n, const EF *a, EF *mixed, hipStream_t s);
// M[y] = matrix_up_folded + a * matrix_down_folded, from E = eq(point, .)
void build_matrix(const EF *E, int log_n, const EF *a, EF *M, hipStream_t s);
// parts[j][b]: block partials of sum_{i<len} A[j*len + i] * E[i], j < nseg; returns n_blocks
int seg_dot(const EF *A, const EF *E, size_t len, int nseg, EF *parts, hipStream_t s);

// number of block partials produced by seg_dot / prod_round_fused / a2b rounds over len items
int dot_blocks(size_t len);
// number of x-chunks produced by whir_T_f
int whir_T_chunks(int m, int k);

// Product sumcheck round on tables of length 2^l (h = 2^(l-1)): parts[0][b] =
// block sums of A[x] B[x], parts[1][b] of (2A[x+h]-A[x])(2B[x+h]-B[x]); when fold
// is set the tables are first folded in place from 2^(l+1) by *c (device memory).
int prod_round_fused(EF *A, EF *B, int l, bool fold, const EF *c, EF *parts, hipStream_t s);
// fold one or two EF tables (B may be nullptr) from 2^l to 2^(l-1) by *c (device memory)
void fold_ef2_p(EF *A, EF *B, int l, const EF *c, hipStream_t s);
// in-place fold of one EF table by (1-c, c)
void fold_ef(EF *A, int l, const EF *c, hipStream_t s);
// nf folds in place of a table of 2^l <= 2^12 entries by c[0 .. nf) in one launch
void fold_ef_tail(EF *A, int l, int nf, const EF *c, hipStream_t s);

// -------------------------------------------------------------- micro --
// fold an F table (2^l) by 2^k EF scalars: out[i] = sum_j s_j in[i + j 2^(l-k)]
void fold_f_large(const F *in, int l, const EF *sc, int k, EF *out, hipStream_t s);
void fold_ef_large(const EF *in, int l, const EF *sc, int k, EF *out, hipStream_t s);
// eq-weighted a^2 b sumcheck round (degree 3, z in {0,2,3}, or {0,1,2,3} when
// all4); F inputs (base) or EF; parts [z][block]
int a2b_round_f(const F *a, const F *b, int l, const EF *eq, EF *parts, hipStream_t s, bool all4);
// a^2 b rounds >= 1 fused with the fold by the previous challenge c and the
// halving of eq: tables of length 2^l after the fold (from a0, b0 when first, in
// place in A, B otherwise), eq of length 2^l halved in place; parts [3][block].
int a2b_round_fused(const F *a0, const F *b0, EF *A, EF *B, int l, bool first, EF c, EF *eq, EF *parts,
                    hipStream_t s);
// the two above with eq(z[j+1..], x) = ehi[x >> L] elo[x mod 2^L] (no round-length
// eq table; eq is not halved); need 2^(l-1) > 2^L
int a2b_round_f_split(const F *a, const F *b, int l, int L, const EF *elo, const EF *ehi, EF *parts, hipStream_t s,
                      bool all4);
int a2b_round_fused_split(const F *a0, const F *b0, EF *A, EF *B, int l, int L, bool first, EF c, const EF *elo,
                          const EF *ehi, EF *parts, hipStream_t s);
// fold an F table by (1-c, c) into an EF table of half size
void fold_f_to_ef(const F *in, int l, const EF *c, EF *out, hipStream_t s);

// ----------------------------------------------------------------- WHIR --
struct P2Dev;  // device Poseidon2 constants (16 and 24)

// T[j][u] partial sums: sum_x f(u*2^(m-k) + x) * E_j(x), E_j EF tables (stride 2^(m-k)).
// parts layout [j][u][chunk]; returns n_chunks.  f[i] = 0 for i >= nz_len (not read).
int whir_T_f(const F *f, int m, int k, const EF *E, int nE, EF *parts, size_t nz_len, hipStream_t s);
// f'(x) = sum_u eqr[u] f(u*2^(m-k) + x)  (f[i] = 0 for i >= nz_len, not read)
void whir_fold_f(const F *f, int m, int k, const EF *eqr, EF *out, size_t nz_len, hipStream_t s);
// W[x] = sum_j scalF[j] * hiF_j[x >> lo] * loF_j[x & mask] + sum_e hiE_e[x >> lo] * loE_e[x & mask]
// (F-point tables; EF-point tables whose hi part already carries the scalar);
// accumulate: W += ..., else W = ...
void whir_build_w(EF *W, int m, int lo, int nF, const EF *scalF, const F *hiF, const F *loF, int nE,
                  const EF *hiE, const EF *loE, bool accumulate, hipStream_t s);
// eq(expand(y_j, n), .) tables over F for nY points y (device), out [j][2^n]
void eq_univ_f_tables(const F *ys, int nY, int n, F *out, hipStream_t s);
// AoS EF (2^m) -> SoA F columns: out[(b*8 + c)*2^(m-k) + i] = in[b*2^(m-k) + i].c[c]
void ef_to_cols(const EF *in, int m, int k, F *out, hipStream_t s);
// RS encode n_cols columns of 2^log_m evaluations (contiguous) into columns of
// 2^log_rows NTT values of the univariate embedding (codeword row r = value at
// w^r).  Returns la: row r is stored at position ((r mod 2^la) << (log_rows - la))
// | (r >> la) of every column (la = 0: natural order).  tmp: n_nonzero x
// 2^log_rows words; tabs: ntt_tables(log_rows) (used when ntt_split holds).
// Columns from n_nonzero on are zero: written as zeros when zero_fill is set, else
// left unwritten (the Merkle leaf and opening kernels then take them as zero).
// variant: the split of the two-pass encoding (ntt_split); tabs must be the tables
// of the same split.
int rs_encode(const F *cols, int n_cols, int n_nonzero, int log_m, int log_rows, F *cw,
              F *tmp, const F *tw_full, const F *tabs, hipStream_t s, bool zero_fill = true, int variant = 0);
// lengths up to 2^12 are encoded in one block per column, 2^14 .. 2^20 by the
// two-pass register kernels with L = LA + LB (LB = 10 by default; variants 1, 2:
// LB = 9, 8), 2^21 .. 2^24 by three register passes (pass A split in two strided
// passes, LB = 10), 2^13 by LDS four-step kernels
constexpr int NTT2_MIN_L = 14, NTT2_MAX_L = 20, NTT3_MAX_L = 24;
constexpr int NTT_VARIANTS = 3, LEAF_VARIANTS = 4, TREE_VARIANTS = 6;
bool ntt_split(int L, int *LA, int *LB, int variant = 0);
size_t ntt_tables_words(int L, int variant = 0);
void ntt_tables(const F *tw_full, int L, F *out, hipStream_t s, int variant = 0);
// twiddle table w_{2^L}^i for i < 2^L
void twiddles(F *out, int L, hipStream_t s);
// Merkle tree over rows of a column-major codeword (width F per leaf), heap layout nodes[2*rows][8]
// (la: codeword storage order as returned by rs_encode)
void merkle_build(const F *cw, int width, int log_rows, int la, const void *pc, F *nodes, hipStream_t s);
// the two halves of merkle_build: leaf hashes (columns from nnz on taken as zero,
// not read), then the compression levels
// variant: LEAF_VARIANTS / TREE_VARIANTS launch shapes (negative: the default)
void merkle_leaves(const F *cw, int width, int nnz, int log_rows, int la, const void *pc, F *nodes, hipStream_t s,
                   int variant = -1);
int merkle_leaves_default(int log_rows);
void merkle_nodes(int log_rows, const void *pc, F *nodes, hipStream_t s, int variant = -1);
// tree schedule `variant`: levels of at least 2^W parents one thread per node, the
// lowest Pn of them on lane pairs, the rest in 16-lane subtrees
void merkle_tree_schedule(int variant, int *W, int *Pn);
// gather rows (width F each; columns from nnz on are zero) and sibling paths for
// sorted query indices
void merkle_open(const F *cw, const F *nodes, int width, int nnz, int log_rows, int la, const uint32_t *idx,
                 int nq, F *rows_out, F *paths_out, hipStream_t s);
// folded query values: out[q] = sum_b eqr[b] * row_q[b], rows of 2^fold F (base)
// or 2^fold EF (8 F each, coefficient-wise) values
void fold_rows(const F *rows, int nq, int fold, bool base, const EF *eqr, EF *out, hipStream_t s);
// minimal w in [base, base + count) with P16(state with w at position pos)[7] & mask == 0,
// skipped when an earlier pass of the group found one; a publishing pass ends the group:
// the group's minimum (0xffffffff: none) to host_dst[0], then seq to *flag.  state =
// {0xffffffff, 0, -, 0xffffffff} on a group's entry and is left so at its end.
void grind_pass(const void *p16, State16 st, int pos, uint32_t mask, uint32_t base, uint32_t count,
                uint32_t *state, uint32_t *host_dst, volatile uint32_t *flag, uint32_t seq, bool publish,
                hipStream_t s);
// device Poseidon2 (tests): permute n states of width W in place
void p2_permute_batch(const void *consts, int width, F *states, int n, hipStream_t s);

}  // namespace wk
