/*
 * oracle.h -- CPU restatement of the Whirlaway prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in whirlaway_amd/: only the test suite, the smoke entry and
 * the CPU-baseline leg of the benchmark may load it.  The product never links,
 * loads or calls anything in this directory.
 *
 * Everything here is plain C11 (+ optional OpenMP), canonical field
 * representation (values in [0, p)), written for clarity.  Each function cites
 * the reference file:line whose behaviour it restates (paths relative to the
 * reference repository root).  Where the behaviour lives in a dependency that
 * is absent from the reference snapshot (Plonky3 @ d0c4a36, whir-p3 @ 9caef91)
 * the restatement follows the dependency's published algorithm and says
 * "parity unpinned" -- see the per-function comments.
 */
#ifndef WL_ORACLE_H
#define WL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- field -- */
/* KoalaBear: p = 2^31 - 2^24 + 1 (Plonky3 p3-koala-bear, MontyField31). */
#define OR_P 0x7f000001u
#define OR_EF_D 8
#define OR_EF_W 3u /* EF = F[X]/(X^8 - 3) (BinomialExtensionField<KoalaBear, 8>) */

typedef uint32_t f_t;
typedef struct { f_t c[OR_EF_D]; } ef_t;

static inline f_t f_add(f_t a, f_t b) { uint32_t s = a + b; return s >= OR_P ? s - OR_P : s; }
static inline f_t f_sub(f_t a, f_t b) { return a >= b ? a - b : a + OR_P - b; }
static inline f_t f_neg(f_t a) { return a ? OR_P - a : 0; }
static inline f_t f_mul(f_t a, f_t b) { return (f_t)(((uint64_t)a * b) % OR_P); }
f_t f_pow(f_t a, uint64_t e);
f_t f_inv(f_t a);
f_t f_from_monty(uint32_t m);     /* m * 2^-32 mod p */
f_t f_two_adic_generator(int bits);

ef_t ef_zero(void);
ef_t ef_one(void);
ef_t ef_from_f(f_t a);
ef_t ef_add(ef_t a, ef_t b);
ef_t ef_sub(ef_t a, ef_t b);
ef_t ef_neg(ef_t a);
ef_t ef_mul(ef_t a, ef_t b);
ef_t ef_mul_f(ef_t a, f_t b);
ef_t ef_square(ef_t a);
ef_t ef_inv(ef_t a);
ef_t ef_pow(ef_t a, uint64_t e);
int ef_eq(ef_t a, ef_t b);
int ef_is_zero(ef_t a);

/* ------------------------------------------------------------------ rng -- */
/* rand 0.9 StdRng = ChaCha12, seeded by rand_core's seed_from_u64 (PCG32
 * expansion).  Parity unpinned (rand_chacha 0.9.0 / rand_core 0.9.3 absent);
 * the block function is pinned by the RFC 7539 ChaCha20 vectors. */
typedef struct {
    uint32_t key[8];
    uint64_t counter;
    uint32_t buf[64];
    int idx;
    int rounds;
} or_rng_t;

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                     int rounds, uint32_t out[16]);
void or_chacha_block_ietf(const uint32_t key[8], uint32_t counter,
                          const uint32_t nonce[3], int rounds, uint32_t out[16]);
void or_rng_seed_u64(or_rng_t *r, uint64_t seed);
uint32_t or_rng_next_u32(or_rng_t *r);
f_t or_rng_sample_f(or_rng_t *r);

/* ------------------------------------------------------------ poseidon2 -- */
/* Poseidon2 over KoalaBear, S-box x^3, 8 full rounds, 20 (width 16) / 23
 * (width 24) partial rounds, Plonky3 layer structure (p3-poseidon2). */
typedef struct {
    int width, rounds_f, rounds_p;
    f_t ext_init[4][24];
    f_t ext_term[4][24];
    f_t internal[23];
} or_p2_t;

void or_p2_init_from_rng(or_p2_t *p, int width, or_rng_t *rng);
void or_p2_permute(const or_p2_t *p, f_t *state);
void or_p2_mds_light(f_t *state, int width);
void or_p2_internal_layer(f_t *state, int width);
const f_t *or_p2_internal_diag(int width);
/* Merkle primitives: PaddingFreeSponge<P24,24,16,8>, TruncatedPermutation<P16,2,8,16> */
void or_hash_leaf(const or_p2_t *p24, const f_t *row, size_t len, f_t out[8]);
void or_compress(const or_p2_t *p16, const f_t l[8], const f_t r[8], f_t out[8]);

/* ------------------------------------------------------------ transcript -- */
/* DuplexChallenger<F, Poseidon16, 16, 8> (p3-challenger) + whir-p3
 * ProverState/VerifierState.  The proof is the Vec<F> proof_data
 * (src/examples/poseidon2.rs:178,191). */
typedef struct {
    const or_p2_t *perm;
    f_t state[16];
    f_t in_buf[8]; int n_in;
    f_t out_buf[8]; int n_out;
} or_chal_t;

typedef struct {
    or_chal_t ch;
    f_t *data; size_t len, cap;   /* prover: proof_data being written */
    size_t pos;                   /* verifier: read cursor */
    int failed;                   /* verifier: read past end */
} or_ts_t;

void or_chal_init(or_chal_t *c, const or_p2_t *p16);
void or_chal_observe(or_chal_t *c, f_t v);
f_t or_chal_sample(or_chal_t *c);
uint32_t or_chal_sample_bits(or_chal_t *c, int bits);
ef_t or_chal_sample_ef(or_chal_t *c);

void or_ts_init(or_ts_t *t, const or_p2_t *p16, const f_t *domsep, size_t domsep_len);
void or_ts_init_verifier(or_ts_t *t, const or_p2_t *p16, const f_t *domsep, size_t domsep_len,
                         const f_t *proof, size_t len);
void or_ts_free(or_ts_t *t);
void or_ts_add_base(or_ts_t *t, const f_t *v, size_t n);
void or_ts_add_ef(or_ts_t *t, const ef_t *v, size_t n);
void or_ts_hint_base(or_ts_t *t, const f_t *v, size_t n);
ef_t or_ts_sample_ef(or_ts_t *t);
void or_ts_pow(or_ts_t *t, int bits);
/* verifier side */
int or_ts_next_base(or_ts_t *t, f_t *v, size_t n);
int or_ts_next_ef(or_ts_t *t, ef_t *v, size_t n);
int or_ts_next_hint_base(or_ts_t *t, f_t *v, size_t n);
int or_ts_check_pow(or_ts_t *t, int bits);
f_t or_grind_witness(const or_chal_t *c, int bits);

/* ------------------------------------------------------------------ poly -- */
size_t or_log2_up(size_t x);  /* crates/utils/src/misc.rs:1-8 */
void or_eval_eq(const ef_t *point, int n, ef_t *out);          /* 2^n, big-endian */
void or_eval_eq_scaled(const ef_t *point, int n, ef_t scale, ef_t *out);
void or_eval_eq_f(const f_t *point, int n, f_t *out);
ef_t or_eq_poly_outside(const ef_t *a, const ef_t *b, int n);
ef_t or_mle_eval_ef(const ef_t *evals, int n, const ef_t *point);
ef_t or_mle_eval_f(const f_t *evals, int n, const ef_t *point);
void or_expand_from_univariate(ef_t x, int n, ef_t *out);
/* dense univariate over EF, coefficients low degree first */
void or_lagrange_interp(const f_t *xs, const ef_t *ys, int n, ef_t *coeffs);
ef_t or_poly_eval(const ef_t *coeffs, int n, ef_t x);
void or_poly_mul(const ef_t *a, int na, const ef_t *b, int nb, ef_t *out);
/* univariate_selectors(k) evaluated at a point: out[i] = L_i(x), i < 2^k */
void or_selectors_at_f(int k, f_t x, f_t *out);
void or_selectors_at_ef(int k, ef_t x, ef_t *out);

/* -------------------------------------------------------- poseidon2 air -- */
#define OR_AIR_WIDTH 16
#define OR_AIR_HALF_FULL 4
#define OR_AIR_PARTIAL 20
#define OR_AIR_NCOLS 165
#define OR_AIR_NCONSTRAINTS 149
typedef struct {
    f_t beginning[OR_AIR_HALF_FULL][OR_AIR_WIDTH];
    f_t partial[OR_AIR_PARTIAL];
    f_t ending[OR_AIR_HALF_FULL][OR_AIR_WIDTH];
} or_air_consts_t;

void or_air_consts_from_rng(or_air_consts_t *c, or_rng_t *rng);
/* columns: OR_AIR_NCOLS column-major arrays of 2^log_n F */
void or_air_generate_trace(const or_air_consts_t *c, const f_t *inputs, int log_n, f_t *columns);
ef_t or_air_eval_f(const or_air_consts_t *c, const f_t *local, const ef_t *alpha_pows);
ef_t or_air_eval_ef(const or_air_consts_t *c, const ef_t *local, const ef_t *alpha_pows);

/* ------------------------------------------------------ constraint programs -- */
/* A SumcheckComputation given as a straight-line program (the symbolic
 * constraints of get_symbolic_constraints, crates/air/src/table.rs:42, in SSA
 * form): value i < n_inputs is point[i]; instruction j (4 words: op, a, b, 0)
 * defines value n_inputs + j.  ASSERT folds value a into the result with the next
 * batching scalar (ConstraintFolder::assert_zero, crates/utils/src/constraints_folder.rs:55-60).
 * CONST: a is the canonical constant. */
enum { OR_OP_CONST = 0, OR_OP_ADD = 1, OR_OP_SUB = 2, OR_OP_MUL = 3, OR_OP_NEG = 4, OR_OP_ASSERT = 5 };
typedef struct {
    int n_inputs, n_instr;
    const uint32_t *code;
} or_prog_t;
int or_prog_check(const or_prog_t *p); /* 0 when every operand refers to an earlier value */
int or_prog_n_asserts(const or_prog_t *p);
ef_t or_prog_eval_f(const or_prog_t *p, const f_t *pt, const ef_t *batching);
ef_t or_prog_eval_ef(const or_prog_t *p, const ef_t *pt, const ef_t *batching);

/* -------------------------------------------------------------- sumcheck -- */
enum { OR_SC_AIR = 0, OR_SC_PRODUCT = 1, OR_SC_A2B = 2, OR_SC_PROGRAM = 3 };
typedef struct {
    int kind;
    const or_air_consts_t *air;
    const ef_t *alpha_pows;  /* batching scalars (AIR and PROGRAM) */
    int n_cols;  /* for AIR: columns per row (up) ; the point has 2*n_cols entries */
    const or_prog_t *prog;   /* PROGRAM */
} or_sc_comp_t;
/* SumcheckGrinding (crates/sumcheck/src/lib.rs:14-31) */
enum { OR_GRIND_NONE = 0, OR_GRIND_AUTO = 1, OR_GRIND_CUSTOM = 2 };

/* crates/sumcheck/src/prove.rs:22-93.  mles: n_mles arrays (F if base_field else EF)
 * of 2^n_vars.  Outputs: challenges (n_rounds), final evals (n_mles EF), final sum. */
void or_sc_prove(int skips, int n_mles, int base_field, const void *const *mles, int n_vars,
                 const or_sc_comp_t *comp, int degree, const ef_t *eq_factor, int is_zerofier,
                 or_ts_t *ts, ef_t sum, int sec_bits, ef_t *challenges, ef_t *final_evals,
                 ef_t *final_sum);
/* The full interface of crates/sumcheck/src/prove.rs:22-35: n_rounds (0 = all,
 * n_vars - skips + 1), grinding (kind OR_GRIND_*, value = security bits for AUTO,
 * bits for CUSTOM), missing_mul_factor (NULL = None).  folded: the n_mles tables
 * after the last round, 2^(n_vars - skips - n_rounds + 1) EF each (may be NULL). */
void or_sc_prove_ex(int skips, int n_mles, int base_field, const void *const *mles, int n_vars,
                    const or_sc_comp_t *comp, int degree, const ef_t *eq_factor, int is_zerofier,
                    or_ts_t *ts, ef_t sum, int n_rounds, int grind_kind, int grind_value,
                    const ef_t *missing_mul_factor, ef_t *challenges, ef_t *folded, ef_t *final_sum);
/* crates/sumcheck/src/verify.rs:23-113 ; returns 0 on success */
int or_sc_verify(or_ts_t *ts, int n_vars, int degree, int sec_bits, ef_t *sum, ef_t *point,
                 ef_t *value);
int or_sc_verify_uskip(or_ts_t *ts, int degree, int n_vars, int skips, int sec_bits, ef_t *sum,
                       ef_t *point, ef_t *value);

/* ------------------------------------------------------------------ whir -- */
#define OR_WHIR_MAX_ROUNDS 16
typedef struct {
    int pow_bits, folding_pow_bits, num_queries, ood_samples, log_inv_rate;
    int num_variables, folding_factor, log_domain_size;
} or_whir_round_t;

/* SecurityAssumption (whir-p3 parameters/errors.rs), in the reference's order */
enum { OR_SA_UNIQUE_DECODING = 0, OR_SA_JOHNSON_BOUND = 1, OR_SA_CAPACITY_BOUND = 2 };
typedef struct {
    int num_variables, security_level, max_pow_bits, soundness;
    int f0, f;                     /* FoldingFactor::ConstantFromSecondRound(f0, f) */
    int starting_log_inv_rate, rs_reduction0;
    int n_rounds, final_sumcheck_rounds;
    int commit_ood_samples, starting_folding_pow_bits;
    or_whir_round_t rounds[OR_WHIR_MAX_ROUNDS];
    int final_queries, final_pow_bits, final_folding_pow_bits, final_log_inv_rate;
    int final_log_domain_size; /* domain of the last committed codeword */
} or_whir_cfg_t;

int or_whir_config(or_whir_cfg_t *c, int num_variables, int security_level, int max_pow_bits,
                   int f0, int f, int log_inv_rate, int rs_reduction0, int soundness);
int or_whir_folding_at(const or_whir_cfg_t *c, int round);
size_t or_whir_domsep(const or_whir_cfg_t *c, f_t *out, size_t cap);

typedef struct {
    int log_rows, width;  /* width in F per leaf */
    f_t *cw;              /* column-major: cw[col * rows + row] */
    f_t *nodes;           /* 8 * (2 * rows - 1) heap order: node 1 = root, leaves at rows.. */
} or_tree_t;

typedef struct {
    const or_whir_cfg_t *cfg;
    const or_p2_t *p16, *p24;
    f_t *evals;           /* 2^v base field evaluations (owned) */
    or_tree_t tree;
    ef_t ood_points[8], ood_answers[8];
    int n_ood;
} or_whir_witness_t;

void or_ntt(f_t *a, int log_n);                         /* natural order in/out */
void or_mobius(f_t *a, int log_n);                      /* evals -> monomial coeffs */
void or_rs_encode_columns(const f_t *evals, int log_m, int n_cols, int log_rows, f_t *cw);
void or_tree_build(or_tree_t *t, const or_p2_t *p16, const or_p2_t *p24);
void or_tree_free(or_tree_t *t);
void or_whir_commit(or_whir_witness_t *w, const or_whir_cfg_t *cfg, const or_p2_t *p16,
                    const or_p2_t *p24, or_ts_t *ts, const f_t *evals);
void or_whir_prove(or_whir_witness_t *w, or_ts_t *ts, const ef_t *point, ef_t value);
void or_whir_witness_free(or_whir_witness_t *w);
typedef struct {
    f_t root[8];
    int n_ood;
    ef_t ood_points[8], ood_answers[8];
} or_whir_commitment_t;
int or_whir_parse_commitment(const or_whir_cfg_t *cfg, or_ts_t *ts, or_whir_commitment_t *out);
int or_whir_verify(const or_whir_cfg_t *cfg, const or_p2_t *p16, const or_p2_t *p24, or_ts_t *ts,
                   const or_whir_commitment_t *cm, const ef_t *point, ef_t value);

/* ------------------------------------------------------------------- air -- */
typedef struct {
    int security_bits;
    int folding_f0, folding_f;
    int log_inv_rate;
    int univariate_skips;
    int rs_reduction0;
    int soundness;  /* OR_SA_* */
} or_air_settings_t;

typedef struct {
    or_air_consts_t air;
    or_p2_t p16, p24;
    f_t *inputs;   /* 2^n * 16 */
    f_t *trace;    /* 165 columns x 2^n */
    int log_n;
} or_p2_bench_t;

/* src/examples/poseidon2.rs:85-151 (inputs, trace, hash constants) */
void or_p2_bench_setup(or_p2_bench_t *b, int log_n, uint64_t seed);
void or_p2_bench_free(or_p2_bench_t *b);
/* crates/air/src/prove.rs:49-226 ; returns proof length in F */
size_t or_air_prove(const or_air_settings_t *s, const or_p2_bench_t *b, f_t **proof_out);
/* the same, appended to a caller's transcript */
void or_air_prove_ts(const or_air_settings_t *s, const or_p2_bench_t *b, or_ts_t *ts);
/* crates/air/src/verify.rs:55-249 ; returns 0 on accept, error code otherwise */
int or_air_verify(const or_air_settings_t *s, const or_p2_bench_t *b, const f_t *proof,
                  size_t len);

/* Generic AIR (any A: Air given as a constraint program; crates/air/src/table.rs:32-56) */
typedef struct {
    int n_columns;            /* air.width(): preprocessed + witness columns */
    int n_preprocessed;       /* the first columns of the point */
    int log_n;
    const f_t *preprocessed;  /* n_preprocessed columns x 2^log_n */
    or_prog_t prog;           /* constraints over [up(0..C), down(0..C)] */
    int degree;               /* constraint_degree */
} or_gair_t;
int or_gair_whir_config(or_whir_cfg_t *c, const or_air_settings_t *s, const or_gair_t *a);
/* witness: (n_columns - n_preprocessed) columns x 2^log_n; appends to ts */
void or_gair_prove(const or_air_settings_t *s, const or_gair_t *a, const f_t *witness, const or_p2_t *p16,
                   const or_p2_t *p24, or_ts_t *ts);
int or_gair_verify(const or_air_settings_t *s, const or_gair_t *a, const or_p2_t *p16, const or_p2_t *p24,
                   or_ts_t *ts);

/* whir config used by AirTable::build_whir_params (crates/air/src/table.rs:72-98) */
int or_air_whir_config(or_whir_cfg_t *c, const or_air_settings_t *s, int log_n);

#ifdef __cplusplus
}
#endif
#endif
