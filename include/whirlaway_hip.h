/*
 * whirlaway_hip.h -- C ABI of the MI355X (gfx950) Whirlaway prover.
 *
 * Drop-in boundary for the reference's prover hot path.  Each entry point
 * names the reference interface it replaces (paths relative to the reference
 * repository root).  Conventions:
 *   - plain pointers and sizes; no torch or HIP types cross the boundary;
 *   - field elements are canonical little-endian u32 in [0, p), p = 2^31 - 2^24 + 1;
 *     an extension element is 8 consecutive u32 (coefficient 0 first);
 *   - every call returns a status (WL_OK == 0) and never unwinds; the message of
 *     the last failure on the calling thread is wl_last_error();
 *   - one HIP stream per context; a context is not thread-safe, separate
 *     contexts are.
 */
#ifndef WHIRLAWAY_HIP_H
#define WHIRLAWAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* WL_ERR_VERIFY: a proof was rejected (verifier entry points) */
enum { WL_OK = 0, WL_ERR_ARG = 1, WL_ERR_DEVICE = 2, WL_ERR_BUFFER = 3, WL_ERR_VERIFY = 4 };

typedef struct wl_ctx wl_ctx;
typedef struct wl_table wl_table;
typedef struct wl_witness wl_witness;
typedef struct wl_transcript wl_transcript;

/* SecurityAssumption of whir-p3 (crates/air/src/lib.rs:3,15), in its order. */
enum { WL_UNIQUE_DECODING = 0, WL_JOHNSON_BOUND = 1, WL_CAPACITY_BOUND = 2 };

/* AirSettings (crates/air/src/lib.rs:13-41) with
 * FoldingFactor::ConstantFromSecondRound(folding_factor_first, folding_factor)
 * (src/main.rs:16-23 passes 128, CapacityBound, (7, 4), 1, 4, 5). */
typedef struct {
    uint32_t security_bits;               /* 128 */
    uint32_t folding_factor_first;        /* 7 */
    uint32_t folding_factor;              /* 4 */
    uint32_t log_inv_rate;                /* 1 */
    uint32_t univariate_skips;            /* 4 (must equal the table's) */
    uint32_t rs_domain_initial_reduction; /* 5 */
    uint32_t soundness;                   /* whir_soudness_type: WL_CAPACITY_BOUND */
} wl_air_settings;

/* Message of the last failed call on this thread. */
const char *wl_last_error(void);

/* Checks settings against a table of 2^log_length rows whose committed
 * polynomial has n_witness_columns columns, without a device: the WHIR schedule
 * of AirTable::build_whir_params (crates/air/src/table.rs:72-98) must exist and
 * every codeword domain must fit the field's 2^24 two-adic subgroup (the
 * reference panics inside WhirConfig::new / the DFT instead).  WL_OK or
 * WL_ERR_ARG with the reason in wl_last_error(). */
int wl_air_settings_validate(uint32_t log_length, uint32_t n_witness_columns, const wl_air_settings *settings);

/* Number of visible HIP devices (0 when none). */
int wl_device_count(void);

/* Context: owns the HIP stream, the device memory pool and twiddle tables.
 * Replaces the implicit global rayon pool of the reference
 * (crates/sumcheck/src/prove.rs:6, crates/utils/src/multilinear.rs:6). */
wl_ctx *wl_create(int device);
/* Frees the context's device memory.  Frees wait for the whole device, so this
 * waits until no proof with device-side waits (below) runs on any context. */
void wl_destroy(wl_ctx *ctx);
/* Device-side waits (default on): each sumcheck round after the first is
 * enqueued before its challenge exists, behind a kernel that spins on mapped
 * memory until the host writes the challenge, so the host's launches stay off the
 * critical path.  While such a proof runs, another thread's device-wide
 * synchronization (hipDeviceSynchronize, hipFree, ...) waits for it, and if that
 * thread holds up this one, both stall until the wait times out (10 s, the proof
 * then fails).  With enabled = 0 the rounds are enqueued once their challenges
 * are known (about 0.15 ms more per proof at 2^14 - 2^17 rows); engine slots run
 * this way.  No counterpart in the reference. */
int wl_set_device_waits(wl_ctx *ctx, int enabled);

/* Bound of each device-side wait, in seconds (0 restores the default, 10 s).  A
 * wait that reaches it ends its kernel without the host value; the proof then
 * fails with "a device wait for a host value timed out", and the context serves
 * later calls.  No counterpart in the reference (its prover does not wait on the
 * device). */
int wl_set_wait_timeout(wl_ctx *ctx, double seconds);

/* AirTable::new for the Poseidon2-16 KoalaBear AIR (crates/air/src/table.rs:32-56,
 * called at src/examples/poseidon2.rs:140-146).  round_constants: the AIR's
 * RoundConstants (148 canonical u32: beginning[4][16], partial[20],
 * ending[4][16]; p3-poseidon2-air RoundConstants::from_rng order).  ctx may be
 * NULL for a table that only verifies (wl_air_verify, no device needed). */
wl_table *wl_air_table_new(wl_ctx *ctx, const uint32_t *round_constants, uint32_t log_length,
                           uint32_t univariate_skips);
void wl_air_table_free(wl_table *table);

/* AirTable::new(air, log_length, univariate_skips, preprocessed_columns,
 * constraint_degree) (crates/air/src/table.rs:32-56) for any AIR: width
 * n_columns, the first n_preprocessed of them preprocessed (n_preprocessed x
 * 2^log_length canonical, column-major: prove.rs:109-113 chains them before the
 * witness), constraints as a program over [local row, next row] (2 n_columns
 * inputs; see wl_sumcheck_prove), the i-th assert receiving alpha^i.  Such a table
 * proves with wl_air_prove / wl_air_prove_into on a witness of n_columns -
 * n_preprocessed columns and verifies with wl_air_verify; with ctx NULL it only
 * verifies. */
wl_table *wl_air_table_new_program(wl_ctx *ctx, uint32_t n_columns, uint32_t n_preprocessed,
                                   const uint32_t *preprocessed, uint32_t log_length, uint32_t univariate_skips,
                                   const uint32_t *program, uint32_t n_instr, uint32_t constraint_degree);

/* AirTable::prove (crates/air/src/prove.rs:49-226) including the transcript set
 * up by DomainSeparator/to_prover_state (src/examples/poseidon2.rs:155-163).
 * p16/p24: Poseidon2KoalaBear<16>/<24> constants (canonical): ext_init[4][W],
 * ext_term[4][W], internal[20 | 23] (new_from_rng_128 order).  witness: 165
 * columns x 2^log_length, column-major (Vec<EvaluationsList<F>> of
 * prove.rs:55).  The proof is proof_data (Vec<F>, src/examples/poseidon2.rs:178):
 * *proof_len always receives its length; proof_out may be NULL to query it. */
int wl_air_prove(wl_table *table, const wl_air_settings *settings, const uint32_t *p16,
                 const uint32_t *p24, const uint32_t *witness, uint32_t *proof_out,
                 size_t proof_cap, size_t *proof_len);

/* The same proof appended to a caller's transcript, as AirTable::prove takes
 * `prover_state: &mut ProverState` (crates/air/src/prove.rs:54) that the caller
 * built from its DomainSeparator (src/examples/poseidon2.rs:155-163): the
 * transcript continues from its current state, and proof_data grows by the
 * proof (wl_transcript_proof_data).  The transcript's Poseidon2-16 must equal p16,
 * as the reference hands the same permutation to the challenger and to the
 * Merkle compression (src/examples/poseidon2.rs:148-163). */
int wl_air_prove_into(wl_table *table, const wl_air_settings *settings, const uint32_t *p16,
                      const uint32_t *p24, const uint32_t *witness, wl_transcript *ts);

/* The domain-separator words this build observes first in wl_air_prove (the
 * counterpart of DomainSeparator::commit_statement + add_whir_proof,
 * src/examples/poseidon2.rs:155-157): one (tag, count) pair per transcript
 * operation of the WHIR commitment and opening, tags 1 observe, 2 sample, 3 hint,
 * 4 proof of work.  A caller whose whir-p3 version encodes the pattern otherwise
 * passes its own words to wl_transcript_new and proves with wl_air_prove_into.
 * *len receives the length; out may be NULL. */
int wl_air_domain_separator(uint32_t log_length, uint32_t n_witness_columns, const wl_air_settings *settings,
                            uint32_t *out, size_t cap, size_t *len);

/* ----------------------------------------------------------------- verify --
 * AirTable::verify(settings, merkle_hash, merkle_compress, verifier_state,
 * log_length) (crates/air/src/verify.rs:55-249) on the host, for either kind of
 * table: verifier_state is a verifier transcript (wl_transcript_new_verifier over
 * the proof_data and the domain separator the prover's transcript started from),
 * read in the prover's order.  WL_OK, or WL_ERR_VERIFY with *error set to the
 * AirVerifError kind (verify.rs:26-34): 1 InvalidPcsCommitment, 2
 * InvalidPcsOpening, 3 Fs, 4 Sumcheck, 5 InvalidBoundaryCondition, 6 SumMismatch. */
int wl_air_verify(const wl_table *table, const wl_air_settings *settings, const uint32_t *p16, const uint32_t *p24,
                  wl_transcript *verifier_state, int *error);
/* Batch verification on n_threads host threads of n proofs of one table made by
 * wl_air_prove (this build's domain separator); results[i]: 0 accepted, else the
 * AirVerifError kind (3 also for words left over after the proof). */
int wl_air_verify_batch(const wl_table *table, const wl_air_settings *settings, const uint32_t *p16,
                        const uint32_t *p24, const uint32_t *const *proofs, const size_t *proof_lens,
                        size_t n_proofs, int n_threads, int *results);

/* Wire format of a proof (proof_data, src/examples/poseidon2.rs:178, whose size
 * the reference counts at log2(p) bits per word, :193): a 48-byte header of
 * little-endian u32 {"WLP1", version 1, log_length, n_witness_columns, the 7
 * settings words, number of words}, the words packed at 31 bits each
 * (little-endian bit order), and a 32-bit FNV-1a checksum of everything before
 * it.  Deserialization refuses a wrong magic, version, length or checksum and
 * non-canonical words; *proof_len receives the word count (proof may be NULL). */
size_t wl_proof_serialized_size(size_t proof_len);
int wl_proof_serialize(const wl_air_settings *settings, uint32_t log_length, uint32_t n_witness_columns,
                       const uint32_t *proof, size_t proof_len, uint8_t *out, size_t cap, size_t *out_len);
int wl_proof_deserialize(const uint8_t *in, size_t in_len, wl_air_settings *settings, uint32_t *log_length,
                         uint32_t *n_witness_columns, uint32_t *proof, size_t cap, size_t *proof_len);

/* Witness resident in HBM (the commitment's packed polynomial,
 * crates/utils/src/multilinear.rs:169-181, built on the device). */
wl_witness *wl_witness_upload(wl_ctx *ctx, const uint32_t *witness, uint32_t log_length);
void wl_witness_free(wl_witness *w);
/* wl_air_prove on a resident witness (no host-to-device copy of the trace). */
int wl_air_prove_resident(wl_table *table, const wl_air_settings *settings, const uint32_t *p16,
                          const uint32_t *p24, const wl_witness *w, uint32_t *proof_out,
                          size_t proof_cap, size_t *proof_len);
/* wl_air_prove_into on a resident witness. */
int wl_air_prove_resident_into(wl_table *table, const wl_air_settings *settings, const uint32_t *p16,
                               const uint32_t *p24, const wl_witness *w, wl_transcript *ts);

/* ------------------------------------------------------------- transcript --
 * ProverState / VerifierState of whir-p3 over DuplexChallenger<F, Poseidon2KoalaBear<16>,
 * 16, 8> (src/examples/poseidon2.rs:27,157-163,183-184), host code.  Created
 * from the Poseidon2-16 constants (as wl_air_prove's p16) and the domain
 * separator's field elements, which the challenger observes first
 * (DomainSeparator::to_prover_state / to_verifier_state).  The sponge: observe
 * overwrites the rate (8 words) and permutes when it is full; sample permutes
 * when input is pending or the output is used up and pops the output's last word. */
wl_transcript *wl_transcript_new(const uint32_t *p16, const uint32_t *domsep, size_t domsep_len);
wl_transcript *wl_transcript_new_verifier(const uint32_t *p16, const uint32_t *domsep, size_t domsep_len,
                                          const uint32_t *proof, size_t proof_len);
void wl_transcript_free(wl_transcript *ts);
/* FieldChallenger::observe (nothing appended) */
int wl_transcript_observe(wl_transcript *ts, const uint32_t *v, size_t n);
/* ProverState::add_base_scalars / add_extension_scalars: append and observe */
int wl_transcript_add_base(wl_transcript *ts, const uint32_t *v, size_t n);
int wl_transcript_add_ext(wl_transcript *ts, const uint32_t *v, size_t n_ef);
/* ProverState hint: append without observing (Merkle openings) */
int wl_transcript_hint(wl_transcript *ts, const uint32_t *v, size_t n);
/* ProverState/VerifierState::sample (EF) and sample_bits */
int wl_transcript_sample_ext(wl_transcript *ts, uint32_t *out, size_t n_ef);
int wl_transcript_sample_bits(wl_transcript *ts, uint32_t bits, uint32_t *out);
/* pow_grinding(bits) (crates/air/src/utils.rs:162-212): the smallest witness w with
 * sample_bits(bits) == 0 after observe(w), appended; bits 0 does nothing.  On ctx's
 * device when ctx is not NULL, else on the host.  *witness (may be NULL). */
int wl_transcript_grind(wl_transcript *ts, wl_ctx *ctx, uint32_t bits, uint32_t *witness);
/* proof_data (src/examples/poseidon2.rs:178): length and copy */
size_t wl_transcript_proof_len(const wl_transcript *ts);
int wl_transcript_proof_data(const wl_transcript *ts, uint32_t *out, size_t cap);
/* The sponge as 34 canonical words: state[16], input[8], input count, output[8],
 * output count; import restores it (a transcript continued elsewhere). */
#define WL_TRANSCRIPT_STATE_WORDS 34
int wl_transcript_export(const wl_transcript *ts, uint32_t *state);
int wl_transcript_import(wl_transcript *ts, const uint32_t *state);
/* Verifier: next_base_scalars / next_extension_scalars (read and observe), hints
 * (read only), check_pow_grinding; WL_ERR_VERIFY when the proof is exhausted, holds
 * a non-canonical word or the witness fails. */
int wl_transcript_next_base(wl_transcript *ts, uint32_t *out, size_t n);
int wl_transcript_next_ext(wl_transcript *ts, uint32_t *out, size_t n_ef);
int wl_transcript_next_hint(wl_transcript *ts, uint32_t *out, size_t n);
int wl_transcript_check_pow(wl_transcript *ts, uint32_t bits);
size_t wl_transcript_remaining(const wl_transcript *ts);

/* Synthetic benchmark inputs exactly as src/examples/poseidon2.rs:104-151 draws
 * them from StdRng::seed_from_u64(seed): air_rc (148, RoundConstants::from_rng),
 * inputs (16 x 2^log_n, row-major), p16 (148) and p24 (215) Poseidon2
 * constants (new_from_rng_128).  Any output pointer may be NULL. */
int wl_p2_bench_inputs(uint64_t seed, uint32_t log_n, uint32_t *air_rc, uint32_t *inputs,
                       uint32_t *p16, uint32_t *p24);

/* generate_trace_rows (src/examples/poseidon2.rs:122-131) on the GPU: one
 * Poseidon2 permutation per row with the table's round constants, written
 * column-major into a resident witness. */
wl_witness *wl_air_generate_witness(wl_table *table, const uint32_t *inputs);

/* Copy a resident witness (165 x 2^n canonical, column-major) to the host. */
int wl_witness_download(const wl_witness *w, uint32_t *out);

/* Accumulated per-phase wall times in ms since the last reset: [commit,
 * zerocheck round 0, zerocheck rounds >= 1, inner sums + batching, inner
 * sumcheck, WHIR opening, host time spent waiting for device results]
 * (no counterpart in the reference; used by the benchmark). */
void wl_phase_times(wl_ctx *ctx, double *out7, int reset);

/* Live timing of kernel classes with HIP events on the context's stream (no
 * counterpart in the reference; used by the benchmark's roofline line).
 * Classes (bit k of class_mask): 1 Merkle leaf hashing (work = Poseidon2-24
 * permutations), 2 zerocheck rounds >= 1 (work = evaluation points), 3 RS
 * encoding (work = codeword elements), 4 zerocheck round 0 (points),
 * 5 grinding (candidates), 6 Merkle compression levels (Poseidon2-16
 * permutations).  wl_kernel_timing resets the counters; _read synchronizes and
 * sums one class. */
int wl_kernel_timing(wl_ctx *ctx, uint32_t class_mask);
int wl_kernel_timing_read(wl_ctx *ctx, int kernel_class, long *launches, double *total_ms,
                          double *work);

/* Measured throughput of the Montgomery product (F x F) on the context's
 * device: 2^20 threads, four independent product chains each (no counterpart in
 * the reference; used by the benchmark beside its VALU roofline). */
int wl_fmul_rate(wl_ctx *ctx, double *fmul_per_s);

/* Host-side check of the device Poseidon2 form (partial rounds as a convolution,
 * paired diagonal entries) against the round-by-round permutation: random round
 * constants from `seed`, n_states random states plus field-extreme ones, for
 * width 16 (mismatches[0]) and 24 (mismatches[1]).  Width 16 also counts states
 * on which the transcript's AVX2 permutation differs, and inputs (field extremes
 * and n_states random ones) on which its partial-round cube differs from the
 * Montgomery cube.  No GPU is used; no
 * counterpart in the reference (it checks this build's own rewriting of
 * Poseidon2Koalabear's permutation). */
int wl_p2_conv_check(uint32_t n_states, uint64_t seed, uint64_t mismatches[2]);

/* Device memory of a context's pool: bytes held (live + cached) and the peak of
 * live bytes since the last reset (no counterpart in the reference). */
void wl_mem_stats(wl_ctx *ctx, uint64_t *held, uint64_t *peak_live, int reset_peak);

/* ----------------------------------------------------- batched and served --
 * The reference proves one table at a time (src/examples/poseidon2.rs:153-174);
 * these entry points have no counterpart there.  An engine runs independent
 * proof jobs concurrently on one GPU: n_slots worker threads, each with its own
 * context (stream + pool).  Jobs are admitted in FIFO order when a slot is free
 * and wl_job_footprint(log_n, settings) fits the memory budget beside the
 * running jobs' footprints (budget 0: 80% of the device's free memory).  Each job runs the
 * single-proof path (wl_air_table_new, wl_air_generate_witness,
 * wl_air_prove_resident), so its proof equals the one proven alone.  Pools keep
 * freed blocks for reuse, up to about one footprint per slot, outside the budget. */
typedef struct wl_engine wl_engine;
typedef struct {
    uint32_t log_n;            /* 2^log_n permutations (trace rows) */
    const uint32_t *air_rc;    /* 148 AIR round constants (as wl_air_table_new) */
    const uint32_t *inputs;    /* 16 x 2^log_n, row-major (as wl_air_generate_witness) */
    const uint32_t *p16;       /* Poseidon2-16 constants (as wl_air_prove) */
    const uint32_t *p24;       /* Poseidon2-24 constants */
    wl_air_settings settings;
} wl_job;                      /* arrays must stay valid until the job has completed */

/* Device bytes reserved for a job of 2^log_n rows under its settings (NULL: the
 * defaults of src/main.rs): witness, first codeword with its encoding scratch and
 * Merkle tree, zerocheck, inner-sumcheck and WHIR-round buffers, plus a margin. */
uint64_t wl_job_footprint(uint32_t log_n, const wl_air_settings *settings);
wl_engine *wl_engine_create(int device, int n_slots, uint64_t mem_budget_bytes);
/* Completes queued jobs with an "engine stopped" error (their waiters return it),
 * waits for running jobs and for every thread inside wl_engine_wait to leave,
 * then frees the slots' contexts. */
void wl_engine_destroy(wl_engine *e);
int wl_engine_submit(wl_engine *e, const wl_job *job, uint64_t *ticket);
/* 1 done, 0 pending, -1 unknown ticket */
int wl_engine_poll(wl_engine *e, uint64_t ticket);
/* Blocks until the job is done; proof as wl_air_prove; times_ms (may be NULL):
 * [queued, running, submit-to-done latency].  The ticket is retired unless the
 * buffer was too small (WL_ERR_BUFFER, *proof_len set). */
int wl_engine_wait(wl_engine *e, uint64_t ticket, uint32_t *proof_out, size_t proof_cap, size_t *proof_len,
                   double *times_ms);
/* [completed jobs, peak concurrent jobs, peak reserved bytes, budget bytes] */
void wl_engine_stats(wl_engine *e, uint64_t *out4);

/* --------------------------------------------------------------- autotuner --
 * No counterpart in the reference (its CPU prover has no launch shapes).  The
 * commitment kernels have launch variants that compute identical values: kind 0
 * the split 2^L = 2^LA x 2^LB of the two-pass Reed-Solomon NTT (variants 0..2:
 * LB = 10, 9, 8), kind 1 Merkle leaf hashing (0: a lane per row, 256-thread
 * blocks; 1: a lane pair per row; 2, 3: the same in 64-thread blocks), kind 2
 * the schedule of the tree levels (0..5).  With tuning on (the default;
 * WHIRLAWAY_AUTOTUNE=0 or mode 0 turns it off) the first calls of a (kind, log
 * size, width) on a device each run the next variant, timed in place by events
 * around the call, until every variant has two samples; the fastest then serves
 * every later call of every context (a process-wide table). */
/* mode 0: defaults only, 1: tune on first use, other values only query; returns
 * the previous mode */
int wl_autotune_mode(int mode);
/* sets the variant of (device, kind, log size, width); width counts only for
 * kind 1 (leaf width in field elements) */
int wl_autotune_force(int device, int kind, int log_size, int width, int variant);
/* table entries, 6 per entry: device, kind, log size, width, variant (-1 while
 * its variants are still being timed), best time in ns (-1: forced); returns the
 * number of entries (written up to cap) */
size_t wl_autotune_entries(int64_t *out, size_t cap);
void wl_autotune_clear(void);

/* ---------------------------------------------------------------- kernels --
 * crates/utils/src/multilinear.rs:90-123 fold_multilinear_in_large_field:
 * out[i] = sum_j scalars[j] * in[i + j * 2^(log_n - log_k)], in F (base) or EF,
 * 2^log_k EF scalars, EF output of 2^(log_n - log_k) elements. */
int wl_fold_large(wl_ctx *ctx, const uint32_t *in, int in_is_ef, uint32_t log_n,
                  const uint32_t *scalars, uint32_t log_k, uint32_t *out);

/* crates/utils/src/multilinear.rs:10-52 fold_multilinear_in_small_field:
 * out[i] = sum_j scalars[j] * in[i + j * 2^(log_n - log_k)] with 2^log_k base-field
 * scalars; in and out both F (in_is_ef = 0) or both EF. */
int wl_fold_small(wl_ctx *ctx, const uint32_t *in, int in_is_ef, uint32_t log_n, const uint32_t *scalars,
                  uint32_t log_k, uint32_t *out);

/* crates/utils/src/multilinear.rs:125-147 multilinears_linear_combination:
 * out[x] = sum_c scalars[c] * cols[c * 2^log_n + x]; n_cols base-field columns,
 * n_cols EF scalars, 2^log_n EF outputs. */
int wl_multilinears_linear_combination(wl_ctx *ctx, const uint32_t *cols, uint32_t n_cols, uint32_t log_n,
                                       const uint32_t *scalars, uint32_t *out);

/* EvaluationsList::evaluate (whir-p3; called at crates/air/src/prove.rs:190 and
 * crates/air/src/verify.rs:114-136): the multilinear extension of 2^n_vars
 * evaluations (F, or EF when is_ef) at an EF point of n_vars coordinates
 * (coordinate 0 binds the most significant index bit), n_vars <= 28; out: 1 EF. */
int wl_mle_evaluate(wl_ctx *ctx, const uint32_t *evals, int is_ef, uint32_t n_vars, const uint32_t *point,
                    uint32_t *out);

/* No counterpart in the reference: wl_mle_evaluate for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc).  evals_dev (2^n_vars canonical F, or EF when
 * is_ef) and out_dev (8 canonical words) are device memory the caller owns; the
 * point (n_vars EF) is host memory, passed in the kernel arguments.  The call
 * allocates nothing: its scratch (eq tables and block partials) is a fixed
 * 1,081,344-byte device array per GPU, held until the process ends, so n_vars is
 * at most 28 (WL_ERR_ARG beyond).  Runs on ctx's stream and returns when the result
 * is in out_dev.  Calls that overlap on one GPU (other threads or contexts) share
 * that array and race. */
int wl_mle_evaluate_noalloc(wl_ctx *ctx, const uint32_t *evals_dev, int is_ef, uint32_t n_vars,
                            const uint32_t *point, uint32_t *out_dev);

/* Kernel micro (a), benches/folding_benchmarks.rs: fold 2^log_n base-field
 * evaluations by n_folds successive EF challenges (1 - c, c), most-significant
 * variable first; after each fold the intermediate table is checked by the
 * caller through out_all (concatenated EF tables of sizes 2^(log_n-1), ...),
 * or only the last table when out_all is NULL (final_out: 2^(log_n-n_folds) EF).
 * Device time of the folds (ms) in *ms when non-NULL. */
int wl_fold_chain(wl_ctx *ctx, const uint32_t *in, uint32_t log_n, const uint32_t *challenges,
                  uint32_t n_folds, uint32_t *out_all, uint32_t *final_out, double *ms);

/* ------------------------------------------------------ generic sumcheck --
 * A SumcheckComputation (crates/sumcheck/src/sc_computation.rs:8-36) given as a
 * constraint program -- the symbolic constraints of an AIR
 * (get_symbolic_constraints, crates/air/src/table.rs:42) or any polynomial of the
 * point -- in straight-line SSA form: 4 words per instruction {op, a, b, 0}.
 * Value i < n_inputs is point[i] (the i-th multilinear; for an AIR, local column
 * i for i < width, next-row column i - width above); instruction j defines value
 * n_inputs + j, except WL_OP_ASSERT, which defines none and adds
 * batching[k] * value(a) for its k-th occurrence (ConstraintFolder::assert_zero,
 * crates/utils/src/constraints_folder.rs:55-60).  WL_OP_CONST: a is a canonical
 * constant.  At most 256 values may be live at once. */
enum { WL_OP_CONST = 0, WL_OP_ADD = 1, WL_OP_SUB = 2, WL_OP_MUL = 3, WL_OP_NEG = 4, WL_OP_ASSERT = 5 };
/* SumcheckGrinding (crates/sumcheck/src/lib.rs:14-31): NONE, AUTO (value = security
 * bits), CUSTOM (value = pow bits) */
enum { WL_GRIND_NONE = 0, WL_GRIND_AUTO = 1, WL_GRIND_CUSTOM = 2 };

/* sumcheck::prove(skips, multilinears, computation, degree, batching_scalars,
 * eq_factor, is_zerofier, prover_state, sum, n_rounds, grinding,
 * missing_mul_factor) (crates/sumcheck/src/prove.rs:22-93, rounds :97-254) on the
 * device.  mles: n_mles tables of 2^n_vars values, contiguous, F (or EF when
 * mles_are_ef); the program reads n_mles inputs.  batching: n_batching EF (at
 * least one per assert).  eq_factor: n_vars - skips + 1 EF, or NULL.  sum: EF.
 * n_rounds: 0 for all (n_vars - skips + 1).  missing_mul_factor: EF or NULL.
 * Round polynomials (coefficients) and grinding witnesses are appended to ts.
 * Outputs (any may be NULL): challenges (n_rounds EF), folded (the n_mles tables
 * after the last round, 2^(n_vars - skips - n_rounds + 1) EF each), the final sum. */
int wl_sumcheck_prove(wl_ctx *ctx, wl_transcript *ts, uint32_t skips, const uint32_t *mles, uint32_t n_mles,
                      int mles_are_ef, uint32_t n_vars, const uint32_t *program, uint32_t n_instr, uint32_t degree,
                      const uint32_t *batching, uint32_t n_batching, const uint32_t *eq_factor, int is_zerofier,
                      const uint32_t *sum, uint32_t n_rounds, int grinding, uint32_t grinding_value,
                      const uint32_t *missing_mul_factor, uint32_t *challenges_out, uint32_t *folded_out,
                      uint32_t *sum_out);

/* Kernel micro (b): sumcheck::prove(1, [a, b], a^2 * b, degree 3, [1],
 * Some(z), false, ...) (crates/sumcheck/src/prove.rs:22-93) over a transcript
 * whose challenger is Poseidon2KoalaBear<16> with constants p16 and an empty
 * domain separator.  a, b: 2^n_vars base-field evaluations; z: n_vars EF.
 * Outputs: challenges (n_vars EF), final evaluations of a and b (2 EF), final
 * sum (1 EF), proof_data (as wl_air_prove). */
int wl_sumcheck_a2b(wl_ctx *ctx, const uint32_t *p16, const uint32_t *a, const uint32_t *b,
                    uint32_t n_vars, const uint32_t *z, uint32_t *challenges,
                    uint32_t *final_evals, uint32_t *final_sum, uint32_t *proof_out,
                    size_t proof_cap, size_t *proof_len, double *ms);

/* Poseidon2 permutations on the device (p3-poseidon2 Poseidon2KoalaBear<W>,
 * src/examples/poseidon2.rs:22-26): n states of width 16 or 24, in place. */
int wl_poseidon2_permute(wl_ctx *ctx, const uint32_t *p16, const uint32_t *p24, int width,
                         uint32_t *states, size_t n);

/* Reed-Solomon encoding of n_cols columns of 2^log_m evaluations (univariate
 * embedding, natural-order NTT over the 2^log_rows-th roots of unity) and the
 * Merkle root of its rows (PaddingFreeSponge<P24,24,16,8> leaves,
 * TruncatedPermutation<P16,2,8,16> nodes) -- the CommitmentWriter::commit
 * core called at crates/air/src/prove.rs:91.  codeword_out (n_cols x
 * 2^log_rows, column-major) may be NULL. */
int wl_rs_commit(wl_ctx *ctx, const uint32_t *p16, const uint32_t *p24, const uint32_t *cols,
                 uint32_t n_cols, uint32_t log_m, uint32_t log_rows, uint32_t *codeword_out,
                 uint32_t *root_out);

#ifdef __cplusplus
}
#endif
#endif
