// sc_generic.h -- the sumcheck of crates/sumcheck/src/prove.rs:22-254 for any
// computation given as a constraint program (host interface of sc_generic.hip).
//
// A program is straight-line SSA over the point (the values of the n_inputs
// multilinears at one hypercube point): 4 words per instruction {op, a, b, 0};
// value i < n_inputs is input i, instruction j defines value n_inputs + j (except
// ASSERT, which defines none); CONST takes the canonical constant in a; ASSERT
// adds batching[k] * value(a) to the result for its k-th occurrence
// (ConstraintFolder::assert_zero, crates/utils/src/constraints_folder.rs:55-60).
// The host compiles it to slot code for the device: operands read from a small
// register file, inputs loaded where first used, slots reused once their value
// is dead.
#pragma once
#include <vector>
#include "ctx.h"

namespace wg {
using namespace kb;

// OP_LOAD and OP_MULC (a product with a constant operand) occur in slot code only
enum Op : uint32_t { OP_CONST = 0, OP_ADD = 1, OP_SUB = 2, OP_MUL = 3, OP_NEG = 4, OP_ASSERT = 5, OP_LOAD = 6, OP_MULC = 7 };

// The constraints split into `groups` groups, each evaluated by its own threads:
// slot code of group g at code[off[g] .. off[g + 1]).
struct Partition {
    int groups = 0;
    int n_slots = 0;
    std::vector<uint32_t> off;
    std::vector<uint32_t> code;
    // the same groups as bundles (k_eval_bundles): a header {op | count << 8, 0} and
    // `count` (1..4) entries {dst, operands} of independent instructions of one op
    int bn_slots = 0;
    std::vector<uint32_t> boff;
    std::vector<uint32_t> bcode;
};
struct Program {
    int n_inputs = 0;
    int n_asserts = 0;
    int n_slots = 0;                 // register-file size the one-group slot code needs
    std::vector<uint32_t> dev;       // one-group slot code, 2 words per instruction
    std::vector<Partition> parts;    // 1, 4 and 16 groups (as many as the constraints allow)
    wl::DBuf<uint32_t> dev_buf;      // every partition on the device: offsets, then code
    std::vector<size_t> part_base;   // word offset of each partition in dev_buf
    std::vector<size_t> bpart_base;  // word offset of each partition's bundles: offsets, then code
};

// Validates the SSA program and compiles it; throws std::runtime_error on a
// malformed program (operand out of range or referring to an ASSERT, unknown op).
void compile(const uint32_t *code, size_t n_instr, int n_inputs, Program &out);
void upload(wl_ctx *c, Program &p);

// An AIR given as a constraint program over [local(width), next(width)]
// (AirTable::new, crates/air/src/table.rs:32-56): the first n_preprocessed columns
// are preprocessed (held here on the device), the others are the witness.
struct AirSpec {
    int n_columns = 0, n_preprocessed = 0, degree = 1;
    Program prog;
    wl::DBuf<F> pre;  // n_preprocessed x 2^log_length, Montgomery
    // host copies for the verifier: the SSA program and the preprocessed columns (canonical)
    std::vector<uint32_t> code;
    std::vector<uint32_t> pre_host;
};

// columns_up_and_down (crates/air/src/utils.rs:140-160) of preprocessed ++
// witness: out = [up(col_0) .. up(col_{C-1}), down(col_0) .. down(col_{C-1})],
// 2C tables of 2^log_n F; up(c)[N-1] = c[N-2], down(c) = c[1..] ++ c[N-1].
void columns_up_down(const F *pre, int n_pre, const F *wit, int n_cols, int log_n, F *out, hipStream_t s);

// SumcheckGrinding (crates/sumcheck/src/lib.rs:14-31)
enum Grinding { GRIND_NONE = 0, GRIND_AUTO = 1, GRIND_CUSTOM = 2 };

struct Args {
    int skips = 1;
    int degree = 1;                          // constraints_degree
    const std::vector<EF> *batching = nullptr;
    const std::vector<EF> *eq_factor = nullptr;  // n_vars - skips + 1 entries, or none
    bool is_zerofier = false;
    EF sum{};
    int n_rounds = 0;                        // 0: n_vars - skips + 1
    int grind_kind = GRIND_AUTO, grind_value = 128;
    const EF *missing_mul_factor = nullptr;
};

// sumcheck::prove on n_mles device tables of 2^n_vars entries, contiguous
// ([m][x], F when base, else EF, Montgomery form).  Appends to ts (grinding on
// the device through pdev's Poseidon2-16).  Returns the challenges; `folded`
// receives the n_mles EF tables after the last round (2^(n_vars - skips -
// n_rounds + 1) entries each, contiguous) and `sum` the final claim.
std::vector<EF> prove(wl_ctx *c, wh::Transcript &ts, const wl::P2Dev *pdev, const void *tables, bool base,
                      int n_mles, int n_vars, const Program &prog, const Args &a, wl::DBuf<EF> &folded, EF &sum);

}  // namespace wg
