// devts.h -- the Fiat-Shamir transcript on the device for pipelined sumcheck
// rounds (host interface of k_devts.hip).
//
// A round's reduction kernel can absorb the round polynomial into a device copy
// of the duplex challenger (DuplexChallenger<F, Poseidon2-16, 16, 8>) and sample
// the next challenge itself, so rounds without grinding run back to back on the
// stream instead of waiting for the host.  The host uploads the sponge before the
// rounds and reads it back, with the proof words and challenges, after them.
#pragma once
#include "kernels.h"

namespace wk {

// the duplex challenger's state (Montgomery form); n_in < 8, n_out <= 8
struct DevSponge {
    F st[16];
    F inb[8];
    F outb[8];
    uint32_t n_in, n_out, pad[2];
};

constexpr int DTS_MAX_ROUNDS = 32;

// Zerocheck rounds 1 .. (crates/sumcheck/src/prove.rs:138-253 with the eq factor
// and skips = 1): round j's values at z = 0, 2, 3 (block partials) become
// e_z = v_z * missing, e_1 = (sum - e_0 (1 - r_j)) / r_j, the cubic through them
// times (1 - r_j) + (2 r_j - 1) X, absorbed (5 EF); the challenge c_j is sampled,
// sum = p(c_j), missing *= (1 - r_j)(1 - c_j) + r_j c_j.
struct ZcRounds {
    DevSponge sp;
    EF sum, missing;
    EF r[DTS_MAX_ROUNDS], rinv[DTS_MAX_ROUNDS];  // zerocheck challenges r_j and their inverses
    EF chal[DTS_MAX_ROUNDS];                      // c_j at index j
    uint32_t words[40 * DTS_MAX_ROUNDS];          // round j's coefficients at 40 (j - 1), canonical
};
// Lagrange basis on the nodes 0..3: b[4 i + d] is coefficient d of L_i
struct Basis4 { F b[16]; };

void zc_ts_round(const EF *parts, int nb, ZcRounds *z, int j, const Basis4 &basis, const void *pc, hipStream_t s);

}  // namespace wk
