// tune.h -- autotuner of the commitment kernels' launch shapes.
//
// Each commitment kernel has launch variants that compute the same values: the
// split L = LA + LB of the two-pass Reed-Solomon encoding, the lane mapping and
// block size of Merkle leaf hashing, and the schedule of the Merkle tree levels
// (one thread, a lane pair or 16 lanes per compression).  Which one is fastest
// depends on the shape (a commit of 2^12 rows is latency-bound, one of 2^19 rows
// issue-bound) and on the machine.  The tuner measures the variants in place: the
// first calls of a (kernel, shape) on a device each run the next candidate,
// timed by two events around that call on the caller's stream, until every
// candidate has two samples; the fastest then serves every later call, from any
// context, through a process-wide table.  Timing the variants alone, back to back,
// ranks them differently from how they run inside a proof on this GPU, so the
// calls themselves are the measurement.  Variants give identical outputs, so a
// proof does not depend on the choice or on the exploration.
#pragma once
#include "ctx.h"

namespace wt {
using kb::F;

enum Kind { TK_NTT = 0, TK_LEAVES = 1, TK_TREE = 2, TK_KINDS = 3 };

// the split variant rs_encode uses for 2^L rows on c's device (no tuning)
int ntt_variant(wl_ctx *c, int L);
// wk::rs_encode with the tuned split; returns la
int rs_encode(wl_ctx *c, const F *cols, int n_cols, int n_nonzero, int log_m, int L, F *cw, F *tmp);
// wk::merkle_leaves / merkle_nodes with the tuned launch variant
void merkle_leaves(wl_ctx *c, const F *cw, int width, int nnz, int log_rows, int la, const void *pc, F *nodes);
void merkle_nodes(wl_ctx *c, int log_rows, const void *pc, F *nodes);

}  // namespace wt
