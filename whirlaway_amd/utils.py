"""crates/utils counterpart: multilinear folds, linear combination, evaluation,
Poseidon2 and the RS/Merkle core.

fold_multilinear_in_small_field / _in_large_field mirror
crates/utils/src/multilinear.rs:10-52 / 90-123, multilinears_linear_combination
:125-147; mle_evaluate is EvaluationsList::evaluate; fold_chain is the kernel
micro-benchmark of benches/folding_benchmarks.rs.
"""
import ctypes

import numpy as np

from . import check, lib, u32ptr


def fold_multilinear_in_large_field(ctx, m, scalars, is_ef=False):
    m = np.ascontiguousarray(m, dtype=np.uint32)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint32).reshape(-1)
    k = (scalars.size // 8).bit_length() - 1
    n_elems = m.size // (8 if is_ef else 1)
    log_n = n_elems.bit_length() - 1
    out = np.zeros((n_elems >> k) * 8, dtype=np.uint32)
    check(lib().wl_fold_large(ctx.ptr, u32ptr(m), int(is_ef), log_n, u32ptr(scalars), k, u32ptr(out)))
    return out


def fold_multilinear_in_small_field(ctx, m, scalars, is_ef=False):
    """out[i] = sum_j scalars[j] m[i + j N/K]: K base-field scalars, F or EF table."""
    m = np.ascontiguousarray(m, dtype=np.uint32)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint32).reshape(-1)
    k = scalars.size.bit_length() - 1
    if scalars.size != 1 << k:
        raise ValueError("the number of scalars must be a power of two")
    w = 8 if is_ef else 1
    n_elems = m.size // w
    out = np.zeros((n_elems >> k) * w, dtype=np.uint32)
    check(lib().wl_fold_small(ctx.ptr, u32ptr(m), int(is_ef), n_elems.bit_length() - 1, u32ptr(scalars), k,
                              u32ptr(out)))
    return out


def multilinears_linear_combination(ctx, cols, scalars):
    """sum_c scalars[c] cols[c]: cols (n_cols, 2^n) base field, scalars (n_cols, 8) EF."""
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    n_cols, N = cols.shape
    sc = np.ascontiguousarray(scalars, dtype=np.uint32).reshape(-1)
    if sc.size != 8 * n_cols:
        raise ValueError("one EF scalar per column")
    out = np.zeros(N * 8, dtype=np.uint32)
    check(lib().wl_multilinears_linear_combination(ctx.ptr, u32ptr(cols), n_cols, N.bit_length() - 1, u32ptr(sc),
                                                   u32ptr(out)))
    return out


def mle_evaluate(ctx, evals, point, is_ef=False):
    """EvaluationsList::evaluate: the multilinear extension at an EF point (n x 8)."""
    evals = np.ascontiguousarray(evals, dtype=np.uint32)
    pt = np.ascontiguousarray(point, dtype=np.uint32).reshape(-1)
    n = pt.size // 8
    if evals.size != (1 << n) * (8 if is_ef else 1):
        raise ValueError("2^n evaluations for an n-coordinate point")
    out = np.zeros(8, dtype=np.uint32)
    check(lib().wl_mle_evaluate(ctx.ptr, u32ptr(evals), int(is_ef), n, u32ptr(pt), u32ptr(out)))
    return out


def fold_chain(ctx, m, challenges, all_tables=False):
    m = np.ascontiguousarray(m, dtype=np.uint32)
    ch = np.ascontiguousarray(challenges, dtype=np.uint32).reshape(-1)
    n_folds = ch.size // 8
    log_n = m.size.bit_length() - 1
    final = np.zeros((1 << (log_n - n_folds)) * 8, dtype=np.uint32)
    total = sum(1 << (log_n - 1 - i) for i in range(n_folds))
    allt = np.zeros(total * 8, dtype=np.uint32) if all_tables else None
    ms = ctypes.c_double(0)
    check(lib().wl_fold_chain(ctx.ptr, u32ptr(m), log_n, u32ptr(ch), n_folds, u32ptr(allt),
                              u32ptr(final), ctypes.byref(ms)))
    return final, allt, ms.value


def poseidon2_permute(ctx, p16, p24, width, states):
    st = np.ascontiguousarray(states, dtype=np.uint32).copy()
    n = st.size // width
    check(lib().wl_poseidon2_permute(ctx.ptr, u32ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                                     u32ptr(np.ascontiguousarray(p24, dtype=np.uint32)), width,
                                     u32ptr(st), n))
    return st


def rs_commit(ctx, p16, p24, cols, log_rows, want_codeword=False):
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    n_cols = cols.shape[0]
    log_m = cols.shape[1].bit_length() - 1
    root = np.zeros(8, dtype=np.uint32)
    cw = np.zeros(n_cols << log_rows, dtype=np.uint32) if want_codeword else None
    check(lib().wl_rs_commit(ctx.ptr, u32ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                             u32ptr(np.ascontiguousarray(p24, dtype=np.uint32)), u32ptr(cols),
                             n_cols, log_m, log_rows, u32ptr(cw), u32ptr(root)))
    return root, cw
