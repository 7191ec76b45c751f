"""crates/sumcheck counterpart: sumcheck::prove for any computation, and the
eq-weighted a^2 b product sumcheck of the kernel micro-benchmark.

prove() mirrors sumcheck::prove(skips, multilinears, computation, degree,
batching_scalars, eq_factor, is_zerofier, prover_state, sum, n_rounds, grinding,
missing_mul_factor) (crates/sumcheck/src/prove.rs:22-35) and returns
(challenges, folded multilinears, final sum); the computation is a
ConstraintProgram (whirlaway_amd.program) and prover_state a ProverState
(whirlaway_amd.transcript).  prove_a2b is the fused kernel path of
sumcheck::prove(1, [a, b], a^2*b, 3, [1], Some(z), false, prover_state, sum, None,
Auto{128}, None) over a fresh transcript (Poseidon16 challenger, empty domain
separator).
"""
import ctypes

import numpy as np

from . import check, lib, u32ptr

# SumcheckGrinding (crates/sumcheck/src/lib.rs:14-31)
GRIND_NONE, GRIND_AUTO, GRIND_CUSTOM = 0, 1, 2


def _ef_or_none(v):
    if v is None:
        return None
    a = np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)
    if a.size % 8:
        raise ValueError("extension-field values are 8 words each")
    return a


def prove(ctx, skips, multilinears, computation, degree, batching_scalars, eq_factor, is_zerofier, prover_state,
          sum, n_rounds=None, grinding=(GRIND_AUTO, 128), missing_mul_factor=None, is_ef=False):
    """multilinears: (n_mles, 2^n) base-field values, or (n_mles, 2^n, 8) with is_ef.
    Returns (challenges (n_rounds, 8), folded (n_mles, 2^(n - skips - n_rounds + 1), 8),
    final sum (8,))."""
    m = np.ascontiguousarray(multilinears, dtype=np.uint32)
    n_mles = m.shape[0]
    N = m.shape[1]
    n = N.bit_length() - 1
    if N != 1 << n or (is_ef and (m.ndim != 3 or m.shape[2] != 8)):
        raise ValueError("multilinears must hold 2^n values each")
    code = np.ascontiguousarray(computation.code(), dtype=np.uint32)
    bat = _ef_or_none(batching_scalars)
    eqf = _ef_or_none(eq_factor)
    if eqf is not None and eqf.size != 8 * (n - skips + 1):
        raise ValueError("eq_factor holds n - skips + 1 extension values")
    mmf = _ef_or_none(missing_mul_factor)
    s = np.ascontiguousarray(sum, dtype=np.uint32).reshape(-1)
    nr = n_rounds or (n - skips + 1)
    ch = np.zeros(8 * nr, dtype=np.uint32)
    fl = (1 << (n - skips - nr + 1))
    folded = np.zeros(n_mles * fl * 8, dtype=np.uint32)
    fs = np.zeros(8, dtype=np.uint32)
    kind, value = grinding
    check(lib().wl_sumcheck_prove(ctx.ptr, prover_state.ptr, skips, u32ptr(m.reshape(-1)), n_mles, int(is_ef), n,
                                  u32ptr(code), computation.n_instr, degree, u32ptr(bat),
                                  0 if bat is None else bat.size // 8, u32ptr(eqf), int(is_zerofier), u32ptr(s),
                                  n_rounds or 0, kind, value, u32ptr(mmf), u32ptr(ch), u32ptr(folded), u32ptr(fs)))
    return ch.reshape(nr, 8), folded.reshape(n_mles, fl, 8), fs


def prove_a2b(ctx, p16, a, b, z):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    z = np.ascontiguousarray(z, dtype=np.uint32).reshape(-1)
    n = int(a.size).bit_length() - 1
    assert a.size == 1 << n and b.size == a.size and z.size == 8 * n
    p16 = np.ascontiguousarray(p16, dtype=np.uint32)
    ch = np.zeros(8 * n, dtype=np.uint32)
    fe = np.zeros(16, dtype=np.uint32)
    fs = np.zeros(8, dtype=np.uint32)
    cap = 64 * n + 64
    proof = np.zeros(cap, dtype=np.uint32)
    ln = ctypes.c_size_t(0)
    ms = ctypes.c_double(0)
    check(lib().wl_sumcheck_a2b(ctx.ptr, u32ptr(p16), u32ptr(a), u32ptr(b), n, u32ptr(z),
                                u32ptr(ch), u32ptr(fe), u32ptr(fs), u32ptr(proof), cap,
                                ctypes.byref(ln), ctypes.byref(ms)))
    return {"challenges": ch.reshape(n, 8), "final_evals": fe.reshape(2, 8), "final_sum": fs,
            "proof": proof[: ln.value].copy(), "ms": ms.value}
