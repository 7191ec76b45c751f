"""crates/sumcheck counterpart: the eq-weighted product sumcheck entry point.

prove_a2b mirrors sumcheck::prove(1, [a, b], a^2*b, 3, [1], Some(z), false,
prover_state, sum, None, Auto{128}, None) (crates/sumcheck/src/prove.rs:22-93)
over a fresh transcript (Poseidon16 challenger, empty domain separator).
"""
import ctypes

import numpy as np

from . import check, lib, u32ptr


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
