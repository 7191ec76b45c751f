"""Proof verification and the proof wire format.

verify() is AirTable::verify(settings, merkle_hash, merkle_compress,
verifier_state, log_length) (crates/air/src/verify.rs:55-249) for a table of
either kind; verify_batch() checks many proofs of one table on host threads.
serialize()/deserialize() carry proof_data (src/examples/poseidon2.rs:178) with
its table shape and settings at 31 bits per field element.
"""
import ctypes

import numpy as np

from . import WhirlawayError, check, lib, u32ptr
from .air import AirSettings

# AirVerifError kinds (crates/air/src/verify.rs:26-34)
VERIFY_ERRORS = {1: "InvalidPcsCommitment", 2: "InvalidPcsOpening", 3: "Fs", 4: "Sumcheck",
                 5: "InvalidBoundaryCondition", 6: "SumMismatch"}


class VerificationError(WhirlawayError):
    def __init__(self, kind):
        super().__init__(VERIFY_ERRORS.get(kind, f"error {kind}"))
        self.kind = kind


def verify(table, settings, p16, p24, verifier_state):
    """Raises VerificationError (with .kind) when the proof is rejected."""
    st = settings.c_struct()
    err = ctypes.c_int(0)
    rc = lib().wl_air_verify(table.ptr, ctypes.byref(st), u32ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                             u32ptr(np.ascontiguousarray(p24, dtype=np.uint32)), verifier_state.ptr, ctypes.byref(err))
    if rc == 4:  # WL_ERR_VERIFY
        raise VerificationError(err.value)
    check(rc)


def verify_batch(table, settings, p16, p24, proofs, n_threads=8):
    """Proofs made by AirTable.prove (this build's domain separator); returns a list of
    0 (accepted) or the AirVerifError kind per proof."""
    arrs = [np.ascontiguousarray(p, dtype=np.uint32) for p in proofs]
    ptrs = (ctypes.POINTER(ctypes.c_uint32) * len(arrs))(*[u32ptr(a) for a in arrs])
    lens = (ctypes.c_size_t * len(arrs))(*[a.size for a in arrs])
    res = (ctypes.c_int * max(1, len(arrs)))()
    st = settings.c_struct()
    check(lib().wl_air_verify_batch(table.ptr, ctypes.byref(st), u32ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                                    u32ptr(np.ascontiguousarray(p24, dtype=np.uint32)), ptrs, lens, len(arrs),
                                    n_threads, res))
    return list(res[:len(arrs)])


def serialize(proof, log_length, n_witness_columns, settings=None):
    p = np.ascontiguousarray(proof, dtype=np.uint32)
    st = (settings or AirSettings()).c_struct()
    n = ctypes.c_size_t(0)
    check(lib().wl_proof_serialize(ctypes.byref(st), log_length, n_witness_columns, u32ptr(p), p.size, None, 0,
                                   ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint8)
    check(lib().wl_proof_serialize(ctypes.byref(st), log_length, n_witness_columns, u32ptr(p), p.size,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), out.size, ctypes.byref(n)))
    return out.tobytes()


def deserialize(data):
    """Returns (proof words, log_length, n_witness_columns, AirSettings)."""
    buf = np.frombuffer(data, dtype=np.uint8).copy()
    bp = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    st = (ctypes.c_uint32 * 7)()
    ll, nw, n = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_size_t(0)
    check(lib().wl_proof_deserialize(bp, buf.size, st, ctypes.byref(ll), ctypes.byref(nw), None, 0,
                                     ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint32)
    check(lib().wl_proof_deserialize(bp, buf.size, st, ctypes.byref(ll), ctypes.byref(nw), u32ptr(out), out.size,
                                     ctypes.byref(n)))
    return out, ll.value, nw.value, AirSettings(*list(st))
