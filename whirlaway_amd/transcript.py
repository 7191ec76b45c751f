"""whir-p3 ProverState / VerifierState counterpart (wl_transcript).

The reference builds the transcript from a DomainSeparator and a
DuplexChallenger<F, Poseidon2-16, 16, 8> (src/examples/poseidon2.rs:155-163) and
passes it to AirTable::prove / verify (crates/air/src/prove.rs:54,
verify.rs:60) and sumcheck::prove (crates/sumcheck/src/prove.rs:29).  Method
names follow whir-p3's: add_base_scalars, add_extension_scalars, sample,
pow_grinding, next_*_vec, check_pow_grinding.
"""
import ctypes

import numpy as np

from . import WhirlawayError, check, lib, u32ptr

STATE_WORDS = 34


def _arr(v):
    return np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)


class _Base:
    ptr = None

    def close(self):
        if self.ptr:
            lib().wl_transcript_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def observe(self, v):
        """FieldChallenger::observe_slice (nothing is appended to the proof)."""
        a = _arr(v)
        check(lib().wl_transcript_observe(self.ptr, u32ptr(a), a.size))

    def sample(self, n=1):
        """n extension-field challenges, (n, 8) canonical."""
        out = np.zeros(8 * n, dtype=np.uint32)
        check(lib().wl_transcript_sample_ext(self.ptr, u32ptr(out), n))
        return out.reshape(n, 8)

    def sample_bits(self, bits):
        out = ctypes.c_uint32(0)
        check(lib().wl_transcript_sample_bits(self.ptr, bits, ctypes.byref(out)))
        return out.value

    def export_state(self):
        out = np.zeros(STATE_WORDS, dtype=np.uint32)
        check(lib().wl_transcript_export(self.ptr, u32ptr(out)))
        return out

    def import_state(self, state):
        a = _arr(state)
        if a.size != STATE_WORDS:
            raise ValueError(f"the sponge state has {STATE_WORDS} words")
        check(lib().wl_transcript_import(self.ptr, u32ptr(a)))

    def proof_data(self):
        n = lib().wl_transcript_proof_len(self.ptr)
        out = np.zeros(n, dtype=np.uint32)
        if n:
            check(lib().wl_transcript_proof_data(self.ptr, u32ptr(out), n))
        return out


class ProverState(_Base):
    """DomainSeparator::to_prover_state: the challenger over Poseidon2-16 constants
    p16 (as AirTable.prove's) after observing the domain separator's words."""

    def __init__(self, p16, domsep=()):
        p = _arr(p16)
        d = _arr(domsep)
        self.ptr = lib().wl_transcript_new(u32ptr(p), u32ptr(d), d.size)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())

    def add_base_scalars(self, v):
        a = _arr(v)
        check(lib().wl_transcript_add_base(self.ptr, u32ptr(a), a.size))

    def add_extension_scalars(self, v):
        a = _arr(v)
        if a.size % 8:
            raise ValueError("extension scalars are 8 words each")
        check(lib().wl_transcript_add_ext(self.ptr, u32ptr(a), a.size // 8))

    def hint_base_scalars(self, v):
        a = _arr(v)
        check(lib().wl_transcript_hint(self.ptr, u32ptr(a), a.size))

    def pow_grinding(self, bits, ctx=None):
        """Appends the smallest witness; on ctx's device when given, else on the host."""
        w = ctypes.c_uint32(0)
        check(lib().wl_transcript_grind(self.ptr, ctx.ptr if ctx is not None else None, bits, ctypes.byref(w)))
        return w.value if bits else None


class VerifierState(_Base):
    """DomainSeparator::to_verifier_state over proof_data."""

    def __init__(self, p16, domsep, proof):
        p, d, pr = _arr(p16), _arr(domsep), _arr(proof)
        self.ptr = lib().wl_transcript_new_verifier(u32ptr(p), u32ptr(d), d.size, u32ptr(pr), pr.size)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())

    def next_base_scalars_vec(self, n):
        out = np.zeros(n, dtype=np.uint32)
        check(lib().wl_transcript_next_base(self.ptr, u32ptr(out), n))
        return out

    def next_extension_scalars_vec(self, n):
        out = np.zeros(8 * n, dtype=np.uint32)
        check(lib().wl_transcript_next_ext(self.ptr, u32ptr(out), n))
        return out.reshape(n, 8)

    def receive_hint_base_scalars(self, n):
        out = np.zeros(n, dtype=np.uint32)
        check(lib().wl_transcript_next_hint(self.ptr, u32ptr(out), n))
        return out

    def check_pow_grinding(self, bits):
        check(lib().wl_transcript_check_pow(self.ptr, bits))

    def remaining(self):
        return int(lib().wl_transcript_remaining(self.ptr))


def air_domain_separator(log_length, settings=None, n_witness_columns=165):
    """The domain-separator words wl_air_prove observes first (this build's
    encoding of DomainSeparator::commit_statement + add_whir_proof)."""
    from .air import AirSettings
    st = (settings or AirSettings()).c_struct()
    n = ctypes.c_size_t(0)
    check(lib().wl_air_domain_separator(log_length, n_witness_columns, ctypes.byref(st), None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint32)
    check(lib().wl_air_domain_separator(log_length, n_witness_columns, ctypes.byref(st), u32ptr(out), out.size,
                                        ctypes.byref(n)))
    return out
