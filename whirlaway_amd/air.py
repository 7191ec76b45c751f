"""crates/air counterpart: AirSettings and AirTable::prove for the Poseidon2 AIR.

Mirrors crates/air/src/lib.rs:13-41 (AirSettings::new) and
crates/air/src/table.rs:32-56 / prove.rs:49-226 (AirTable::new / prove).  The
reference prove() panics on failure; here a WhirlawayError is raised.
"""
import ctypes
import threading
from dataclasses import dataclass

import numpy as np

from . import check, lib, u32ptr, WhirlawayError

NCOLS = 165  # Poseidon2Air<16, 3, 0, 4, 20> width

# SecurityAssumption (whir-p3), in the reference's order
UNIQUE_DECODING, JOHNSON_BOUND, CAPACITY_BOUND = 0, 1, 2


@dataclass
class AirSettings:
    """AirSettings::new(security_bits, soundness, ConstantFromSecondRound(f0, f),
    log_inv_rate, univariate_skips, initial_domain_reduction) -- crates/air/src/lib.rs:13-41,
    defaults from src/main.rs:16-23."""
    security_bits: int = 128
    folding_factor_first: int = 7
    folding_factor: int = 4
    log_inv_rate: int = 1
    univariate_skips: int = 4
    rs_domain_initial_reduction: int = 5
    soundness: int = CAPACITY_BOUND

    def fields(self):
        return (self.security_bits, self.folding_factor_first, self.folding_factor, self.log_inv_rate,
                self.univariate_skips, self.rs_domain_initial_reduction, self.soundness)

    def c_struct(self):
        return (ctypes.c_uint32 * 7)(*self.fields())

    def validate(self, log_length, n_witness_columns=NCOLS):
        """wl_air_settings_validate: raises WhirlawayError when no valid schedule exists."""
        st = self.c_struct()
        check(lib().wl_air_settings_validate(log_length, n_witness_columns, ctypes.byref(st)))


class Context:
    """One HIP stream + memory pool on one device (wl_create)."""

    def __init__(self, device=0):
        L = lib()
        self.ptr = L.wl_create(device)
        if not self.ptr:
            raise WhirlawayError(L.wl_last_error().decode())

    def close(self):
        if self.ptr:
            lib().wl_destroy(self.ptr)
            self.ptr = None

    def set_device_waits(self, enabled):
        """wl_set_device_waits: enqueue sumcheck rounds ahead of their challenges behind
        device-side waits (the default), or only once the challenge is known."""
        check(lib().wl_set_device_waits(self.ptr, int(bool(enabled))))

    def set_wait_timeout(self, seconds):
        """wl_set_wait_timeout: the bound of each device-side wait (0: the default)."""
        check(lib().wl_set_wait_timeout(self.ptr, float(seconds)))

    def phase_times(self, reset=True):
        out = (ctypes.c_double * 7)()
        lib().wl_phase_times(self.ptr, out, int(reset))
        return list(out)

    def mem_stats(self, reset_peak=False):
        """(bytes held by the pool, peak live bytes since the last reset)."""
        held, peak = ctypes.c_uint64(0), ctypes.c_uint64(0)
        lib().wl_mem_stats(self.ptr, ctypes.byref(held), ctypes.byref(peak), int(reset_peak))
        return held.value, peak.value

    def fmul_rate(self):
        """Measured Montgomery products per second on this context's device."""
        r = ctypes.c_double(0.0)
        check(lib().wl_fmul_rate(self.ptr, ctypes.byref(r)))
        return r.value

    KERNEL_CLASSES = {1: "merkle_leaves", 2: "zerocheck_rounds", 3: "rs_encode_ntt",
                      4: "zerocheck_round0", 5: "grinding", 6: "merkle_nodes"}

    def kernel_timing(self, classes):
        """Start live HIP-event timing of the given kernel classes (wl_kernel_timing)."""
        mask = 0
        for k in classes:
            mask |= 1 << k
        check(lib().wl_kernel_timing(self.ptr, mask))

    def kernel_timing_read(self, kernel_class):
        n = ctypes.c_long(0)
        ms = ctypes.c_double(0)
        work = ctypes.c_double(0)
        check(lib().wl_kernel_timing_read(self.ptr, kernel_class, ctypes.byref(n), ctypes.byref(ms),
                                          ctypes.byref(work)))
        return n.value, ms.value, work.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class AirTable:
    """AirTable::<KoalaBear, EF8, Poseidon2Air>::new(air, log_length, skips, [], 3)."""

    def __init__(self, ctx, round_constants, log_length, univariate_skips=4):
        rc = np.ascontiguousarray(round_constants, dtype=np.uint32)
        if rc.size != 148:
            raise ValueError("Poseidon2 AIR round constants: expected 148 values")
        self.ctx = ctx
        self.log_length = log_length
        self._rc = rc
        # proof output buffer kept across calls: a fresh multi-MiB array per proof
        # costs its page faults on every call (the returned proof is a copy)
        self._out = np.empty(1 << 20, dtype=np.uint32)
        self._out_lock = threading.Lock()
        # ctx None: a table that only verifies (no device)
        self.ptr = lib().wl_air_table_new(ctx.ptr if ctx is not None else None, u32ptr(rc), log_length,
                                          univariate_skips)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())

    def close(self):
        if self.ptr:
            lib().wl_air_table_free(self.ptr)
            self.ptr = None

    def prove_into(self, settings, p16, p24, witness, prover_state):
        """AirTable::prove(settings, merkle_hash, merkle_compress, prover_state, witness)
        (crates/air/src/prove.rs:49-66): appends the proof to a caller's ProverState
        (whirlaway_amd.transcript), continuing its Fiat-Shamir state."""
        p16 = np.ascontiguousarray(p16, dtype=np.uint32)
        p24 = np.ascontiguousarray(p24, dtype=np.uint32)
        st = settings.c_struct()
        L = lib()
        if isinstance(witness, Witness):
            check(L.wl_air_prove_resident_into(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24), witness.ptr,
                                               prover_state.ptr))
        else:
            w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
            if w.size != NCOLS << self.log_length:
                raise ValueError("witness must hold 165 columns of 2^log_length values")
            check(L.wl_air_prove_into(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24), u32ptr(w),
                                      prover_state.ptr))

    def prove(self, settings, p16, p24, witness):
        """Returns proof_data (uint32 canonical field elements).  witness: 165 x 2^n
        column-major (uint32) or a resident Witness."""
        p16 = np.ascontiguousarray(p16, dtype=np.uint32)
        p24 = np.ascontiguousarray(p24, dtype=np.uint32)
        st = settings.c_struct()
        n = ctypes.c_size_t(0)
        L = lib()
        own = self._out_lock.acquire(blocking=False)
        out = self._out if own else np.empty(self._out.size, dtype=np.uint32)
        cap = out.size
        try:
            if isinstance(witness, Witness):
                check(L.wl_air_prove_resident(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24),
                                              witness.ptr, u32ptr(out), cap, ctypes.byref(n)))
            else:
                w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
                if w.size != NCOLS << self.log_length:
                    raise ValueError("witness must hold 165 columns of 2^log_length values")
                check(L.wl_air_prove(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24), u32ptr(w),
                                     u32ptr(out), cap, ctypes.byref(n)))
            return out[: n.value].copy()
        finally:
            if own:
                self._out_lock.release()


class ProgramAirTable:
    """AirTable::<F, EF, A>::new(air, log_length, univariate_skips, preprocessed_columns,
    constraint_degree) (crates/air/src/table.rs:32-56) for an AIR given as a
    ConstraintProgram over [local row, next row] (whirlaway_amd.program)."""

    def __init__(self, ctx, program, log_length, constraint_degree, preprocessed=None, univariate_skips=4):
        pre = None if preprocessed is None else np.ascontiguousarray(preprocessed, dtype=np.uint32)
        n_pre = 0 if pre is None else pre.reshape(-1, 1 << log_length).shape[0]
        if pre is not None and pre.size != n_pre << log_length:
            raise ValueError("preprocessed columns hold 2^log_length values each")
        self.ctx = ctx
        self.log_length = log_length
        self.n_columns = program.width
        self.n_preprocessed = n_pre
        self.n_witness = self.n_columns - n_pre
        self._code = np.ascontiguousarray(program.code(), dtype=np.uint32)
        self.ptr = lib().wl_air_table_new_program(ctx.ptr if ctx is not None else None, self.n_columns, n_pre,
                                                  u32ptr(pre), log_length,
                                                  univariate_skips, u32ptr(self._code), program.n_instr,
                                                  constraint_degree)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())

    def close(self):
        if self.ptr:
            lib().wl_air_table_free(self.ptr)
            self.ptr = None

    def _witness(self, witness):
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
        if w.size != self.n_witness << self.log_length:
            raise ValueError(f"witness must hold {self.n_witness} columns of 2^log_length values")
        return w

    def prove(self, settings, p16, p24, witness):
        """proof_data over a fresh transcript (this build's domain separator)."""
        p16 = np.ascontiguousarray(p16, dtype=np.uint32)
        p24 = np.ascontiguousarray(p24, dtype=np.uint32)
        w = self._witness(witness)
        st = settings.c_struct()
        n = ctypes.c_size_t(0)
        check(lib().wl_air_prove(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24), u32ptr(w), None, 0,
                                 ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint32)
        check(lib().wl_air_prove(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24), u32ptr(w), u32ptr(out),
                                 out.size, ctypes.byref(n)))
        return out

    def prove_into(self, settings, p16, p24, witness, prover_state):
        p16 = np.ascontiguousarray(p16, dtype=np.uint32)
        p24 = np.ascontiguousarray(p24, dtype=np.uint32)
        st = settings.c_struct()
        check(lib().wl_air_prove_into(self.ptr, ctypes.byref(st), u32ptr(p16), u32ptr(p24),
                                      u32ptr(self._witness(witness)), prover_state.ptr))


class Witness:
    """Trace resident in HBM (wl_witness_upload)."""

    def __init__(self, ctx, witness, log_length):
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
        if w.size != NCOLS << log_length:
            raise ValueError("witness must hold 165 columns of 2^log_length values")
        self.log_length = log_length
        self.ptr = lib().wl_witness_upload(ctx.ptr, u32ptr(w), log_length)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())

    def close(self):
        if self.ptr:
            lib().wl_witness_free(self.ptr)
            self.ptr = None


def bench_inputs(seed, log_n, with_inputs=True):
    """src/examples/poseidon2.rs:104-151: (air round constants, inputs, p16, p24)."""
    rc = np.zeros(148, dtype=np.uint32)
    p16 = np.zeros(148, dtype=np.uint32)
    p24 = np.zeros(215, dtype=np.uint32)
    inputs = np.zeros(16 << log_n, dtype=np.uint32) if with_inputs else None
    check(lib().wl_p2_bench_inputs(seed, log_n, u32ptr(rc), u32ptr(inputs), u32ptr(p16), u32ptr(p24)))
    return rc, inputs, p16, p24


def generate_witness(table, inputs):
    """generate_trace_rows on the GPU into a resident Witness."""
    inputs = np.ascontiguousarray(inputs, dtype=np.uint32)
    if inputs.size != 16 << table.log_length:
        raise ValueError("inputs must hold 16 x 2^log_length values")
    ptr = lib().wl_air_generate_witness(table.ptr, u32ptr(inputs))
    if not ptr:
        raise WhirlawayError(lib().wl_last_error().decode())
    w = Witness.__new__(Witness)
    w.ptr = ptr
    w.log_length = table.log_length
    return w


def download_witness(w):
    out = np.zeros(NCOLS << w.log_length, dtype=np.uint32)
    check(lib().wl_witness_download(w.ptr, u32ptr(out)))
    return out
