"""MI355X-native Whirlaway prover hot path.

The product is the HIP/C++ library ``libwhirlaway_hip.so`` (C ABI in
``include/whirlaway_hip.h``); this package is a thin ctypes mirror of the
reference's Rust API (``crates/air``: AirSettings, AirTable::prove;
``crates/sumcheck``: prove; ``crates/utils``: folds) for tests and the
benchmark.  There is no CPU fallback: importing a compute entry point without
the built library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwhirlaway_hip.so")
P = 0x7F000001

_lib = None


class WhirlawayError(RuntimeError):
    pass


def lib():
    """Load the HIP extension; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WhirlawayError(
                f"{LIB_PATH} is missing: build it with `python -m whirlaway_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        szp = ctypes.POINTER(ctypes.c_size_t)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.wl_last_error.restype = ctypes.c_char_p
        L.wl_device_count.restype = ctypes.c_int
        L.wl_create.restype = vp
        L.wl_create.argtypes = [ctypes.c_int]
        L.wl_destroy.argtypes = [vp]
        L.wl_set_device_waits.argtypes = [vp, ctypes.c_int]
        L.wl_set_wait_timeout.argtypes = [vp, ctypes.c_double]
        L.wl_air_table_new.restype = vp
        L.wl_air_table_new.argtypes = [vp, u32p, ctypes.c_uint32, ctypes.c_uint32]
        L.wl_air_table_free.argtypes = [vp]
        L.wl_air_table_new_program.restype = vp
        L.wl_air_table_new_program.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32,
                                               ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32]
        L.wl_air_prove.argtypes = [vp, vp, u32p, u32p, u32p, u32p, ctypes.c_size_t, szp]
        L.wl_witness_upload.restype = vp
        L.wl_witness_upload.argtypes = [vp, u32p, ctypes.c_uint32]
        L.wl_witness_free.argtypes = [vp]
        L.wl_air_prove_resident.argtypes = [vp, vp, u32p, u32p, vp, u32p, ctypes.c_size_t, szp]
        L.wl_phase_times.argtypes = [vp, dp, ctypes.c_int]
        L.wl_fold_large.argtypes = [vp, u32p, ctypes.c_int, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p]
        L.wl_fold_small.argtypes = [vp, u32p, ctypes.c_int, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p]
        L.wl_multilinears_linear_combination.argtypes = [vp, u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
        L.wl_mle_evaluate.argtypes = [vp, u32p, ctypes.c_int, ctypes.c_uint32, u32p, u32p]
        L.wl_mle_evaluate_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_uint32, u32p, vp]
        L.wl_fold_chain.argtypes = [vp, u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p, u32p, dp]
        L.wl_sumcheck_a2b.argtypes = [vp, u32p, u32p, u32p, ctypes.c_uint32, u32p, u32p, u32p, u32p,
                                      u32p, ctypes.c_size_t, szp, dp]
        L.wl_poseidon2_permute.argtypes = [vp, u32p, u32p, ctypes.c_int, u32p, ctypes.c_size_t]
        L.wl_rs_commit.argtypes = [vp, u32p, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, u32p, u32p]
        L.wl_p2_bench_inputs.argtypes = [ctypes.c_uint64, ctypes.c_uint32, u32p, u32p, u32p, u32p]
        L.wl_air_generate_witness.restype = vp
        L.wl_air_generate_witness.argtypes = [vp, u32p]
        L.wl_witness_download.argtypes = [vp, u32p]
        u64 = ctypes.c_uint64
        u64p = ctypes.POINTER(u64)
        L.wl_mem_stats.argtypes = [vp, u64p, u64p, ctypes.c_int]
        L.wl_fmul_rate.argtypes = [vp, dp]
        L.wl_p2_conv_check.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.wl_job_footprint.restype = u64
        L.wl_job_footprint.argtypes = [ctypes.c_uint32, vp]
        L.wl_air_settings_validate.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp]
        L.wl_engine_create.restype = vp
        L.wl_engine_create.argtypes = [ctypes.c_int, ctypes.c_int, u64]
        L.wl_engine_destroy.argtypes = [vp]
        L.wl_engine_submit.argtypes = [vp, vp, u64p]
        L.wl_engine_poll.argtypes = [vp, u64]
        L.wl_engine_wait.argtypes = [vp, u64, u32p, ctypes.c_size_t, szp, dp]
        L.wl_engine_stats.argtypes = [vp, u64p]
        L.wl_air_prove_into.argtypes = [vp, vp, u32p, u32p, u32p, vp]
        L.wl_air_prove_resident_into.argtypes = [vp, vp, u32p, u32p, vp, vp]
        L.wl_air_domain_separator.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp, u32p, ctypes.c_size_t, szp]
        L.wl_transcript_new.restype = vp
        L.wl_transcript_new.argtypes = [u32p, u32p, ctypes.c_size_t]
        L.wl_transcript_new_verifier.restype = vp
        L.wl_transcript_new_verifier.argtypes = [u32p, u32p, ctypes.c_size_t, u32p, ctypes.c_size_t]
        L.wl_transcript_free.argtypes = [vp]
        for fn in ("observe", "add_base", "add_ext", "hint", "sample_ext", "next_base", "next_ext", "next_hint"):
            getattr(L, "wl_transcript_" + fn).argtypes = [vp, u32p, ctypes.c_size_t]
        L.wl_transcript_sample_bits.argtypes = [vp, ctypes.c_uint32, u32p]
        L.wl_transcript_grind.argtypes = [vp, vp, ctypes.c_uint32, u32p]
        L.wl_transcript_proof_len.restype = ctypes.c_size_t
        L.wl_transcript_proof_len.argtypes = [vp]
        L.wl_transcript_proof_data.argtypes = [vp, u32p, ctypes.c_size_t]
        L.wl_transcript_export.argtypes = [vp, u32p]
        L.wl_transcript_import.argtypes = [vp, u32p]
        L.wl_transcript_check_pow.argtypes = [vp, ctypes.c_uint32]
        L.wl_transcript_remaining.restype = ctypes.c_size_t
        L.wl_transcript_remaining.argtypes = [vp]
        ui = ctypes.c_uint32
        L.wl_sumcheck_prove.argtypes = [vp, vp, ui, u32p, ui, ctypes.c_int, ui, u32p, ui, ui, u32p, ui, u32p,
                                        ctypes.c_int, u32p, ui, ctypes.c_int, ui, u32p, u32p, u32p, u32p]
        L.wl_air_verify.argtypes = [vp, vp, u32p, u32p, vp, ctypes.POINTER(ctypes.c_int)]
        L.wl_air_verify_batch.argtypes = [vp, vp, u32p, u32p, ctypes.POINTER(u32p), ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.wl_proof_serialized_size.restype = ctypes.c_size_t
        L.wl_proof_serialized_size.argtypes = [ctypes.c_size_t]
        L.wl_proof_serialize.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t, u8p,
                                         ctypes.c_size_t, szp]
        L.wl_proof_deserialize.argtypes = [u8p, ctypes.c_size_t, vp, u32p, u32p, u32p, ctypes.c_size_t, szp]
        L.wl_kernel_timing.argtypes = [vp, ctypes.c_uint32]
        L.wl_kernel_timing_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long), dp, dp]
        L.wl_autotune_mode.argtypes = [ctypes.c_int]
        L.wl_autotune_force.argtypes = [ctypes.c_int] * 5
        L.wl_autotune_entries.restype = ctypes.c_size_t
        L.wl_autotune_entries.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t]
        _lib = L
    return _lib


def u32ptr(a):
    """Pointer to a C-contiguous uint32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"], "expected contiguous uint32"
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def check(rc):
    if rc != 0:
        raise WhirlawayError(lib().wl_last_error().decode() or f"error {rc}")


from .air import AirSettings, AirTable, Context  # noqa: E402
from . import autotune, engine, program, proof, sumcheck, transcript, utils  # noqa: E402,F401
