"""Batched and served proving (wl_engine_*): several independent Poseidon2 AIR
proofs in flight on one GPU, admitted by device-memory footprint.

The reference proves one table at a time (src/examples/poseidon2.rs:153-174);
this module has no counterpart there.  Each job runs the single-proof path, so
its proof is the one `AirTable.prove` returns for the same inputs.
"""
import ctypes
import threading

import numpy as np

from . import WhirlawayError, check, lib, u32ptr
from .air import AirSettings


class _Settings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("security_bits", "folding_factor_first", "folding_factor",
                                               "log_inv_rate", "univariate_skips",
                                               "rs_domain_initial_reduction", "soundness")]


class _Job(ctypes.Structure):
    _fields_ = [("log_n", ctypes.c_uint32),
                ("air_rc", ctypes.POINTER(ctypes.c_uint32)),
                ("inputs", ctypes.POINTER(ctypes.c_uint32)),
                ("p16", ctypes.POINTER(ctypes.c_uint32)),
                ("p24", ctypes.POINTER(ctypes.c_uint32)),
                ("settings", _Settings)]


def job_footprint(log_n, settings=None):
    """Device bytes reserved for a proof of 2^log_n permutations under `settings`."""
    st = (settings or AirSettings()).c_struct()
    return int(lib().wl_job_footprint(log_n, ctypes.byref(st)))


class Engine:
    """wl_engine: n_slots worker threads, one HIP stream and pool each."""

    def __init__(self, device=0, n_slots=4, mem_budget=0):
        self.ptr = lib().wl_engine_create(device, n_slots, mem_budget)
        if not self.ptr:
            raise WhirlawayError(lib().wl_last_error().decode())
        self._keep = {}
        self._tls = threading.local()

    def submit(self, log_n, air_rc, inputs, p16, p24, settings=None):
        """Queue one proof; returns a ticket.  The arrays are kept alive until wait()."""
        s = settings or AirSettings()
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (air_rc, inputs, p16, p24)]
        if arrs[1].size != 16 << log_n:
            raise ValueError("inputs must hold 16 x 2^log_n values")
        job = _Job(log_n, u32ptr(arrs[0]), u32ptr(arrs[1]), u32ptr(arrs[2]), u32ptr(arrs[3]),
                   _Settings(*s.fields()))
        t = ctypes.c_uint64(0)
        check(lib().wl_engine_submit(self.ptr, ctypes.byref(job), ctypes.byref(t)))
        self._keep[t.value] = arrs
        return t.value

    def done(self, ticket):
        return lib().wl_engine_poll(self.ptr, ticket) == 1

    def wait(self, ticket):
        """Returns (proof_data, {"queued_ms", "run_ms", "latency_ms"})."""
        # per-thread output buffer kept across calls (a fresh array per proof costs
        # its page faults every time); the returned proof is a copy
        out = getattr(self._tls, "out", None)
        if out is None:
            out = self._tls.out = np.empty(1 << 18, dtype=np.uint32)
        n = ctypes.c_size_t(0)
        times = (ctypes.c_double * 3)()
        while True:
            rc = lib().wl_engine_wait(self.ptr, ticket, u32ptr(out), out.size, ctypes.byref(n), times)
            if rc == 3 and n.value > out.size:  # WL_ERR_BUFFER
                out = self._tls.out = np.empty(n.value, dtype=np.uint32)
                continue
            break
        self._keep.pop(ticket, None)
        check(rc)
        return out[: n.value].copy(), {"queued_ms": times[0], "run_ms": times[1], "latency_ms": times[2]}

    def stats(self):
        out = (ctypes.c_uint64 * 4)()
        lib().wl_engine_stats(self.ptr, out)
        return {"completed": out[0], "peak_running": out[1], "peak_reserved": out[2], "budget": out[3]}

    def close(self):
        if self.ptr:
            lib().wl_engine_destroy(self.ptr)
            self.ptr = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
