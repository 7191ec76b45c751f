"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the test suite, the smoke entry and the benchmark's CPU-baseline leg load
this module; the product (whirlaway_amd) never does.  build() compiles
oracle/build/liboracle.so with the Makefile next to this file.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
P = 0x7F000001

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        vp = ctypes.c_void_p
        L.or_api_bench_new.restype = vp
        L.or_api_bench_new.argtypes = [ctypes.c_int, ctypes.c_uint64]
        L.or_api_bench_free.argtypes = [vp]
        L.or_api_bench_trace.restype = u32p
        L.or_api_bench_trace.argtypes = [vp]
        L.or_api_bench_inputs.restype = u32p
        L.or_api_bench_inputs.argtypes = [vp]
        L.or_api_bench_air_consts.argtypes = [vp, u32p]
        L.or_api_bench_p2_consts.restype = ctypes.c_int
        L.or_api_bench_p2_consts.argtypes = [vp, ctypes.c_int, u32p]
        L.or_api_permute.argtypes = [vp, ctypes.c_int, u32p]
        L.or_api_hash_leaf.argtypes = [vp, u32p, ctypes.c_size_t, u32p]
        L.or_api_compress.argtypes = [vp, u32p, u32p, u32p]
        L.or_api_prove.restype = ctypes.c_size_t
        L.or_api_prove.argtypes = [vp] + [ctypes.c_int] * 7 + [ctypes.POINTER(u32p)]
        L.or_api_verify.restype = ctypes.c_int
        L.or_api_verify.argtypes = [vp] + [ctypes.c_int] * 7 + [u32p, ctypes.c_size_t]
        L.or_api_free.argtypes = [vp]
        L.or_api_whir_config.restype = ctypes.c_int
        L.or_api_whir_config.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(ctypes.c_int)]
        L.or_api_ef_mul.argtypes = [u32p, u32p, u32p]
        L.or_api_ef_inv.argtypes = [u32p, u32p]
        L.or_api_two_adic_generator.restype = ctypes.c_uint32
        L.or_api_two_adic_generator.argtypes = [ctypes.c_int]
        L.or_api_rng_u32.argtypes = [ctypes.c_uint64, u32p, ctypes.c_size_t]
        L.or_api_rng_f.argtypes = [ctypes.c_uint64, u32p, ctypes.c_size_t]
        L.or_api_eval_eq.argtypes = [u32p, ctypes.c_int, u32p]
        L.or_api_mle_eval_f.argtypes = [u32p, ctypes.c_int, u32p, u32p]
        L.or_api_mle_eval_ef.argtypes = [u32p, ctypes.c_int, u32p, u32p]
        L.or_api_air_eval_f.argtypes = [vp, u32p, u32p, u32p]
        L.or_api_air_eval_ef.argtypes = [vp, u32p, u32p, u32p]
        L.or_api_selectors_at_ef.argtypes = [ctypes.c_int, u32p, u32p]
        L.or_api_lagrange.argtypes = [u32p, u32p, ctypes.c_int, u32p]
        L.or_api_fold_f_large.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.or_api_fold_ef_large.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.or_api_sc_a2b.restype = ctypes.c_size_t
        L.or_api_sc_a2b.argtypes = [vp, u32p, u32p, ctypes.c_int, u32p, u32p, u32p, u32p,
                                    ctypes.POINTER(u32p)]
        L.or_api_fold_small.argtypes = [u32p, ctypes.c_int, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.or_api_lincomb.argtypes = [u32p, ctypes.c_int, ctypes.c_int, u32p, u32p]
        szt = ctypes.c_size_t
        L.or_api_ts_new.restype = vp
        L.or_api_ts_new.argtypes = [u32p, u32p, szt]
        L.or_api_ts_free.argtypes = [vp]
        for fn in ("observe", "add_base", "hint"):
            getattr(L, "or_api_ts_" + fn).argtypes = [vp, u32p, szt]
        L.or_api_ts_sample_ef.argtypes = [vp, u32p, szt]
        L.or_api_ts_sample_bits.restype = ctypes.c_uint32
        L.or_api_ts_sample_bits.argtypes = [vp, ctypes.c_int]
        L.or_api_ts_pow.argtypes = [vp, ctypes.c_int]
        L.or_api_ts_len.restype = szt
        L.or_api_ts_len.argtypes = [vp]
        L.or_api_ts_data.restype = u32p
        L.or_api_ts_data.argtypes = [vp]
        L.or_api_prove_ts.argtypes = [vp] + [ctypes.c_int] * 7 + [vp]
        L.or_api_domsep.restype = szt
        L.or_api_domsep.argtypes = [ctypes.c_int] * 8 + [u32p, szt]
        ci = ctypes.c_int
        L.or_api_sc_prove.argtypes = [vp, ci, u32p, ci, ci, ci, u32p, ci, ci, u32p, u32p, ci, u32p, ci, ci, ci,
                                      u32p, u32p, u32p, u32p]
        L.or_api_prog_check.restype = ci
        L.or_api_prog_check.argtypes = [u32p, ci, ci]
        L.or_api_prog_eval.argtypes = [u32p, ci, ci, ci, u32p, u32p, u32p]
        L.or_api_gair_prove.argtypes = [ci, ci, ci, u32p, u32p, ci, ci] + [ci] * 7 + [u32p, u32p, u32p, vp]
        L.or_api_gair_verify.restype = ci
        L.or_api_gair_verify.argtypes = [ci, ci, ci, u32p, u32p, ci, ci] + [ci] * 7 + [u32p, u32p, u32p, szt, u32p,
                                                                                        szt]
        L.or_api_gair_domsep.restype = szt
        L.or_api_gair_domsep.argtypes = [ci] * 9 + [u32p, szt]
        L.or_ntt.argtypes = [u32p, ctypes.c_int]
        L.or_mobius.argtypes = [u32p, ctypes.c_int]
        L.or_chacha_block.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, u32p]
        L.or_chacha_block_ietf.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_int, u32p]
        L.or_log2_up.restype = ctypes.c_size_t
        L.or_log2_up.argtypes = [ctypes.c_size_t]
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


# SecurityAssumption codes in the reference's order
UNIQUE_DECODING, JOHNSON_BOUND, CAPACITY_BOUND = 0, 1, 2
DEFAULT = dict(sec=128, f0=7, f=4, rate=1, skips=4, rs0=5, sa=CAPACITY_BOUND)


class Bench:
    """src/examples/poseidon2.rs:85-151: seeded constants, inputs and trace."""

    def __init__(self, log_n, seed=0):
        self.log_n = log_n
        self.h = lib().or_api_bench_new(log_n, seed)

    def close(self):
        if self.h:
            lib().or_api_bench_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def trace(self):
        N = 1 << self.log_n
        p = lib().or_api_bench_trace(self.h)
        return np.ctypeslib.as_array(p, shape=(165 * N,)).copy()

    def inputs(self):
        N = 1 << self.log_n
        p = lib().or_api_bench_inputs(self.h)
        return np.ctypeslib.as_array(p, shape=(16 * N,)).copy()

    def air_consts(self):
        out = np.zeros(148, dtype=np.uint32)
        lib().or_api_bench_air_consts(self.h, ptr(out))
        return out

    def p2_consts(self, width):
        out = np.zeros(4 * 24 * 2 + 23, dtype=np.uint32)
        k = lib().or_api_bench_p2_consts(self.h, width, ptr(out))
        return out[:k].copy()

    def permute(self, width, state):
        s = np.ascontiguousarray(state, dtype=np.uint32).copy()
        lib().or_api_permute(self.h, width, ptr(s))
        return s

    def prove(self, **kw):
        s = dict(DEFAULT, **kw)
        p = ctypes.POINTER(ctypes.c_uint32)()
        n = lib().or_api_prove(self.h, s["sec"], s["f0"], s["f"], s["rate"], s["skips"], s["rs0"],
                               s["sa"], ctypes.byref(p))
        arr = np.ctypeslib.as_array(p, shape=(n,)).copy()
        lib().or_api_free(p)
        return arr

    def verify(self, proof, **kw):
        s = dict(DEFAULT, **kw)
        pr = np.ascontiguousarray(proof, dtype=np.uint32)
        return lib().or_api_verify(self.h, s["sec"], s["f0"], s["f"], s["rate"], s["skips"],
                                   s["rs0"], s["sa"], ptr(pr), pr.size)

    def prove_into(self, ts, **kw):
        """AirTable::prove appended to the Transcript ts (its Poseidon2-16 must be this
        bench's)."""
        s = dict(DEFAULT, **kw)
        lib().or_api_prove_ts(self.h, s["sec"], s["f0"], s["f"], s["rate"], s["skips"], s["rs0"], s["sa"], ts.h)

    def sc_a2b(self, a, b, z):
        n = a.size.bit_length() - 1
        ch = np.zeros(8 * n, dtype=np.uint32)
        fe = np.zeros(16, dtype=np.uint32)
        fs = np.zeros(8, dtype=np.uint32)
        p = ctypes.POINTER(ctypes.c_uint32)()
        ln = lib().or_api_sc_a2b(self.h, ptr(a), ptr(b), n, ptr(np.ascontiguousarray(z)),
                                 ptr(ch), ptr(fe), ptr(fs), ctypes.byref(p))
        proof = np.ctypeslib.as_array(p, shape=(ln,)).copy()
        lib().or_api_free(p)
        return {"challenges": ch.reshape(n, 8), "final_evals": fe.reshape(2, 8),
                "final_sum": fs, "proof": proof}


def whir_config(v, sec=128, max_pow=16, f0=7, f=4, rate=1, rs0=5, sa=CAPACITY_BOUND):
    out = (ctypes.c_int * 256)()
    k = lib().or_api_whir_config(v, sec, max_pow, f0, f, rate, rs0, sa, out)
    if k < 0:
        raise ValueError("no WHIR configuration for these parameters")
    vals = list(out[:k])
    head = dict(zip(["n_rounds", "final_sumcheck_rounds", "commit_ood", "starting_folding_pow",
                     "final_queries", "final_pow", "final_folding_pow", "final_log_domain"],
                    vals[:8]))
    rounds = []
    for r in range(head["n_rounds"]):
        rounds.append(dict(zip(["pow_bits", "folding_pow_bits", "num_queries", "ood_samples",
                                "log_inv_rate", "num_variables", "folding_factor",
                                "log_domain_size"], vals[8 + 8 * r: 16 + 8 * r])))
    head["rounds"] = rounds
    return head


def rng_f(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().or_api_rng_f(seed, ptr(out), n)
    return out


def rng_u32(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().or_api_rng_u32(seed, ptr(out), n)
    return out


def fold_f_large(m, scalars, k):
    log_n = m.size.bit_length() - 1
    out = np.zeros((m.size >> k) * 8, dtype=np.uint32)
    lib().or_api_fold_f_large(ptr(m), log_n, ptr(np.ascontiguousarray(scalars)), k, ptr(out))
    return out


def fold_ef_large(m, scalars, k):
    log_n = (m.size // 8).bit_length() - 1
    out = np.zeros(((m.size // 8) >> k) * 8, dtype=np.uint32)
    lib().or_api_fold_ef_large(ptr(m), log_n, ptr(np.ascontiguousarray(scalars)), k, ptr(out))
    return out


def fold_small(m, scalars, k, is_ef=False):
    """fold_multilinear_in_small_field: F scalars, F or EF table."""
    w = 8 if is_ef else 1
    log_n = (m.size // w).bit_length() - 1
    out = np.zeros(((m.size // w) >> k) * w, dtype=np.uint32)
    lib().or_api_fold_small(ptr(m), w, log_n, ptr(np.ascontiguousarray(scalars)), k, ptr(out))
    return out


def lincomb(cols, scalars):
    """multilinears_linear_combination: cols (n_cols x 2^n F), scalars (n_cols x 8)."""
    n_cols, N = cols.shape
    out = np.zeros(N * 8, dtype=np.uint32)
    lib().or_api_lincomb(ptr(np.ascontiguousarray(cols).reshape(-1)), n_cols, N.bit_length() - 1,
                         ptr(np.ascontiguousarray(scalars).reshape(-1)), ptr(out))
    return out


def mle_eval(evals, point, is_ef=False):
    """EvaluationsList::evaluate at an EF point (big-endian variables)."""
    n = point.size // 8
    out = np.zeros(8, dtype=np.uint32)
    fn = lib().or_api_mle_eval_ef if is_ef else lib().or_api_mle_eval_f
    fn(ptr(np.ascontiguousarray(evals)), n, ptr(np.ascontiguousarray(point).reshape(-1)), ptr(out))
    return out


def domain_separator(log_n, **kw):
    s = dict(DEFAULT, **kw)
    out = np.zeros(4096, dtype=np.uint32)
    k = lib().or_api_domsep(log_n, s["sec"], s["f0"], s["f"], s["rate"], s["skips"], s["rs0"], s["sa"], ptr(out),
                            out.size)
    return out[:k].copy()


class Transcript:
    """The oracle's ProverState over DuplexChallenger<F, Poseidon2-16, 16, 8>."""

    def __init__(self, p16, domsep=()):
        d = np.ascontiguousarray(domsep, dtype=np.uint32)
        self.h = lib().or_api_ts_new(ptr(np.ascontiguousarray(p16, dtype=np.uint32)), ptr(d), d.size)

    def close(self):
        if self.h:
            lib().or_api_ts_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _call(self, fn, v):
        a = np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)
        getattr(lib(), "or_api_ts_" + fn)(self.h, ptr(a), a.size)

    def observe(self, v):
        self._call("observe", v)

    def add_base(self, v):
        self._call("add_base", v)

    def hint(self, v):
        self._call("hint", v)

    def sample_ext(self, n=1):
        out = np.zeros(8 * n, dtype=np.uint32)
        lib().or_api_ts_sample_ef(self.h, ptr(out), n)
        return out.reshape(n, 8)

    def sample_bits(self, bits):
        return int(lib().or_api_ts_sample_bits(self.h, bits))

    def grind(self, bits):
        lib().or_api_ts_pow(self.h, bits)

    def proof_data(self):
        n = lib().or_api_ts_len(self.h)
        if n == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(lib().or_api_ts_data(self.h), shape=(n,)).copy()


def sc_prove(ts, skips, mles, code, n_instr, degree, batching, eq_factor, is_zerofier, sum, n_rounds=0,
             grinding=(1, 128), missing=None, is_ef=False):
    """sumcheck::prove with a constraint program on the oracle; mles (n_mles, 2^n[, 8])."""
    m = np.ascontiguousarray(mles, dtype=np.uint32)
    n_mles, N = m.shape[0], m.shape[1]
    n = N.bit_length() - 1
    nr = n_rounds or (n - skips + 1)
    ch = np.zeros(8 * nr, dtype=np.uint32)
    fl = 1 << (n - skips - nr + 1)
    folded = np.zeros(n_mles * fl * 8, dtype=np.uint32)
    fs = np.zeros(8, dtype=np.uint32)
    conv = lambda v: None if v is None else np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)
    lib().or_api_sc_prove(ts.h, skips, ptr(m.reshape(-1)), n_mles, 0 if is_ef else 1, n,
                          ptr(np.ascontiguousarray(code, dtype=np.uint32)), n_instr, degree, ptr(conv(batching)),
                          ptr(conv(eq_factor)), int(is_zerofier), ptr(conv(sum)), n_rounds, grinding[0], grinding[1],
                          ptr(conv(missing)), ptr(ch), ptr(folded), ptr(fs))
    return ch.reshape(nr, 8), folded.reshape(n_mles, fl, 8), fs


def prog_check(code, n_instr, n_inputs):
    return lib().or_api_prog_check(ptr(np.ascontiguousarray(code, dtype=np.uint32)), n_instr, n_inputs)


def prog_eval(code, n_instr, n_inputs, point, batching, is_ef=False):
    out = np.zeros(8, dtype=np.uint32)
    lib().or_api_prog_eval(ptr(np.ascontiguousarray(code, dtype=np.uint32)), n_instr, n_inputs, 0 if is_ef else 1,
                           ptr(np.ascontiguousarray(point, dtype=np.uint32).reshape(-1)),
                           ptr(np.ascontiguousarray(batching, dtype=np.uint32).reshape(-1)), ptr(out))
    return out


class GenericAir:
    """AirTable for an AIR given as a constraint program (oracle): width C, the first
    n_pre columns preprocessed (pre: n_pre x 2^log_n), code over [local, next]."""

    def __init__(self, C, n_pre, log_n, pre, code, n_instr, degree):
        self.C, self.n_pre, self.log_n, self.n_instr, self.degree = C, n_pre, log_n, n_instr, degree
        self.pre = np.ascontiguousarray(pre if pre is not None else np.zeros(1, np.uint32), dtype=np.uint32)
        self.code = np.ascontiguousarray(code, dtype=np.uint32)

    def _args(self, kw):
        s = dict(DEFAULT, **kw)
        return (self.C, self.n_pre, self.log_n, ptr(self.pre.reshape(-1)), ptr(self.code), self.n_instr, self.degree,
                s["sec"], s["f0"], s["f"], s["rate"], s["skips"], s["rs0"], s["sa"])

    def domain_separator(self, **kw):
        s = dict(DEFAULT, **kw)
        out = np.zeros(4096, dtype=np.uint32)
        k = lib().or_api_gair_domsep(self.C - self.n_pre, self.log_n, s["sec"], s["f0"], s["f"], s["rate"],
                                     s["skips"], s["rs0"], s["sa"], ptr(out), out.size)
        return out[:k].copy()

    def prove_into(self, ts, witness, p16, p24, **kw):
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
        lib().or_api_gair_prove(*self._args(kw), ptr(w), ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                                ptr(np.ascontiguousarray(p24, dtype=np.uint32)), ts.h)

    def prove(self, witness, p16, p24, **kw):
        ts = Transcript(p16, self.domain_separator(**kw))
        self.prove_into(ts, witness, p16, p24, **kw)
        return ts.proof_data()

    def verify(self, proof, p16, p24, domsep=None, **kw):
        d = self.domain_separator(**kw) if domsep is None else np.ascontiguousarray(domsep, dtype=np.uint32)
        pr = np.ascontiguousarray(proof, dtype=np.uint32)
        return lib().or_api_gair_verify(*self._args(kw), ptr(np.ascontiguousarray(p16, dtype=np.uint32)),
                                        ptr(np.ascontiguousarray(p24, dtype=np.uint32)), ptr(d), d.size, ptr(pr),
                                        pr.size)
