"""Pins the oracle's arithmetic primitives against mathematics and published vectors.

- log2_up: the reference's own unit tests (crates/utils/src/misc.rs:10-44).
- EF = F[X]/(X^8 - 3): big-int schoolbook product, inverse, irreducibility (sympy).
- two-adic generators: order exactly 2^k.
- ChaCha block function: RFC 7539 section 2.3.2 and appendix A.1 vectors.
"""
import numpy as np
import pytest
import sympy

import pyfield as pf

P = pf.P


def test_log2_up_reference_vectors(oracle):
    L = oracle.lib()
    # crates/utils/src/misc.rs:10-44
    assert L.or_log2_up(0) == 0
    for x, e in [(1, 0), (2, 1), (4, 2), (8, 3), (16, 4), (1024, 10)]:
        assert L.or_log2_up(x) == e
    for x, e in [(3, 2), (5, 3), (9, 4), (17, 5), (1025, 11)]:
        assert L.or_log2_up(x) == e
    assert L.or_log2_up(2 ** 64 - 1) == 64


def test_modulus_and_irreducibility():
    assert P == 2 ** 31 - 2 ** 24 + 1 and sympy.isprime(P)
    x = sympy.symbols("x")
    assert sympy.Poly(x ** 8 - 3, x, modulus=P).is_irreducible


def test_ef_mul_and_inverse(oracle):
    rng = np.random.default_rng(1)
    L = oracle.lib()
    for _ in range(50):
        a = rng.integers(0, P, 8, dtype=np.uint32)
        b = rng.integers(0, P, 8, dtype=np.uint32)
        out = np.zeros(8, dtype=np.uint32)
        L.or_api_ef_mul(oracle.ptr(a), oracle.ptr(b), oracle.ptr(out))
        assert [int(v) for v in out] == pf.emul(pf.ef(a), pf.ef(b))
        inv = np.zeros(8, dtype=np.uint32)
        L.or_api_ef_inv(oracle.ptr(a), oracle.ptr(inv))
        assert pf.emul(pf.ef(a), pf.ef(inv)) == pf.ef(1)
    # a planted mistake (W = 2 instead of 3) must be visible in the product
    a, b = pf.ef([0] * 7 + [1]), pf.ef([0, 1] + [0] * 6)
    assert pf.emul(a, b) == pf.ef(3)


def test_two_adic_generators(oracle):
    L = oracle.lib()
    for bits in range(0, 25):
        g = L.or_api_two_adic_generator(bits)
        assert pow(g, 1 << bits, P) == 1
        if bits:
            assert pow(g, 1 << (bits - 1), P) != 1
        assert g == pow(3, (P - 1) >> bits, P)


def _words(hexstr):
    b = bytes.fromhex(hexstr.replace(" ", "").replace(":", ""))
    return [int.from_bytes(b[i:i + 4], "little") for i in range(0, len(b), 4)]


def test_chacha20_rfc7539_vectors(oracle):
    L = oracle.lib()
    # RFC 7539 2.3.2: key 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00
    key = np.array(_words("".join(f"{i:02x}" for i in range(32))), dtype=np.uint32)
    nonce = np.array(_words("000000090000004a00000000"), dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    L.or_chacha_block_ietf(oracle.ptr(key), 1, oracle.ptr(nonce), 20, oracle.ptr(out))
    expect = _words("10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
                    "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert [int(v) for v in out] == expect
    # RFC 7539 A.1 #1: all-zero key, counter 0, nonce 0 (same layout as the 64-bit
    # counter/stream variant used by StdRng)
    zk = np.zeros(8, dtype=np.uint32)
    L.or_chacha_block(oracle.ptr(zk), 0, 0, 20, oracle.ptr(out))
    expect = _words("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                    "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert [int(v) for v in out] == expect


def test_rng_sampling_is_canonical(oracle):
    vals = oracle.rng_f(0, 4096)
    assert (vals < P).all()
    # the reference draws F as next_u32() >> 1 read as Montgomery form; the sampled
    # canonical value is (u >> 1) * 2^-32 mod p for accepted u
    raw = oracle.rng_u32(0, 64)
    rinv = pow(2 ** 32, P - 2, P)
    accepted = [(int(u) >> 1) * rinv % P for u in raw if (int(u) >> 1) < P]
    assert [int(v) for v in vals[: len(accepted)]] == accepted


def test_f_arithmetic_against_bigint(oracle):
    rng = np.random.default_rng(2)
    L = oracle.lib()
    a = rng.integers(0, P, 8, dtype=np.uint32)
    b = np.zeros(8, dtype=np.uint32)
    b[0] = 5
    out = np.zeros(8, dtype=np.uint32)
    L.or_api_ef_mul(oracle.ptr(a), oracle.ptr(b), oracle.ptr(out))
    assert [int(v) for v in out] == [int(v) * 5 % P for v in a]
