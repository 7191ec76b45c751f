"""Poseidon2 as the device kernels compute it -- the partial rounds as a
convolution with the internal diagonal's entries taken in pairs (d, -d) -- must
equal the round-by-round permutation (kb.h p2_permute, which the proof parity
tests pin against the oracle), and so must the host transcript's AVX2
Poseidon2-16 (host_p16.h, whose partial rounds cube by one reduction mod 2^64).
The check runs on the host through the library, for widths 16 and 24, on random
states and round constants and on field-extreme states (and, for the AVX2 cube,
on field-extreme inputs).  Reference:
Poseidon2KoalaBear<16>/<24> as named at src/examples/poseidon2.rs:22-23."""
import ctypes

import pytest


@pytest.mark.parametrize("seed", [1, 2])
def test_partial_rounds_by_convolution_equal_round_by_round(seed):
    from whirlaway_amd import lib
    mm = (ctypes.c_uint64 * 2)()
    assert lib().wl_p2_conv_check(50000, seed, mm) == 0
    assert list(mm) == [0, 0]
