"""Synthetic benchmark inputs: the product's host StdRng (ChaCha12) generator
against the oracle's (src/examples/poseidon2.rs:104-151 draw order).  Host-only
entry point, so this runs without a GPU."""
import numpy as np
import pytest


@pytest.mark.parametrize("n,seed", [(5, 0), (9, 7), (12, 0)])
def test_bench_inputs_match_oracle(oracle, n, seed):
    from whirlaway_amd.air import bench_inputs
    rc, inp, p16, p24 = bench_inputs(seed, n)
    b = oracle.Bench(n, seed)
    assert np.array_equal(rc, b.air_consts())
    assert np.array_equal(inp, b.inputs())
    assert np.array_equal(p16, b.p2_consts(16))
    assert np.array_equal(p24, b.p2_consts(24))
