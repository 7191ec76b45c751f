"""Autotuner: every launch variant of the commitment kernels (NTT split, leaf
hashing, tree schedule) gives the same codeword and Merkle root as the defaults,
and a tuned commit records its choices and commits identically."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 0x7F000001


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("log_rows", [11, 14, 15, 16, 18])
def test_autotune_variants_commit_identically(gpu_ctx, oracle, log_rows):
    from whirlaway_amd import autotune, utils
    n_cols, log_m = 32, log_rows - 1
    b = oracle.Bench(5, seed=21)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    cols = np.random.default_rng(log_rows).integers(0, P, (n_cols, 1 << log_m), dtype=np.uint32)
    prev = autotune.mode(0)
    try:
        autotune.clear()
        root0, cw0 = utils.rs_commit(gpu_ctx, p16, p24, cols, log_rows, True)
        # the default codeword's first column against the oracle's encoding
        L = oracle.lib()
        col = np.zeros(1 << log_rows, dtype=np.uint32)
        col[: 1 << log_m] = cols[0]
        L.or_mobius(oracle.ptr(col), log_m)
        L.or_ntt(oracle.ptr(col), log_rows)
        assert np.array_equal(cw0.reshape(n_cols, -1)[0], col)
        for kind, nvar in ((autotune.NTT, 3), (autotune.LEAVES, 4), (autotune.TREE, 6)):
            for v in range(nvar):
                _say(f"2^{log_rows} rows: {autotune.KINDS[kind]} variant {v}")
                autotune.clear()
                autotune.force(kind, log_rows, v, width=n_cols)
                root, cw = utils.rs_commit(gpu_ctx, p16, p24, cols, log_rows, True)
                assert np.array_equal(root, root0), (kind, v)
                assert np.array_equal(cw, cw0), (kind, v)
        _say(f"2^{log_rows} rows: tuned in place")
        autotune.clear()
        autotune.mode(1)
        calls = 0
        while True:  # every call runs one candidate until all have their samples
            root, cw = utils.rs_commit(gpu_ctx, p16, p24, cols, log_rows, calls % 4 == 0)
            assert np.array_equal(root, root0)
            if calls % 4 == 0:
                assert np.array_equal(cw, cw0)
            calls += 1
            ent = [e for e in autotune.entries() if e["log_size"] == log_rows]
            if ent and all(e["variant"] >= 0 for e in ent):
                break
            assert calls < 16, ent
        kinds = {e["kind"] for e in ent}
        assert {"merkle_leaves", "merkle_tree"} <= kinds
        if log_rows >= 14:
            assert "rs_encode_split" in kinds
        _say(f"  decided after {calls} calls: {ent}")
        # the table serves the next call
        root, _ = utils.rs_commit(gpu_ctx, p16, p24, cols, log_rows, False)
        assert np.array_equal(root, root0)
    finally:
        autotune.clear()
        autotune.mode(prev)
