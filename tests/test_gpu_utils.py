"""crates/utils counterparts on the GPU against the oracle: small-field folds,
the linear combination of columns and multilinear evaluation (bit-exact)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 0x7F000001


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("log_n,k,is_ef", [(10, 1, False), (12, 4, False), (17, 3, True), (9, 9, True),
                                           (20, 2, False)])
def test_fold_small(gpu_ctx, oracle, log_n, k, is_ef):
    from whirlaway_amd import utils
    _say(f"fold_small log_n={log_n} k={k} ef={is_ef}")
    rng = np.random.default_rng(log_n + 100 * k)
    w = 8 if is_ef else 1
    m = rng.integers(0, P, (1 << log_n) * w, dtype=np.uint32)
    m[:8] = P - 1  # field extremes
    sc = rng.integers(0, P, 1 << k, dtype=np.uint32)
    sc[0] = P - 1
    assert np.array_equal(utils.fold_multilinear_in_small_field(gpu_ctx, m, sc, is_ef=is_ef),
                          oracle.fold_small(m, sc, k, is_ef=is_ef))


@pytest.mark.parametrize("n_cols,log_n", [(1, 4), (5, 11), (165, 12), (20, 17)])
def test_linear_combination(gpu_ctx, oracle, n_cols, log_n):
    from whirlaway_amd import utils
    _say(f"lincomb {n_cols} x 2^{log_n}")
    rng = np.random.default_rng(n_cols * 7 + log_n)
    cols = rng.integers(0, P, (n_cols, 1 << log_n), dtype=np.uint32)
    sc = rng.integers(0, P, (n_cols, 8), dtype=np.uint32)
    assert np.array_equal(utils.multilinears_linear_combination(gpu_ctx, cols, sc), oracle.lincomb(cols, sc))


@pytest.mark.parametrize("n,is_ef", [(0, False), (1, True), (5, False), (8, False), (9, True), (15, False),
                                     (16, True), (22, False)])
def test_mle_evaluate(gpu_ctx, oracle, n, is_ef):
    from whirlaway_amd import utils
    _say(f"mle_evaluate n={n} ef={is_ef}")
    rng = np.random.default_rng(n * 3 + int(is_ef))
    w = 8 if is_ef else 1
    ev = rng.integers(0, P, (1 << n) * w, dtype=np.uint32)
    pt = rng.integers(0, P, (n, 8), dtype=np.uint32)
    assert np.array_equal(utils.mle_evaluate(gpu_ctx, ev, pt, is_ef=is_ef), oracle.mle_eval(ev, pt, is_ef=is_ef))
