"""wl_sumcheck_prove (sumcheck::prove for a constraint program) on the GPU against
the oracle's sumcheck on the same inputs and transcript: proof words, challenges,
folded multilinears and final sum bit-exact, over the interface's options (skips,
eq factor, zerofier, partial rounds, missing_mul_factor, grinding modes, F and EF
multilinears)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 0x7F000001


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_program(n_inputs, seed):
    from whirlaway_amd.program import ConstraintProgram
    rng = np.random.default_rng(seed)
    p = ConstraintProgram(n_inputs=n_inputs)
    xs = [p.input(i) for i in range(n_inputs)]
    for _ in range(4):
        i, j = rng.integers(0, n_inputs, 2)
        p.assert_zero(xs[i] * xs[j] - xs[(i + 1) % n_inputs] + int(rng.integers(0, P)))
    p.assert_zero(-(xs[0] + 3) * xs[-1])
    return p


def _run(gpu_ctx, oracle, p16, dom, skips, mles, prog, degree, batching, eq, zerofier, s, n_rounds=0,
         grinding=(1, 128), missing=None, is_ef=False):
    from whirlaway_amd import sumcheck
    from whirlaway_amd.transcript import ProverState
    ours = ProverState(p16, dom)
    ch, fo, fs = sumcheck.prove(gpu_ctx, skips, mles, prog, degree, batching, eq, zerofier, ours, s,
                                n_rounds=n_rounds or None, grinding=grinding, missing_mul_factor=missing,
                                is_ef=is_ef)
    ref = oracle.Transcript(p16, dom)
    rch, rfo, rfs = oracle.sc_prove(ref, skips, mles, prog.code(), prog.n_instr, degree, batching, eq, zerofier, s,
                                    n_rounds=n_rounds, grinding=grinding, missing=missing, is_ef=is_ef)
    assert np.array_equal(ours.proof_data(), ref.proof_data())
    assert np.array_equal(ch, rch)
    assert np.array_equal(fo, rfo)
    assert np.array_equal(fs, rfs)
    assert np.array_equal(ours.sample(1), ref.sample_ext(1))


def _consts(oracle):
    b = oracle.Bench(5, seed=17)
    return b, b.p2_consts(16)


def test_a2b_with_eq_factor(gpu_ctx, oracle):
    from whirlaway_amd.program import a2b_program
    _say("generic sumcheck a^2 b, eq factor")
    b, p16 = _consts(oracle)
    rng = np.random.default_rng(1)
    n = 9
    mles = rng.integers(0, P, (2, 1 << n), dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    one = np.zeros(8, dtype=np.uint32)
    one[0] = 1
    s = rng.integers(0, P, 8, dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [4, 5], 1, mles, a2b_program(), 3, one, z, False, s)


def test_product_ef_inputs(gpu_ctx, oracle):
    from whirlaway_amd.program import product_program
    _say("generic sumcheck product of EF tables")
    b, p16 = _consts(oracle)
    rng = np.random.default_rng(2)
    n = 8
    mles = rng.integers(0, P, (2, 1 << n, 8), dtype=np.uint32)
    one = np.zeros(8, dtype=np.uint32)
    one[0] = 1
    s = rng.integers(0, P, 8, dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [], 1, mles, product_program(), 2, one, None, False, s, is_ef=True)


@pytest.mark.parametrize("n,skips", [(6, 4), (8, 2)])
def test_poseidon2_zerocheck_program(gpu_ctx, oracle, n, skips):
    """The zerocheck of the Poseidon2 AIR (crates/air/src/prove.rs:114-131) through the
    program path: 330 up/down columns of a trace, skips, eq factor, zerofier, sum 0."""
    from whirlaway_amd.program import poseidon2_air_program
    _say(f"generic sumcheck Poseidon2 zerocheck n={n} skips={skips}")
    b = oracle.Bench(n, seed=5)
    p16 = b.p2_consts(16)
    prog = poseidon2_air_program(b.air_consts())
    tr = b.trace().reshape(165, 1 << n)
    up = tr.copy()
    up[:, -1] = tr[:, -2]
    dn = np.concatenate([tr[:, 1:], tr[:, -1:]], axis=1)
    mles = np.concatenate([up, dn])
    rng = np.random.default_rng(n)
    alpha = rng.integers(0, P, (149, 8), dtype=np.uint32)
    zc = rng.integers(0, P, (n - skips + 1, 8), dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [1], skips, mles, prog, 3, alpha, zc, True, np.zeros(8, np.uint32))
    b.close()


@pytest.mark.parametrize("skips,eq,rounds,grind,missing", [(2, True, 2, (2, 4), True), (3, False, 0, (0, 0), False),
                                                           (1, True, 3, (1, 100), True), (2, False, 1, (2, 0), False)])
def test_program_options(gpu_ctx, oracle, skips, eq, rounds, grind, missing):
    _say(f"generic sumcheck options skips={skips} eq={eq} rounds={rounds} grind={grind} missing={missing}")
    b, p16 = _consts(oracle)
    rng = np.random.default_rng(skips * 10 + rounds)
    n = 7
    prog = _random_program(3, skips)
    mles = rng.integers(0, P, (3, 1 << n), dtype=np.uint32)
    bat = rng.integers(0, P, (prog.n_constraints, 8), dtype=np.uint32)
    z = rng.integers(0, P, (n - skips + 1, 8), dtype=np.uint32) if eq else None
    mm = rng.integers(0, P, 8, dtype=np.uint32) if missing else None
    s = rng.integers(0, P, 8, dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [9], skips, mles, prog, 2, bat, z, False, s, n_rounds=rounds, grinding=grind,
         missing=mm)


def test_program_partitions(gpu_ctx, oracle):
    """A program of 12 constraints over 2^17 points: the first round evaluates it
    whole (2^17 threads), later rounds in 4 constraint groups (16 groups need 32
    constraints; the Poseidon2 program above covers them), the smallest ones with
    the register file in LDS; every round bit-exact."""
    from whirlaway_amd.program import ConstraintProgram
    _say("generic sumcheck, 12 constraints, n = 17 (1 and 4 constraint groups)")
    b, p16 = _consts(oracle)
    rng = np.random.default_rng(77)
    n = 17
    p = ConstraintProgram(n_inputs=3)
    xs = [p.input(i) for i in range(3)]
    for k in range(12):
        i, j = rng.integers(0, 3, 2)
        p.assert_zero(xs[i] * xs[j] * int(rng.integers(1, P)) - xs[(i + k) % 3] + int(rng.integers(0, P)))
    mles = rng.integers(0, P, (3, 1 << n), dtype=np.uint32)
    bat = rng.integers(0, P, (p.n_constraints, 8), dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    s = rng.integers(0, P, 8, dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [3], 1, mles, p, 2, bat, z, False, s)


def test_program_wide_sums(gpu_ctx, oracle):
    """Constraints over a 300-term sum written as a chain of additions (which the
    compiler rebalances into a tree for the bundled extension-field rounds) and a
    chain with subtractions: 41 constraints, so 1, 4 and 16 constraint groups;
    every round bit-exact against the oracle, which evaluates the program as
    written."""
    from whirlaway_amd.program import ConstraintProgram
    _say("generic sumcheck, 300-term sums, n = 9")
    b, p16 = _consts(oracle)
    rng = np.random.default_rng(303)
    n, m = 9, 300
    p = ConstraintProgram(n_inputs=m)
    xs = [p.input(i) for i in range(m)]
    acc = xs[0]
    for i in range(1, m):
        acc = acc + xs[i]
    alt = xs[1]
    for i in range(2, 60):
        alt = alt - xs[i] if i % 3 else alt + xs[i]
    for k in range(40):
        p.assert_zero(acc * xs[k] - xs[k + 1] + int(rng.integers(0, P)))
    p.assert_zero(alt * acc)
    mles = rng.integers(0, P, (m, 1 << n), dtype=np.uint32)
    bat = rng.integers(0, P, (p.n_constraints, 8), dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    s = rng.integers(0, P, 8, dtype=np.uint32)
    _run(gpu_ctx, oracle, p16, [5], 1, mles, p, 2, bat, z, False, s)
