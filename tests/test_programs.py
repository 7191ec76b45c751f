"""Constraint programs (the build's form of an AIR's symbolic constraints), on
the CPU: the Poseidon2 AIR written as a program evaluates exactly as the oracle's
hand-written AIR evaluator (both fields, random points), malformed programs are
refused, and the oracle's sumcheck driven by a program equals its built-in
product and a^2 b computations (different code paths, same transcript)."""
import numpy as np
import pytest

P = 0x7F000001


def _bench(oracle, n=5, seed=3):
    return oracle.Bench(n, seed=seed)


def test_poseidon2_program_matches_air_evaluator(oracle):
    from whirlaway_amd.program import poseidon2_air_program
    b = _bench(oracle)
    prog = poseidon2_air_program(b.air_consts())
    assert prog.n_constraints == 149 and prog.width == 165
    code = prog.code()
    assert oracle.prog_check(code, prog.n_instr, prog.n_inputs) == 0
    rng = np.random.default_rng(0)
    L = oracle.lib()
    for trial in range(6):
        alpha = rng.integers(0, P, (149, 8), dtype=np.uint32)
        pt = rng.integers(0, P, 330, dtype=np.uint32)
        if trial == 0:  # a trace row: every constraint vanishes
            tr = b.trace().reshape(165, -1)
            pt = np.concatenate([tr[:, 3], tr[:, 4]]).astype(np.uint32)
        want = np.zeros(8, dtype=np.uint32)
        L.or_api_air_eval_f(b.h, oracle.ptr(pt[:165].copy()), oracle.ptr(alpha.reshape(-1)), oracle.ptr(want))
        got = oracle.prog_eval(code, prog.n_instr, prog.n_inputs, pt, alpha)
        assert np.array_equal(got, want), trial
        if trial == 0:
            assert not got.any()
        pte = rng.integers(0, P, (330, 8), dtype=np.uint32)
        want = np.zeros(8, dtype=np.uint32)
        L.or_api_air_eval_ef(b.h, oracle.ptr(pte[:165].reshape(-1).copy()), oracle.ptr(alpha.reshape(-1)),
                             oracle.ptr(want))
        assert np.array_equal(oracle.prog_eval(code, prog.n_instr, prog.n_inputs, pte, alpha, is_ef=True), want)
    b.close()


def test_malformed_programs_are_refused(oracle):
    from whirlaway_amd.program import ConstraintProgram
    ok = ConstraintProgram(n_inputs=2)
    ok.assert_zero(ok.input(0) * ok.input(1) - 5)
    code = ok.code()
    assert oracle.prog_check(code, ok.n_instr, 2) == 0
    fwd = code.copy().reshape(-1, 4)
    fwd[0, 1] = 2 + ok.n_instr  # refers to a later value
    assert oracle.prog_check(fwd.reshape(-1), ok.n_instr, 2) != 0
    p = ConstraintProgram(n_inputs=1)
    p.assert_zero(p.input(0))
    c = list(p.code()) + [1, 1, 0, 0]  # ADD whose operand is the ASSERT
    assert oracle.prog_check(np.array(c, dtype=np.uint32), 2, 1) != 0
    bad = np.array([0, P, 0, 0], dtype=np.uint32)  # non-canonical constant
    assert oracle.prog_check(bad, 1, 1) != 0


def test_program_sumcheck_equals_builtin_computations(oracle):
    """sumcheck::prove with the product / a^2 b programs equals the oracle's built-in
    computations: the inner-sumcheck shape (product, EF tables, no eq factor) inside a
    whole proof is covered by proofs; here the a^2 b micro shape with eq factor z."""
    from whirlaway_amd.program import a2b_program
    b = _bench(oracle, 5, 11)
    rng = np.random.default_rng(9)
    n = 7
    a = rng.integers(0, P, 1 << n, dtype=np.uint32)
    bb = rng.integers(0, P, 1 << n, dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    want = b.sc_a2b(a, bb, z.reshape(-1).copy())
    g = (a.astype(np.uint64) * a % P * bb % P).astype(np.uint32)
    s = oracle.mle_eval(g, z)
    prog = a2b_program()
    ts = oracle.Transcript(b.p2_consts(16))
    one = np.zeros(8, dtype=np.uint32)
    one[0] = 1
    ch, folded, fs = oracle.sc_prove(ts, 1, np.stack([a, bb]), prog.code(), prog.n_instr, 3, one, z, False, s)
    assert np.array_equal(ts.proof_data(), want["proof"])
    assert np.array_equal(ch, want["challenges"])
    assert np.array_equal(folded[:, 0, :], want["final_evals"])
    assert np.array_equal(fs, want["final_sum"])
    b.close()


def test_program_builder_operators():
    from whirlaway_amd.program import ConstraintProgram
    p = ConstraintProgram(width=3)
    x, y = p.local(0), p.next(2)
    e = 2 * x - (y ** 3) + 7 - (-x) * 1
    p.assert_zero(e)
    code = p.code().reshape(-1, 4)
    assert p.n_inputs == 6 and p.n_constraints == 1
    assert code[-1, 0] == 5 and code[-1, 1] == e.id
    c2 = p.const(2)
    assert c2.id == p.const(2 + P).id  # constants are shared
    with pytest.raises(IndexError):
        p.local(6)
