"""The oracle's generic AIR (constraints as a program, preprocessed columns) on
the CPU: with the Poseidon2 AIR's program and no preprocessed column its proof
equals the hand-specialised Poseidon2 prover's byte for byte (two independent
code paths); with a preprocessed column, and for the MulFib AIR, proofs verify,
changed words are rejected, and a witness that breaks a constraint is rejected."""
import numpy as np
import pytest

from airs import P, mul_fib_columns, mul_fib_program


def test_poseidon2_program_proof_equals_specialised(oracle):
    from whirlaway_amd.program import poseidon2_air_program
    n = 8
    b = oracle.Bench(n, seed=2)
    prog = poseidon2_air_program(b.air_consts())
    g = oracle.GenericAir(165, 0, n, None, prog.code(), prog.n_instr, 3)
    got = g.prove(b.trace(), b.p2_consts(16), b.p2_consts(24))
    assert np.array_equal(got, b.prove())
    b.close()


def test_poseidon2_program_with_preprocessed_export_column(oracle):
    """prove_poseidon2(.., n_preprocessed_columns = 1, ..) (src/examples/poseidon2.rs:87,134):
    the export column becomes preprocessed; 164 committed columns."""
    from whirlaway_amd.program import poseidon2_air_program
    n = 7
    b = oracle.Bench(n, seed=4)
    prog = poseidon2_air_program(b.air_consts())
    tr = b.trace().reshape(165, 1 << n)
    g = oracle.GenericAir(165, 1, n, tr[:1], prog.code(), prog.n_instr, 3)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    proof = g.prove(tr[1:], p16, p24)
    assert g.verify(proof, p16, p24) == 0
    bad = proof.copy()
    bad[len(bad) // 2] = (int(bad[len(bad) // 2]) + 1) % P
    assert g.verify(bad, p16, p24) != 0
    b.close()


@pytest.mark.parametrize("n,kw", [(9, {}), (10, dict(skips=2)), (8, dict(skips=1, sa=0)),
                                  (11, dict(sec=100, f0=5, f=3, rate=2, skips=3, rs0=3))])
def test_mul_fib_proves_and_verifies(oracle, n, kw):
    b = oracle.Bench(5, seed=1)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    prog = mul_fib_program()
    pre, wit = mul_fib_columns(n, n)
    g = oracle.GenericAir(4, 1, n, pre, prog.code(), prog.n_instr, 3)
    proof = g.prove(wit, p16, p24, **kw)
    assert g.verify(proof, p16, p24, **kw) == 0
    rng = np.random.default_rng(n)
    for pos in rng.integers(0, proof.size, 6):
        bad = proof.copy()
        bad[pos] = (int(bad[pos]) + 1) % P
        assert g.verify(bad, p16, p24, **kw) != 0, pos
    broken = wit.copy()
    broken[2, 5] = (int(broken[2, 5]) + 1) % P  # c != a b k on row 5
    assert g.verify(g.prove(broken, p16, p24, **kw), p16, p24, **kw) != 0
    # the preprocessed column is the verifier's own: another one is rejected
    pre2 = pre.copy()
    pre2[0, 3] ^= 1
    g2 = oracle.GenericAir(4, 1, n, pre2, prog.code(), prog.n_instr, 3)
    assert g2.verify(proof, p16, p24, **kw) != 0
    b.close()
