"""Generic AIR tables on the GPU (constraints as a program, preprocessed columns)
against the oracle: proofs byte-identical, accepted by the oracle's verifier."""
import sys

import numpy as np
import pytest

from airs import P, mul_fib_columns, mul_fib_program

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _settings(kw):
    import whirlaway_amd as wa
    return wa.AirSettings(security_bits=kw.get("sec", 128), folding_factor_first=kw.get("f0", 7),
                          folding_factor=kw.get("f", 4), log_inv_rate=kw.get("rate", 1),
                          univariate_skips=kw.get("skips", 4), rs_domain_initial_reduction=kw.get("rs0", 5),
                          soundness=kw.get("sa", 2))


@pytest.mark.parametrize("n", [8, 11])
def test_poseidon2_program_table_equals_specialised(gpu_ctx, oracle, n):
    """The Poseidon2 AIR through the program path proves byte-identically to the
    oracle's specialised Poseidon2 prover (and so to the GPU's specialised path)."""
    import whirlaway_amd as wa
    from whirlaway_amd.program import poseidon2_air_program
    _say(f"program table Poseidon2 n={n}")
    b = oracle.Bench(n, seed=6)
    prog = poseidon2_air_program(b.air_consts())
    t = wa.air.ProgramAirTable(gpu_ctx, prog, n, 3)
    got = t.prove(wa.AirSettings(), b.p2_consts(16), b.p2_consts(24), b.trace())
    t.close()
    assert np.array_equal(got, b.prove())
    b.close()


def test_poseidon2_program_table_preprocessed_export(gpu_ctx, oracle):
    import whirlaway_amd as wa
    from whirlaway_amd.program import poseidon2_air_program
    n = 9
    _say("program table Poseidon2 with the export column preprocessed")
    b = oracle.Bench(n, seed=8)
    prog = poseidon2_air_program(b.air_consts())
    tr = b.trace().reshape(165, 1 << n)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    t = wa.air.ProgramAirTable(gpu_ctx, prog, n, 3, preprocessed=tr[:1])
    got = t.prove(wa.AirSettings(), p16, p24, tr[1:])
    t.close()
    g = oracle.GenericAir(165, 1, n, tr[:1], prog.code(), prog.n_instr, 3)
    assert np.array_equal(got, g.prove(tr[1:], p16, p24))
    assert g.verify(got, p16, p24) == 0
    b.close()


@pytest.mark.parametrize("n,kw", [(9, {}), (12, dict(skips=2)), (10, dict(skips=1, sa=0)),
                                  (13, dict(sec=100, f0=5, f=3, rate=2, skips=3, rs0=3)), (16, {})])
def test_mul_fib_table(gpu_ctx, oracle, n, kw):
    import whirlaway_amd as wa
    _say(f"program table MulFib n={n} {kw}")
    b = oracle.Bench(5, seed=1)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    prog = mul_fib_program()
    pre, wit = mul_fib_columns(n, n)
    t = wa.air.ProgramAirTable(gpu_ctx, prog, n, 3, preprocessed=pre, univariate_skips=kw.get("skips", 4))
    got = t.prove(_settings(kw), p16, p24, wit)
    t.close()
    g = oracle.GenericAir(4, 1, n, pre, prog.code(), prog.n_instr, 3)
    assert np.array_equal(got, g.prove(wit, p16, p24, **kw))
    assert g.verify(got, p16, p24, **kw) == 0
    b.close()


def test_mul_fib_into_caller_transcript(gpu_ctx, oracle):
    import whirlaway_amd as wa
    from whirlaway_amd.transcript import ProverState
    n = 10
    _say("program table MulFib into a caller's transcript")
    b = oracle.Bench(5, seed=1)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    prog = mul_fib_program()
    pre, wit = mul_fib_columns(n, 3)
    dom = np.arange(20, dtype=np.uint32)
    ours = ProverState(p16, dom)
    ours.add_base_scalars([7, 8, 9])
    t = wa.air.ProgramAirTable(gpu_ctx, prog, n, 3, preprocessed=pre)
    t.prove_into(wa.AirSettings(), p16, p24, wit, ours)
    t.close()
    ref = oracle.Transcript(p16, dom)
    ref.add_base(np.array([7, 8, 9], dtype=np.uint32))
    oracle.GenericAir(4, 1, n, pre, prog.code(), prog.n_instr, 3).prove_into(ref, wit, p16, p24)
    assert np.array_equal(ours.proof_data(), ref.proof_data())
    b.close()
