"""The product verifier (wl_air_verify, host code) and the proof wire format, on
the CPU: it accepts the oracle's proofs of the Poseidon2 AIR and of program AIRs
(with preprocessed columns, under every soundness assumption), rejects every
changed word with the same AirVerifError kind as the oracle's verifier, and the
batch verifier agrees with one-by-one verification."""
import numpy as np
import pytest

from airs import P, mul_fib_columns, mul_fib_program

KW = {"sec": "security_bits", "f0": "folding_factor_first", "f": "folding_factor", "rate": "log_inv_rate",
      "skips": "univariate_skips", "rs0": "rs_domain_initial_reduction", "sa": "soundness"}


def _settings(kw):
    import whirlaway_amd as wa
    return wa.AirSettings(**{KW[k]: v for k, v in kw.items()})


def _verify(table, st, p16, p24, dom, proof):
    from whirlaway_amd.proof import VerificationError, verify
    from whirlaway_amd.transcript import VerifierState
    vs = VerifierState(p16, dom, proof)
    try:
        verify(table, st, p16, p24, vs)
        return 0
    except VerificationError as e:
        return e.kind


@pytest.mark.parametrize("n,kw", [(8, {}), (9, dict(sa=0)), (9, dict(sa=1)),
                                  (10, dict(sec=100, f0=5, f=3, rate=2, skips=2, rs0=3))])
def test_poseidon2_oracle_proofs(oracle, n, kw):
    import whirlaway_amd as wa
    from whirlaway_amd.transcript import air_domain_separator
    b = oracle.Bench(n, seed=7)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    st = _settings(kw)
    proof = b.prove(**kw)
    table = wa.AirTable(None, b.air_consts(), n, univariate_skips=st.univariate_skips)
    dom = air_domain_separator(n, st)
    assert _verify(table, st, p16, p24, dom, proof) == 0
    rng = np.random.default_rng(n)
    positions = sorted(set(rng.integers(0, proof.size, 40).tolist()) | {0, 8, proof.size - 1})
    for pos in positions:
        bad = proof.copy()
        bad[pos] = (int(bad[pos]) + 1 + int(rng.integers(0, 1000))) % P
        want = b.verify(bad, **kw)
        assert want != 0
        assert _verify(table, st, p16, p24, dom, bad) == want, pos
    assert _verify(table, st, p16, p24, dom, proof[:-3]) != 0
    table.close()
    b.close()


def test_program_air_oracle_proofs(oracle):
    import whirlaway_amd as wa
    n = 10
    b = oracle.Bench(5, seed=1)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    prog = mul_fib_program()
    pre, wit = mul_fib_columns(n, 11)
    g = oracle.GenericAir(4, 1, n, pre, prog.code(), prog.n_instr, 3)
    proof = g.prove(wit, p16, p24)
    table = wa.air.ProgramAirTable(None, prog, n, 3, preprocessed=pre)
    dom = g.domain_separator()
    st = wa.AirSettings()
    assert _verify(table, st, p16, p24, dom, proof) == 0
    for pos in (0, 20, proof.size // 2, proof.size - 2):
        bad = proof.copy()
        bad[pos] = (int(bad[pos]) + 5) % P
        assert _verify(table, st, p16, p24, dom, bad) == g.verify(bad, p16, p24) != 0
    # the verifier's own preprocessed column must match the prover's
    pre2 = pre.copy()
    pre2[0, 1] = (int(pre2[0, 1]) + 1) % P
    t2 = wa.air.ProgramAirTable(None, prog, n, 3, preprocessed=pre2)
    assert _verify(t2, st, p16, p24, dom, proof) != 0
    # a broken witness gives a proof that is rejected
    broken = wit.copy()
    broken[1, 7] = (int(broken[1, 7]) + 1) % P
    assert _verify(table, st, p16, p24, dom, g.prove(broken, p16, p24)) == 6
    table.close()
    t2.close()
    b.close()


def test_batch_verifier_matches_single(oracle):
    import whirlaway_amd as wa
    from whirlaway_amd.proof import verify_batch
    n = 7
    b = oracle.Bench(n, seed=2)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    good = b.prove()
    proofs = []
    for i in range(12):
        p = good.copy()
        if i % 3 == 1:
            p[(97 * i) % p.size] = (int(p[(97 * i) % p.size]) + 1) % P
        if i % 3 == 2:
            p = np.concatenate([p, np.zeros(2, np.uint32)])  # trailing words
        proofs.append(p)
    table = wa.AirTable(None, b.air_consts(), n)
    res = verify_batch(table, wa.AirSettings(), p16, p24, proofs, n_threads=4)
    for i, r in enumerate(res):
        if i % 3 == 0:
            assert r == 0
        elif i % 3 == 1:
            assert r == b.verify(proofs[i]) != 0
        else:
            assert r == 3
    table.close()
    b.close()


def test_wire_format_round_trip(oracle):
    import whirlaway_amd as wa
    from whirlaway_amd import WhirlawayError
    from whirlaway_amd.proof import deserialize, serialize
    b = oracle.Bench(6, seed=3)
    proof = b.prove()
    st = wa.AirSettings(soundness=1)
    data = serialize(proof, 6, 165, st)
    assert len(data) == 48 + (31 * proof.size + 7) // 8 + 4
    back, ll, nw, st2 = deserialize(data)
    assert np.array_equal(back, proof) and ll == 6 and nw == 165 and st2 == st
    extremes = np.array([0, 1, P - 1, P - 2, 1 << 30, (1 << 31) - (1 << 24)], dtype=np.uint32)
    assert np.array_equal(deserialize(serialize(extremes, 3, 2))[0], extremes)
    for i in (0, 5, 60, len(data) - 1):
        bad = bytearray(data)
        bad[i] ^= 0x10
        with pytest.raises(WhirlawayError):
            deserialize(bytes(bad))
    with pytest.raises(WhirlawayError):
        deserialize(data[:-1])
    with pytest.raises(WhirlawayError):
        serialize(np.array([P], dtype=np.uint32), 3, 2)
    b.close()
