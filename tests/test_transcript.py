"""The transcript handle (wl_transcript, host code) against the oracle's
ProverState restatement: the same operations give the same proof_data and the
same challenges; export/import continues a transcript; the verifier side reads a
proof back and rejects a bad proof-of-work witness.  No GPU needed."""
import numpy as np
import pytest

P = 0x7F000001


def _p16():
    import whirlaway_amd as wa
    from whirlaway_amd.air import bench_inputs
    return bench_inputs(7, 4, with_inputs=False)[2]


def test_operation_sequence_matches_oracle(oracle):
    from whirlaway_amd.transcript import ProverState
    p16 = _p16()
    rng = np.random.default_rng(3)
    dom = rng.integers(0, P, 37, dtype=np.uint32)
    ours, ref = ProverState(p16, dom), oracle.Transcript(p16, dom)
    for step in range(60):
        op = step % 6
        if op == 0:
            v = rng.integers(0, P, int(rng.integers(1, 20)), dtype=np.uint32)
            ours.add_base_scalars(v)
            ref.add_base(v)
        elif op == 1:
            v = rng.integers(0, P, (int(rng.integers(1, 4)), 8), dtype=np.uint32)
            ours.add_extension_scalars(v)
            ref.add_base(v)
        elif op == 2:
            v = rng.integers(0, P, int(rng.integers(1, 30)), dtype=np.uint32)
            ours.hint_base_scalars(v)
            ref.hint(v)
        elif op == 3:
            k = int(rng.integers(1, 3))
            assert np.array_equal(ours.sample(k), ref.sample_ext(k))
        elif op == 4:
            b = int(rng.integers(1, 25))
            assert ours.sample_bits(b) == ref.sample_bits(b)
        else:
            bits = int(rng.integers(0, 9))
            ours.pow_grinding(bits)
            ref.grind(bits)
        v = rng.integers(0, P, int(rng.integers(0, 5)), dtype=np.uint32)
        ours.observe(v)
        ref.observe(v)
    assert np.array_equal(ours.proof_data(), ref.proof_data())
    assert np.array_equal(ours.sample(3), ref.sample_ext(3))


def test_export_import_continues(oracle):
    from whirlaway_amd.transcript import ProverState
    p16 = _p16()
    a = ProverState(p16, [1, 2, 3])
    a.add_base_scalars(np.arange(11, dtype=np.uint32))
    a.sample(1)
    a.observe([5])
    st = a.export_state()
    b = ProverState(p16)
    b.import_state(st)
    assert np.array_equal(a.sample(4), b.sample(4))
    bad = st.copy()
    bad[24] = 9  # input count beyond the rate
    with pytest.raises(Exception):
        b.import_state(bad)


def test_verifier_reads_back_and_checks_pow(oracle):
    from whirlaway_amd import WhirlawayError
    from whirlaway_amd.transcript import ProverState, VerifierState
    p16 = _p16()
    dom = [9, 8, 7]
    pr = ProverState(p16, dom)
    pr.add_base_scalars([1, 2, 3])
    c1 = pr.sample(1)
    pr.add_extension_scalars(c1)
    pr.hint_base_scalars([P - 1, 0])
    w = pr.pow_grinding(10)
    c2 = pr.sample(1)
    proof = pr.proof_data()
    vs = VerifierState(p16, dom, proof)
    assert np.array_equal(vs.next_base_scalars_vec(3), [1, 2, 3])
    assert np.array_equal(vs.sample(1), c1)
    assert np.array_equal(vs.next_extension_scalars_vec(1), c1)
    assert np.array_equal(vs.receive_hint_base_scalars(2), [P - 1, 0])
    vs.check_pow_grinding(10)
    assert np.array_equal(vs.sample(1), c2)
    assert vs.remaining() == 0
    bad = proof.copy()
    bad[-1] = (w + 1) % P
    vb = VerifierState(p16, dom, bad)
    vb.next_base_scalars_vec(3)
    vb.sample(1)
    vb.next_extension_scalars_vec(1)
    vb.receive_hint_base_scalars(2)
    with pytest.raises(WhirlawayError):
        vb.check_pow_grinding(10)
    with pytest.raises(WhirlawayError):
        VerifierState(p16, dom, proof[:2]).next_base_scalars_vec(3)


@pytest.mark.parametrize("n,kw", [(12, {}), (17, {}), (10, dict(sec=100, f0=5, f=3, rate=2, skips=2, rs0=3)),
                                  (9, dict(sa=0)), (9, dict(sa=1))])
def test_domain_separator_matches_oracle(oracle, n, kw):
    import whirlaway_amd as wa
    from whirlaway_amd.transcript import air_domain_separator
    st = wa.AirSettings(security_bits=kw.get("sec", 128), folding_factor_first=kw.get("f0", 7),
                        folding_factor=kw.get("f", 4), log_inv_rate=kw.get("rate", 1),
                        univariate_skips=kw.get("skips", 4), rs_domain_initial_reduction=kw.get("rs0", 5),
                        soundness=kw.get("sa", 2))
    assert np.array_equal(air_domain_separator(n, st), oracle.domain_separator(n, **kw))
