"""GPU parity: every HIP entry point against the CPU oracle on the same inputs.

Integer/field work, so the bar is bit-exact equality.  Full proofs are compared
byte for byte (proof_data) for n up to 12 and checked against the oracle's
verifier at the benchmark size n = 17.
"""
import sys

import numpy as np
import pytest

import pyfield as pf

pytestmark = pytest.mark.gpu
P = pf.P


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edge_states(rng, width, n_random):
    """Random states plus ones built from the extremes of the field: 0, 1, p - 1,
    p - 2, values just below and above powers of two (the shift-based internal
    diagonal splits x at bit k), and rows mixing them."""
    special = np.array([0, 1, 2, P - 1, P - 2, (P + 1) // 2, (P - 1) // 2]
                       + [(1 << k) + d for k in range(1, 31) for d in (-1, 0, 1)], dtype=np.int64)
    special = special[(special >= 0) & (special < P)].astype(np.uint32)
    rows = [np.full(width, v, dtype=np.uint32) for v in special[:16]]
    rows += [rng.choice(special, width) for _ in range(240)]
    rows += list(rng.integers(0, P, (n_random, width), dtype=np.uint32))
    return np.array(rows, dtype=np.uint32)


def test_poseidon2_permutations(gpu_ctx, oracle):
    from whirlaway_amd import utils
    b = oracle.Bench(5, seed=3)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    rng = np.random.default_rng(5)
    for width in (16, 24):
        _say(f"poseidon2 width {width}")
        st = _edge_states(rng, width, 1024)
        got = utils.poseidon2_permute(gpu_ctx, p16, p24, width, st).reshape(-1, width)
        for i in range(st.shape[0]):
            assert np.array_equal(got[i], b.permute(width, st[i])), (width, i)


@pytest.mark.parametrize("log_n,k", [(10, 1), (12, 4), (16, 4), (18, 7)])
def test_fold_large(gpu_ctx, oracle, log_n, k):
    from whirlaway_amd import utils
    _say(f"fold_large log_n={log_n} k={k}")
    rng = np.random.default_rng(log_n * 31 + k)
    m = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    sc = rng.integers(0, P, 8 << k, dtype=np.uint32)
    assert np.array_equal(utils.fold_multilinear_in_large_field(gpu_ctx, m, sc),
                          oracle.fold_f_large(m, sc, k))
    me = rng.integers(0, P, (1 << (log_n - 2)) * 8, dtype=np.uint32)
    assert np.array_equal(utils.fold_multilinear_in_large_field(gpu_ctx, me, sc, is_ef=True),
                          oracle.fold_ef_large(me, sc, k))


@pytest.mark.parametrize("log_n,folds", [(14, 14), (13, 9), (12, 5), (2, 2)])
def test_fold_chain_per_round(gpu_ctx, oracle, log_n, folds):
    """Kernel micro (a) shape at reduced size: every intermediate table bit-exact, and
    the final table of the timed chain, whose folds of tables of at most 2^12 entries
    run in one launch (ending at one entry, at 16 and at 128 entries, and a chain of
    two folds)."""
    from whirlaway_amd import utils
    _say(f"fold_chain log_n={log_n}")
    rng = np.random.default_rng(7)
    m = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    ch = rng.integers(0, P, (folds, 8), dtype=np.uint32)
    final, allt, _ = utils.fold_chain(gpu_ctx, m, ch, all_tables=True)
    cur, off = None, 0
    for i in range(folds):
        sc = np.concatenate([(np.array(pf.esub(pf.ef(1), pf.ef(ch[i])), dtype=np.uint32)), ch[i]])
        cur = oracle.fold_f_large(m, sc, 1) if i == 0 else oracle.fold_ef_large(cur, sc, 1)
        ln = cur.size
        assert np.array_equal(allt[off: off + ln], cur), i
        off += ln
    assert np.array_equal(final, cur)


@pytest.mark.parametrize("n", [4, 9, 14, 15, 16, 17])
def test_sumcheck_a2b(gpu_ctx, oracle, n):
    """Kernel micro (b) shape at reduced size: transcript, challenges and final values.
    From n = 15 the first rounds take the eq weight split over a 2^13-entry table of
    the last variables and a table of the others (n = 15: round 0 split, round 1 the
    whole table; n = 16: rounds 0-1 split, then the whole table; n = 17: rounds 0-2
    split)."""
    from whirlaway_amd import sumcheck
    _say(f"sumcheck a2b n={n}")
    b = oracle.Bench(5, seed=11)
    rng = np.random.default_rng(n)
    a = rng.integers(0, P, 1 << n, dtype=np.uint32)
    bb = rng.integers(0, P, 1 << n, dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    want = b.sc_a2b(a, bb, z.reshape(-1).copy())
    got = sumcheck.prove_a2b(gpu_ctx, b.p2_consts(16), a, bb, z)
    assert np.array_equal(got["proof"], want["proof"])
    assert np.array_equal(got["challenges"], want["challenges"])
    assert np.array_equal(got["final_evals"], want["final_evals"])
    assert np.array_equal(got["final_sum"], want["final_sum"])


@pytest.mark.parametrize("n_cols,log_m,log_rows", [(4, 5, 6), (8, 10, 12), (3, 10, 13), (3, 13, 14), (2, 13, 16),
                                                   (16, 6, 15), (2, 18, 19), (1, 19, 20), (3, 13, 21),
                                                   (2, 18, 21), (1, 20, 22), (2, 22, 23), (1, 23, 24)])
def test_rs_commit(gpu_ctx, oracle, n_cols, log_m, log_rows):
    from whirlaway_amd import utils
    _say(f"rs_commit {n_cols}x2^{log_m} -> 2^{log_rows}")
    b = oracle.Bench(5, seed=13)
    rng = np.random.default_rng(log_rows)
    cols = rng.integers(0, P, (n_cols, 1 << log_m), dtype=np.uint32)
    root, cw = utils.rs_commit(gpu_ctx, b.p2_consts(16), b.p2_consts(24), cols, log_rows, True)
    L = oracle.lib()
    cw = cw.reshape(n_cols, 1 << log_rows)
    for c in range(n_cols):
        col = np.zeros(1 << log_rows, dtype=np.uint32)
        col[: 1 << log_m] = cols[c]
        L.or_mobius(oracle.ptr(col), log_m)
        L.or_ntt(oracle.ptr(col), log_rows)
        assert np.array_equal(cw[c], col), c
    if log_rows > 16:
        return  # the Python tree walk below is too slow; full proofs cover the tree at these sizes
    # (15 and 16 rows hash their leaves on lane pairs, 16 also its 2^15-node level)
    # Merkle root from the oracle's sponge/compression
    rows = cw.T.copy()
    level = []
    for r in range(rows.shape[0]):
        out = np.zeros(8, dtype=np.uint32)
        L.or_api_hash_leaf(b.h, oracle.ptr(np.ascontiguousarray(rows[r])), n_cols, oracle.ptr(out))
        level.append(out)
    while len(level) > 1:
        nxt = []
        for i in range(0, len(level), 2):
            out = np.zeros(8, dtype=np.uint32)
            L.or_api_compress(b.h, oracle.ptr(level[i]), oracle.ptr(level[i + 1]), oracle.ptr(out))
            nxt.append(out)
        level = nxt
    assert np.array_equal(root, level[0])


@pytest.mark.parametrize("n", [5, 6, 8, 10, 12])
def test_proof_bytes_identical(gpu_ctx, oracle, n):
    import whirlaway_amd as wa
    _say(f"full proof n={n}")
    b = oracle.Bench(n, seed=0)
    want = b.prove()
    table = wa.AirTable(gpu_ctx, b.air_consts(), n)
    got = table.prove(wa.AirSettings(), b.p2_consts(16), b.p2_consts(24), b.trace())
    table.close()
    assert got.size == want.size
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first mismatch at {diff[:5]}"
    assert b.verify(got) == 0


# Parameter variants (other security levels, folding factors, rates, skip counts
# and initial domain reductions); skips above 4 take the generic round-0 pass.
SETTINGS_VARIANTS = [
    (10, dict(sec=100, f0=5, f=3, rate=2, skips=2, rs0=3)),
    (11, dict(sec=128, f0=6, f=4, rate=1, skips=3, rs0=4)),
    (9, dict(sec=128, f0=7, f=4, rate=2, skips=1, rs0=1)),
    (12, dict(sec=128, f0=7, f=4, rate=1, skips=5, rs0=5)),
    (13, dict(sec=90, f0=4, f=2, rate=3, skips=6, rs0=2)),
]


@pytest.mark.parametrize("n,kw", SETTINGS_VARIANTS)
def test_proof_settings_variants(gpu_ctx, oracle, n, kw):
    import whirlaway_amd as wa
    _say(f"full proof n={n} settings {kw}")
    b = oracle.Bench(n, seed=3)
    want = b.prove(**kw)
    st = wa.AirSettings(security_bits=kw["sec"], folding_factor_first=kw["f0"], folding_factor=kw["f"],
                        log_inv_rate=kw["rate"], univariate_skips=kw["skips"],
                        rs_domain_initial_reduction=kw["rs0"])
    table = wa.AirTable(gpu_ctx, b.air_consts(), n, univariate_skips=kw["skips"])
    got = table.prove(st, b.p2_consts(16), b.p2_consts(24), b.trace())
    table.close()
    assert got.size == want.size
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first mismatch at {diff[:5]}"
    assert b.verify(got, **kw) == 0


# Rate 1/8 with a first folding factor of 4: the first codeword has
# 2^(n + 8 + 3 - 4) rows, 2^21 at n = 14 and 2^22 at n = 15, beyond the two-pass
# register NTT, so these proofs take the three-pass encoding.
@pytest.mark.parametrize("n,kw", [(14, dict(sec=90, f0=4, f=2, rate=3, skips=4, rs0=2)),
                                  (15, dict(sec=100, f0=4, f=4, rate=3, skips=4, rs0=3))])
def test_proof_three_pass_encoding(gpu_ctx, oracle, n, kw):
    import whirlaway_amd as wa
    _say(f"full proof n={n} settings {kw} (first codeword beyond 2^20 rows)")
    b = oracle.Bench(n, seed=5)
    st = wa.AirSettings(security_bits=kw["sec"], folding_factor_first=kw["f0"], folding_factor=kw["f"],
                        log_inv_rate=kw["rate"], univariate_skips=kw["skips"],
                        rs_domain_initial_reduction=kw["rs0"])
    table = wa.AirTable(gpu_ctx, b.air_consts(), n, univariate_skips=kw["skips"])
    got = table.prove(st, b.p2_consts(16), b.p2_consts(24), b.trace())
    table.close()
    _say(f"  oracle proof n={n}")
    want = b.prove(**kw)
    assert got.size == want.size
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first mismatch at {diff[:5]}"
    assert b.verify(got, **kw) == 0


@pytest.mark.parametrize("n,kw", [(8, {}), (12, {}), SETTINGS_VARIANTS[1]])
def test_device_transcript_rounds(oracle, monkeypatch, n, kw):
    """Zerocheck rounds whose transcript runs on the device (one wave's duplex
    sponge, opt-in per context) give the oracle's proof bytes."""
    import whirlaway_amd as wa
    _say(f"full proof n={n} {kw}, device-side zerocheck transcript")
    monkeypatch.setenv("WHIRLAWAY_DEVICE_ROUNDS", "1")
    ctx = wa.Context(0)
    b = oracle.Bench(n, seed=4)
    want = b.prove(**kw)
    st = wa.AirSettings(security_bits=kw.get("sec", 128), folding_factor_first=kw.get("f0", 7),
                        folding_factor=kw.get("f", 4), log_inv_rate=kw.get("rate", 1),
                        univariate_skips=kw.get("skips", 4), rs_domain_initial_reduction=kw.get("rs0", 5))
    table = wa.AirTable(ctx, b.air_consts(), n, univariate_skips=kw.get("skips", 4))
    got = table.prove(st, b.p2_consts(16), b.p2_consts(24), b.trace())
    table.close()
    ctx.close()
    assert np.array_equal(got, want)


def test_proof_benchmark_size_identical(gpu_ctx, oracle):
    """n = 17 (the benchmark workload): the GPU proof is byte-identical to the oracle's
    proof of the same inputs, the oracle verifier accepts it and rejects a changed
    word, and a resident-witness proof equals the host-witness proof."""
    import whirlaway_amd as wa
    from whirlaway_amd.air import Witness
    n = 17
    _say("full proof n=17")
    b = oracle.Bench(n, seed=0)
    table = wa.AirTable(gpu_ctx, b.air_consts(), n)
    tr = b.trace()
    got = table.prove(wa.AirSettings(), b.p2_consts(16), b.p2_consts(24), tr)
    _say("  oracle proof n=17")
    want = b.prove()
    assert got.size == want.size
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, f"first mismatch at {diff[:5]}"
    assert b.verify(got) == 0
    w = Witness(gpu_ctx, tr, n)
    again = table.prove(wa.AirSettings(), b.p2_consts(16), b.p2_consts(24), w)
    w.close()
    table.close()
    assert np.array_equal(got, again)
    bad = got.copy()
    bad[len(bad) // 2] = (int(bad[len(bad) // 2]) + 1) % P
    assert b.verify(bad) != 0


@pytest.mark.parametrize("sa", [0, 1])
def test_proof_soundness_variants(gpu_ctx, oracle, sa):
    """UniqueDecoding (no OOD samples, more queries) and JohnsonBound schedules:
    byte-identical to the oracle, which verifies them."""
    import whirlaway_amd as wa
    n = 10
    _say(f"full proof n={n} soundness {sa}")
    b = oracle.Bench(n, seed=5)
    want = b.prove(sa=sa)
    table = wa.AirTable(gpu_ctx, b.air_consts(), n)
    got = table.prove(wa.AirSettings(soundness=sa), b.p2_consts(16), b.p2_consts(24), b.trace())
    table.close()
    assert np.array_equal(got, want)
    assert b.verify(got, sa=sa) == 0


def test_settings_beyond_two_adicity_rejected(gpu_ctx):
    """A shape whose first codeword would need a 2^25 domain is refused before any
    device work (rate 1/8, first folding 4, n = 18)."""
    import whirlaway_amd as wa
    n = 18
    rc = np.zeros(148, dtype=np.uint32)
    table = wa.AirTable(gpu_ctx, rc, n, univariate_skips=4)
    st = wa.AirSettings(log_inv_rate=3, folding_factor_first=4, folding_factor=2, rs_domain_initial_reduction=2,
                        security_bits=90)
    with pytest.raises(wa.WhirlawayError, match="two-adicity"):
        table.prove(st, np.zeros(148, np.uint32), np.zeros(215, np.uint32), np.zeros(165 << n, np.uint32))
    table.close()


@pytest.mark.parametrize("n", [5, 11, 17])
def test_trace_generation(gpu_ctx, oracle, n):
    """generate_trace_rows on the GPU equals the oracle's trace."""
    import whirlaway_amd as wa
    from whirlaway_amd.air import bench_inputs, download_witness, generate_witness
    _say(f"trace generation n={n}")
    rc, inp, p16, p24 = bench_inputs(0, n)
    table = wa.AirTable(gpu_ctx, rc, n)
    w = generate_witness(table, inp)
    got = download_witness(w)
    w.close()
    table.close()
    assert np.array_equal(got, oracle.Bench(n, 0).trace())


def test_fold_chain_full_size(gpu_ctx, oracle):
    """Kernel micro (a) at its full size: 2^24 base-field values folded 24 times by
    extension challenges, every intermediate table bit-exact against the oracle."""
    from whirlaway_amd import utils
    log_n = folds = 24
    _say(f"fold_chain full size log_n={log_n}")
    rng = np.random.default_rng(24)
    m = rng.integers(0, P, 1 << log_n, dtype=np.uint32)
    ch = rng.integers(0, P, (folds, 8), dtype=np.uint32)
    final, allt, ms = utils.fold_chain(gpu_ctx, m, ch, all_tables=True)
    cur, off = None, 0
    for i in range(folds):
        sc = np.concatenate([(np.array(pf.esub(pf.ef(1), pf.ef(ch[i])), dtype=np.uint32)), ch[i]])
        cur = oracle.fold_f_large(m, sc, 1) if i == 0 else oracle.fold_ef_large(cur, sc, 1)
        ln = cur.size
        assert np.array_equal(allt[off: off + ln], cur), i
        off += ln
    assert np.array_equal(final, cur)


def test_sumcheck_a2b_full_size(gpu_ctx, oracle):
    """Kernel micro (b) at its full size: n = 24, proof words, challenges and final
    values bit-exact against the oracle."""
    from whirlaway_amd import sumcheck
    n = 24
    _say(f"sumcheck a2b full size n={n}")
    b = oracle.Bench(5, seed=11)
    rng = np.random.default_rng(1024)
    a = rng.integers(0, P, 1 << n, dtype=np.uint32)
    bb = rng.integers(0, P, 1 << n, dtype=np.uint32)
    z = rng.integers(0, P, (n, 8), dtype=np.uint32)
    got = sumcheck.prove_a2b(gpu_ctx, b.p2_consts(16), a, bb, z)
    want = b.sc_a2b(a, bb, z.reshape(-1).copy())
    for k in ("proof", "challenges", "final_evals", "final_sum"):
        assert np.array_equal(got[k], want[k]), k
    b.close()


@pytest.mark.parametrize("n", [8, 11])
def test_prove_into_caller_transcript(gpu_ctx, oracle, n):
    """AirTable::prove on a caller's ProverState (crates/air/src/prove.rs:54): the
    transcript already holds a domain separator and earlier messages, and the proof
    continues it.  proof_data equals the oracle's for the same prefix; the device
    grinding of a transcript equals the host grinding."""
    import whirlaway_amd as wa
    from whirlaway_amd.transcript import ProverState
    _say(f"prove_into n={n}")
    b = oracle.Bench(n, seed=21)
    p16, p24 = b.p2_consts(16), b.p2_consts(24)
    rng = np.random.default_rng(n)
    dom = rng.integers(0, P, 50, dtype=np.uint32)
    prefix = rng.integers(0, P, 13, dtype=np.uint32)
    ours = ProverState(p16, dom)
    ours.add_base_scalars(prefix)
    ours.pow_grinding(6, ctx=gpu_ctx)
    ref = oracle.Transcript(p16, dom)
    ref.add_base(prefix)
    ref.grind(6)
    table = wa.AirTable(gpu_ctx, b.air_consts(), n)
    table.prove_into(wa.AirSettings(), p16, p24, b.trace(), ours)
    table.close()
    b.prove_into(ref)
    got, want = ours.proof_data(), ref.proof_data()
    assert got.size == want.size and np.array_equal(got, want)
    # the transcripts end in the same state
    assert np.array_equal(ours.sample(2), ref.sample_ext(2))
    ours.close()
    # a transcript whose Poseidon2-16 differs from the Merkle compression's is refused
    other = ProverState(oracle.Bench(n, seed=22).p2_consts(16))
    table = wa.AirTable(gpu_ctx, b.air_consts(), n)
    with pytest.raises(wa.WhirlawayError):
        table.prove_into(wa.AirSettings(), p16, p24, b.trace(), other)
    table.close()


def test_product_verifier_accepts_gpu_proofs(gpu_ctx, oracle):
    """GPU proofs (2^17 Poseidon2 and a program AIR) pass the product verifier,
    alone and in a batch, and a changed word is rejected."""
    import whirlaway_amd as wa
    from whirlaway_amd.proof import verify_batch
    _say("product verifier on GPU proofs")
    n = 17
    rc, inp, p16, p24 = wa.air.bench_inputs(0, n)
    table = wa.AirTable(gpu_ctx, rc, n)
    w = wa.air.generate_witness(table, inp)
    proof = table.prove(wa.AirSettings(), p16, p24, w)
    w.close()
    bad = proof.copy()
    bad[1000] = (int(bad[1000]) + 1) % P
    assert verify_batch(table, wa.AirSettings(), p16, p24, [proof, bad, proof], n_threads=3)[::2] == [0, 0]
    assert verify_batch(table, wa.AirSettings(), p16, p24, [bad])[0] != 0
    table.close()
