"""AirSettings validation and the WHIR soundness assumptions (no GPU needed).

wl_air_settings_validate rejects shapes whose codeword domains exceed KoalaBear's
two-adicity (2^24): the reference panics there (WhirConfig::new / the DFT's
two_adic_generator), so the build returns WL_ERR_ARG before any device work.
The oracle proves and verifies under each SecurityAssumption.
"""
import numpy as np
import pytest

P = 0x7F000001


def _settings(**kw):
    import whirlaway_amd as wa
    return wa.AirSettings(**kw)


def test_default_settings_valid_up_to_two_adicity():
    # first codeword: 2^(n + 8 + rate - f0) rows; defaults: n + 2 <= 24
    for n in (5, 12, 17, 22):
        _settings().validate(n)
    from whirlaway_amd import WhirlawayError
    for n in (23, 24):
        with pytest.raises(WhirlawayError, match="two-adicity"):
            _settings().validate(n)


def test_rate3_folding4_domain_limit():
    from whirlaway_amd import WhirlawayError
    s = _settings(log_inv_rate=3, folding_factor_first=4, folding_factor=2, rs_domain_initial_reduction=2,
                  security_bits=90, univariate_skips=6)
    s.validate(17)  # 17 + 8 + 3 - 4 = 24
    with pytest.raises(WhirlawayError, match="two-adicity"):
        s.validate(18)


def test_later_round_domain_checked():
    """With a small initial reduction and a small second folding factor the second
    codeword is larger than the first: 2^(v + rate - rs0 - f) > 2^(v + rate - f0)."""
    from whirlaway_amd import WhirlawayError
    s = _settings(folding_factor_first=7, folding_factor=2, rs_domain_initial_reduction=1)
    s.validate(15)  # v = 23: first 2^17, second 2^21
    with pytest.raises(WhirlawayError, match="two-adicity"):
        s.validate(19)  # first 2^21, second 2^25


def test_argument_errors():
    from whirlaway_amd import WhirlawayError
    with pytest.raises(WhirlawayError):
        _settings(univariate_skips=12).validate(12)
    with pytest.raises(WhirlawayError):
        _settings(rs_domain_initial_reduction=8).validate(12)
    with pytest.raises(WhirlawayError):
        _settings(soundness=3).validate(12)
    with pytest.raises(WhirlawayError):
        _settings(log_inv_rate=0).validate(12)


def test_oracle_config_agrees_on_the_limit(oracle):
    oracle.whir_config(24 + 8 - 1 - 1)  # n = 22 with defaults: 2^24 rows
    with pytest.raises(ValueError):
        oracle.whir_config(23 + 8)
    with pytest.raises(ValueError):
        oracle.whir_config(26, rate=3, f0=4, f=2, rs0=2, sec=90)


def test_footprint_grows_with_rate_and_small_folding():
    from whirlaway_amd.engine import job_footprint
    base = job_footprint(16)
    wide = job_footprint(16, _settings(log_inv_rate=3, folding_factor_first=4, folding_factor=2,
                                       rs_domain_initial_reduction=2, security_bits=90, univariate_skips=6))
    assert wide > 2 * base
    assert job_footprint(17) > job_footprint(16) > job_footprint(10)


@pytest.mark.parametrize("sa", [0, 1, 2])
def test_oracle_soundness_variants_prove_and_verify(oracle, sa):
    """Each SecurityAssumption gives a schedule the oracle proves and verifies; a
    changed word is rejected."""
    b = oracle.Bench(7, seed=9)
    proof = b.prove(sa=sa)
    assert b.verify(proof, sa=sa) == 0
    bad = proof.copy()
    bad[len(bad) // 3] = (int(bad[len(bad) // 3]) + 1) % P
    assert b.verify(bad, sa=sa) != 0
    b.close()


def test_soundness_schedules_differ(oracle):
    """Unique decoding sends no OOD samples and needs the most queries; the Johnson
    bound needs twice the capacity-bound queries (a query detects a far word with
    probability 1 - sqrt(rho) instead of 1 - rho)."""
    ud = oracle.whir_config(20, sa=0)
    jb = oracle.whir_config(20, sa=1)
    cb = oracle.whir_config(20, sa=2)
    assert ud["commit_ood"] == 0 and all(r["ood_samples"] == 0 for r in ud["rounds"])
    assert cb["commit_ood"] >= 1 and jb["commit_ood"] >= 1
    for a, b, c in zip(ud["rounds"], jb["rounds"], cb["rounds"]):
        assert a["num_queries"] > b["num_queries"] > c["num_queries"]
        assert b["num_queries"] == -(-2 * 112 // b["log_inv_rate"])
