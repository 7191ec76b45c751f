"""A device-side wait that reaches its bound ends its kernel without the host
value: the proof fails with an error instead of hanging or returning a wrong
proof, and the same context then proves correctly.  The bound is forced from the
host (wl_set_wait_timeout with a nanosecond limit, so every wait of the proof
ends before the host could answer); every kernel still exits normally."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_wait_timeout_fails_the_proof_and_the_context_recovers(oracle):
    import whirlaway_amd as wa
    n = 10
    b = oracle.Bench(n, seed=21)
    ctx = wa.Context(0)
    table = wa.AirTable(ctx, b.air_consts(), n)
    st = wa.AirSettings()
    p16, p24, tr = b.p2_consts(16), b.p2_consts(24), b.trace()
    try:
        ctx.set_wait_timeout(1e-9)
        with pytest.raises(wa.WhirlawayError, match="timed out"):
            table.prove(st, p16, p24, tr)
        ctx.set_wait_timeout(0)
        got = table.prove(st, p16, p24, tr)
        assert np.array_equal(got, b.prove())
        again = table.prove(st, p16, p24, tr)
        assert np.array_equal(got, again)
    finally:
        table.close()
        ctx.close()


def test_wait_timeout_rejects_out_of_range_values():
    import whirlaway_amd as wa
    ctx = wa.Context(0)
    try:
        for bad in (-1.0, 1e9, float("nan")):
            with pytest.raises(wa.WhirlawayError):
                ctx.set_wait_timeout(bad)
    finally:
        ctx.close()
