"""Batched / served proving (wl_engine): each proof equals the one proven alone,
admission respects the memory budget, and the footprint estimate covers the
measured peak of a proof."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inputs(seed, n):
    from whirlaway_amd.air import bench_inputs
    return bench_inputs(seed, n)


def _alone(ctx, n, rc, inputs, p16, p24):
    import whirlaway_amd as wa
    from whirlaway_amd.air import generate_witness
    table = wa.AirTable(ctx, rc, n)
    w = generate_witness(table, inputs)
    proof = table.prove(wa.AirSettings(), p16, p24, w)
    w.close()
    table.close()
    return proof


def test_engine_mix_equals_alone(gpu_ctx, oracle):
    from whirlaway_amd.engine import Engine
    sizes = [6, 9, 11, 7, 10, 8, 11, 9, 6, 10]
    jobs = [(n, 100 + i) + _inputs(100 + i, n) for i, n in enumerate(sizes)]
    _say(f"engine mix sizes {sizes}")
    want = [_alone(gpu_ctx, n, rc, inp, p16, p24) for n, _, rc, inp, p16, p24 in jobs]
    # pin the single-proof path itself on two of them
    for k in (0, 3):
        b = oracle.Bench(sizes[k], seed=jobs[k][1])
        assert np.array_equal(b.prove(), want[k])
        b.close()
    with Engine(0, n_slots=4) as eng:
        tickets = [eng.submit(n, rc, inp, p16, p24) for n, _, rc, inp, p16, p24 in jobs]
        got = [eng.wait(t) for t in tickets]
        st = eng.stats()
    _say(f"engine stats {st}")
    assert st["completed"] == len(jobs)
    for (proof, times), w in zip(got, want):
        assert np.array_equal(proof, w)
        assert times["latency_ms"] >= times["run_ms"] >= 0


def test_engine_admission_respects_budget(gpu_ctx):
    from whirlaway_amd.engine import Engine, job_footprint
    n = 9
    jobs = [_inputs(200 + i, n) for i in range(5)]
    budget = int(job_footprint(n) * 1.5)  # room for one job at a time
    _say("engine admission")
    with Engine(0, n_slots=4, mem_budget=budget) as eng:
        tickets = [eng.submit(n, *j) for j in jobs]
        got = [eng.wait(t) for t in tickets]
        st = eng.stats()
        with pytest.raises(Exception):
            eng.submit(16, *_inputs(1, 16))  # footprint above the budget
    assert st["peak_running"] == 1 and st["peak_reserved"] <= budget
    for (proof, _), j in zip(got, jobs):
        assert np.array_equal(proof, _alone(gpu_ctx, n, *j))


@pytest.mark.parametrize("n", [8, 12, 15])
def test_footprint_covers_measured_peak(n):
    import whirlaway_amd as wa
    from whirlaway_amd.engine import job_footprint
    _say(f"footprint n={n}")
    with wa.Context(0) as ctx:
        rc, inp, p16, p24 = _inputs(3, n)
        ctx.mem_stats(reset_peak=True)
        _alone(ctx, n, rc, inp, p16, p24)
        held, peak = ctx.mem_stats()
    _say(f"  peak {peak / 2**20:.1f} MiB, footprint {job_footprint(n) / 2**20:.1f} MiB")
    assert peak <= job_footprint(n)
