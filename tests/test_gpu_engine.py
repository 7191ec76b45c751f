"""Batched / served proving (wl_engine): each proof equals the one proven alone,
admission respects the memory budget, and the footprint estimate covers the
measured peak of a proof."""
import sys
import time

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


FOOTPRINT_CASES = [
    (8, {}), (12, {}), (15, {}), (16, {}), (18, {}),
    (14, dict(log_inv_rate=3, folding_factor_first=4, folding_factor=2, rs_domain_initial_reduction=2,
              security_bits=90, univariate_skips=4)),
    (15, dict(log_inv_rate=2, folding_factor_first=5, folding_factor=3, rs_domain_initial_reduction=3,
              security_bits=100, univariate_skips=2)),
]


@pytest.mark.parametrize("n,kw", FOOTPRINT_CASES)
def test_footprint_covers_measured_peak(n, kw):
    import whirlaway_amd as wa
    from whirlaway_amd.air import generate_witness
    from whirlaway_amd.engine import job_footprint
    st = wa.AirSettings(**kw)
    _say(f"footprint n={n} {kw}")
    with wa.Context(0) as ctx:
        rc, inp, p16, p24 = _inputs(3, n)
        ctx.mem_stats(reset_peak=True)
        table = wa.AirTable(ctx, rc, n, univariate_skips=st.univariate_skips)
        w = generate_witness(table, inp)
        table.prove(st, p16, p24, w)
        w.close()
        table.close()
        held, peak = ctx.mem_stats()
    fp = job_footprint(n, st)
    _say(f"  peak {peak / 2**20:.1f} MiB, footprint {fp / 2**20:.1f} MiB")
    assert peak <= fp


def _seeded_requests(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(lo, hi + 1, count)]
    return [(n, 5000 + seed * 100 + i) for i, n in enumerate(sizes)]


def test_batched_mix_64_against_oracle(gpu_ctx, oracle):
    """The batched workload: 64 seeded proofs with log_n ~ U{10..16} through the engine.
    Every proof equals the one proven alone on a plain context, and for each size
    10..16 one proof equals the oracle's proof of the same inputs."""
    from whirlaway_amd.engine import Engine
    reqs = _seeded_requests(1, 64, 10, 16)
    for n in range(10, 17):  # every size present at least once
        if not any(m == n for m, _ in reqs):
            reqs[n - 10] = (n, reqs[n - 10][1])
    jobs = [(n, sd) + _inputs(sd, n) for n, sd in reqs]
    _say(f"batched mix: {len(jobs)} proofs, sizes {sorted(set(n for n, _ in reqs))}")
    with Engine(0, n_slots=8) as eng:
        tickets = [eng.submit(n, rc, inp, p16, p24) for n, _, rc, inp, p16, p24 in jobs]
        got = [eng.wait(t)[0] for t in tickets]
        st = eng.stats()
    _say(f"engine stats {st}")
    assert st["completed"] == len(jobs) and st["peak_running"] >= 2
    checked = set()
    for (n, sd, rc, inp, p16, p24), proof in zip(jobs, got):
        assert np.array_equal(proof, _alone(gpu_ctx, n, rc, inp, p16, p24)), (n, sd)
        if n not in checked:
            _say(f"  oracle proof n={n} seed={sd}")
            b = oracle.Bench(n, seed=sd)
            assert np.array_equal(proof, b.prove()), (n, sd)
            b.close()
            checked.add(n)
    assert checked == set(range(10, 17))


def test_serving_replay_against_oracle(gpu_ctx, oracle):
    """The serving workload: requests with log_n ~ U{10..18} arriving faster than one
    slot serves them (1 ms apart), replayed open-loop against a 4-slot engine.  Jobs
    overlap on the device, every proof equals the one proven alone, and an n = 18
    proof equals the oracle's."""
    import time
    from whirlaway_amd.engine import Engine
    reqs = _seeded_requests(2, 24, 10, 18)
    reqs[5] = (18, reqs[5][1])
    jobs = [(n, sd) + _inputs(sd, n) for n, sd in reqs]
    _say(f"serving replay: {len(jobs)} requests, sizes {[n for n, _ in reqs]}")
    with Engine(0, n_slots=4) as eng:
        t0 = time.perf_counter()
        tickets = []
        for i, (n, sd, rc, inp, p16, p24) in enumerate(jobs):
            while time.perf_counter() < t0 + 0.001 * i:
                pass
            tickets.append(eng.submit(n, rc, inp, p16, p24))
        res = [eng.wait(t) for t in tickets]
        st = eng.stats()
    _say(f"engine stats {st}")
    assert st["peak_running"] >= 2
    for (n, sd, rc, inp, p16, p24), (proof, times) in zip(jobs, res):
        assert times["latency_ms"] >= times["run_ms"] >= 0
        assert np.array_equal(proof, _alone(gpu_ctx, n, rc, inp, p16, p24)), (n, sd)
    n, sd = reqs[5]
    _say(f"  oracle proof n={n} seed={sd}")
    b = oracle.Bench(n, seed=sd)
    assert np.array_equal(res[5][0], b.prove())
    b.close()


def test_engine_destroy_completes_queued_waiters():
    """Destroying an engine with queued tickets returns an error to their waiters
    instead of leaving them blocked; the running job finishes normally."""
    import threading
    from whirlaway_amd import WhirlawayError
    from whirlaway_amd.engine import Engine
    n = 15
    job = _inputs(300, n)
    jobs = [job] * 30  # ~0.3 s of work on one slot: most tickets are still queued below
    eng = Engine(0, n_slots=1)
    tickets = [eng.submit(n, *j) for j in jobs]
    outcome = {}

    def waiter(t):
        try:
            eng.wait(t)
            outcome[t] = "proof"
        except WhirlawayError as e:
            outcome[t] = str(e)

    ths = [threading.Thread(target=waiter, args=(t,)) for t in tickets[1:]]
    for th in ths:
        th.start()
    time.sleep(0.03)  # the waiters are inside wl_engine_wait
    _say("destroying the engine with waiters blocked")
    eng.close()
    for th in ths:
        th.join(timeout=60)
        assert not th.is_alive(), "a waiter stayed blocked after wl_engine_destroy"
    assert len(outcome) == len(tickets) - 1
    assert any("stopped" in v for v in outcome.values())
