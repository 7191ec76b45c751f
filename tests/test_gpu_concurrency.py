"""Independent contexts proving on separate threads, some destroying their
context while others still prove: every proof equals the one proven alone.
(A context's destroy frees device memory, which waits for the whole device; it
must wait for the other contexts' proofs with device-side waits instead of
stalling them.)"""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("device_waits", [True, False])
def test_threads_with_own_contexts(gpu_ctx, device_waits):
    import whirlaway_amd as wa
    from whirlaway_amd.air import bench_inputs, generate_witness
    n = 10
    reps = [2, 5, 9, 14]  # threads finish (and destroy their contexts) at different times
    jobs = [bench_inputs(400 + i, n) for i in range(len(reps))]
    want = []
    for rc, inp, p16, p24 in jobs:
        t = wa.AirTable(gpu_ctx, rc, n)
        w = generate_witness(t, inp)
        want.append(t.prove(wa.AirSettings(), p16, p24, w))
        w.close()
        t.close()
    errors, got = [], [[] for _ in reps]
    start = threading.Barrier(len(reps))

    def worker(i):
        try:
            rc, inp, p16, p24 = jobs[i]
            ctx = wa.Context(0)
            ctx.set_device_waits(device_waits)
            t = wa.AirTable(ctx, rc, n)
            w = generate_witness(t, inp)
            start.wait()
            for _ in range(reps[i]):
                got[i].append(t.prove(wa.AirSettings(), p16, p24, w))
            w.close()
            t.close()
            ctx.close()
        except Exception as e:  # reported below
            errors.append(f"thread {i}: {e}")

    _say(f"{len(reps)} threads, device waits {device_waits}")
    ths = [threading.Thread(target=worker, args=(i,)) for i in range(len(reps))]
    for th in ths:
        th.start()
    for th in ths:
        th.join(timeout=120)
        assert not th.is_alive(), "a proving thread did not finish"
    assert not errors, errors
    for i, proofs in enumerate(got):
        assert len(proofs) == reps[i]
        for p in proofs:
            assert np.array_equal(p, want[i]), i
