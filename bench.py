#!/usr/bin/env python3
"""Benchmark: Poseidon2 permutations proven per second on one MI355X.

Workload: the Poseidon2-16 KoalaBear AIR with 2^17 permutations (one trace row
each), inputs from StdRng::seed_from_u64(0) as in src/examples/poseidon2.rs,
default AirSettings of src/main.rs (128-bit, CapacityBound, folding (7, 4),
rate 1/2, 4 univariate skips, RS reduction 5).  One step = one full proof
(AirTable::prove, crates/air/src/prove.rs:49-226) with the witness already
resident in HBM; proof bytes land on the host at the end of every step.

Prints one JSON line.  The CPU baseline is the oracle (a C restatement of the
same prover, OpenMP) timed on this host on one proof of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Per-kernel-class roofline models: (name, bound, algorithmic work per unit, unit name).
# VALU classes count field multiplications (F-mul) as in the per-unit figures of the
# work model: Poseidon2-24 982, Poseidon2-16 616 F-mul per permutation; one EF
# AIR evaluation point 31168 F-mul; one skip-round point 5700 F-mul.
# VALU peak: 256 CUs x 64 lanes x 2.4 GHz full-rate instructions (the 32-bit integer
# multiplies v_mad_u64_u32 / v_mul_lo_u32 issue at full rate on gfx950, as the live
# Montgomery-product probe below shows) over the 5 instructions a reduced Montgomery
# product needs at least (product, m = lo (-1/p), the multiply-add m p + x, then a
# subtraction of p and an unsigned minimum).
VALU_PEAK_TFMUL = 256 * 64 * 2.4e9 / 5 / 1e12
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each kernel class at 2^17 permutations, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE passes (profiles/prove_n17_hbm_traffic.md), fetch scaled x2.0
# and write x1.0 by the same counters on a 256 MiB device copy.  (read MiB, write MiB)
MEASURED_TRAFFIC_N17 = {
    1: (70.03, 5.76),     # k_merkle_leaves (3 launches) + k_merkle_leaves_pair (1), averaged
    3: (92.86, 139.00),   # k_ntt2_a + k_ntt2_b (one rs_encode, averaged over the 4 commits)
}
# VALU instructions per Poseidon2-24 leaf permutation (SQ_INSTS_VALU per wave of the
# 2^19-row commit / 8 permutations per thread, profiles/prove_n17_valu_issue.md).  The
# leaf class's live permutation rate times this count, over the VALU issue rate of
# 1024 SIMDs x 2.4 GHz at 4 cycles per wave64 instruction, is `valu_issue`: the share of
# the issue rate the leaf hashing uses (the F-mul `frac` counts only its products).
VALU_INSTR_PER_P24 = 6438
PHASES = ["commit", "zerocheck_round0", "zerocheck_rounds", "batching", "inner_sumcheck", "whir_open",
          "host_wait_for_device"]
MODELS = {
    1: ("merkle_leaves (Poseidon2-24 sponge)", "valu", 982.0),
    2: ("zerocheck_rounds (EF AIR eval, rounds>=1)", "valu", 31168.0),
    3: ("rs_encode (Mobius + NTT)", "hbm", 1.0),          # work already in bytes
    4: ("zerocheck_round0 (skip LDE + F AIR eval)", "valu", 5700.0),
    5: ("grinding (Poseidon2-16)", "valu", 616.0),
    6: ("merkle_nodes (Poseidon2-16 compress)", "valu", 616.0),
}


def _progress(msg):
    """One line on stderr per finished leg (the JSON line alone goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def run_extra(args, ctx, table, rc, p16, p24, wit, proof):
    """The same 2^n workload through the generic AIR path (the Poseidon2 AIR as a
    constraint program, proved by the generic sumcheck) with the proof compared to
    the specialised one, and the product verifier (AirTable::verify) on that proof:
    one verification, and a batch on host threads."""
    import numpy as np
    import whirlaway_amd as wa
    from whirlaway_amd.air import download_witness
    from whirlaway_amd.program import poseidon2_air_program
    from whirlaway_amd.proof import verify_batch
    n = args.log_n
    settings = wa.AirSettings()
    host_w = download_witness(wit)
    gtable = wa.air.ProgramAirTable(ctx, poseidon2_air_program(rc), n, 3)
    gp = gtable.prove(settings, p16, p24, host_w)  # warm-up
    steps = 5
    t = time.perf_counter()
    for _ in range(steps):
        gp = gtable.prove(settings, p16, p24, host_w)
    tg = (time.perf_counter() - t) / steps
    gtable.close()
    del host_w
    r = verify_batch(table, settings, p16, p24, [proof])
    t = time.perf_counter()
    for _ in range(steps):
        r = verify_batch(table, settings, p16, p24, [proof], n_threads=1)
    tv = (time.perf_counter() - t) / steps
    batch = [proof] * 64
    t = time.perf_counter()
    rb = verify_batch(table, settings, p16, p24, batch, n_threads=16)
    tb = time.perf_counter() - t
    return {"generic_air": {"ms_per_proof": round(tg * 1e3, 3), "value": round((1 << n) / tg, 1),
                            "unit": "perms/s", "witness": "host memory",
                            "identical_to_specialised": bool(gp.size == proof.size and (gp == proof).all())},
            "verifier": {"ms_per_proof": round(tv * 1e3, 3), "accepted": r[0] == 0,
                         "batch": {"proofs": len(batch), "threads": 16, "proofs_per_s": round(len(batch) / tb, 1),
                                   "all_accepted": all(x == 0 for x in rb)}}}


def run_micro(args):
    """Kernel micro configs at full size: (a) 2^24 base-field values folded 24 times
    by extension challenges (benches/folding_benchmarks.rs shape), (b) the eq-weighted
    a^2 b product sumcheck over n = 24 variables.  Device time with HIP events;
    algorithmic bytes: (a) read 2^24 F + write/read each EF table once, (b) rounds 0
    and 1 read the two base-field tables, round 1 writes the two folded EF tables of
    2^23 entries, each later round reads the two tables of the round before and writes
    them folded (the eq weights come from tables of at most 2^14 entries)."""
    import numpy as np
    import whirlaway_amd as wa
    from whirlaway_amd import sumcheck, utils
    from whirlaway_amd.air import bench_inputs
    P = wa.P
    rng = np.random.default_rng(args.seed + 24)
    ctx = wa.Context(0)
    m = rng.integers(0, P, 1 << 24, dtype=np.uint32)
    ch = rng.integers(0, P, (24, 8), dtype=np.uint32)
    utils.fold_chain(ctx, m, ch)
    fold_ms = min(utils.fold_chain(ctx, m, ch)[2] for _ in range(3))
    ef_tables = sum(1 << (23 - i) for i in range(24))
    fold_bytes = 4 * (1 << 24) + 32 * ef_tables + 32 * (ef_tables - (1 << 23))
    a = rng.integers(0, P, 1 << 24, dtype=np.uint32)
    b = rng.integers(0, P, 1 << 24, dtype=np.uint32)
    z = rng.integers(0, P, (24, 8), dtype=np.uint32)
    _, _, p16, _ = bench_inputs(args.seed, 4, with_inputs=False)
    sumcheck.prove_a2b(ctx, p16, a, b, z)
    sc_ms = min(sumcheck.prove_a2b(ctx, p16, a, b, z)["ms"] for _ in range(2))
    sc_bytes = 2 * (2 * 4 << 24) + 2 * 32 * (1 << 23) + sum(2 * 32 * ((1 << (25 - j)) + (1 << (24 - j)))
                                                             for j in range(2, 24))
    # F-mul of the product sumcheck (an EF x EF product is 64, EF x F 8): round 0 on the
    # base-field tables, per pair and z in {0, 2, 3}: a_z^2 b_z (2) and the eq weight (8);
    # later rounds, per pair: the four folds by the previous challenge (EF x F in round 1,
    # EF x EF after) and for each z a_z^2, a_z^2 b_z and the eq weight (3 EF x EF)
    sc_fmul = ((1 << 23) * 3 * 10 + (1 << 22) * (4 * 8 + 3 * 3 * 64)
               + sum((1 << (23 - k)) * (4 * 64 + 3 * 3 * 64) for k in range(2, 24)))
    ctx.close()
    return {"fold_2^24_x24": {"ms": round(fold_ms, 4), "GB_per_s": round(fold_bytes / fold_ms / 1e6, 1),
                               "hbm_frac": round(fold_bytes / fold_ms / 1e6 / HBM_PEAK_GBS, 4)},
            "sumcheck_a2b_n24": {"ms": round(sc_ms, 4), "GB_per_s": round(sc_bytes / sc_ms / 1e6, 1),
                                 "hbm_frac": round(sc_bytes / sc_ms / 1e6 / HBM_PEAK_GBS, 4),
                                 "bound": "valu", "TFmul_per_s": round(sc_fmul / sc_ms / 1e9, 3),
                                 "valu_frac": round(sc_fmul / sc_ms / 1e9 / VALU_PEAK_TFMUL, 4)}}


def _requests(seed, count, lo, hi):
    """count proof requests with log_n uniform in {lo..hi} and inputs drawn per request
    (src/examples/poseidon2.rs:104-151 with seed = request seed)."""
    import numpy as np
    from whirlaway_amd.air import bench_inputs
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, count)
    return [(int(n), bench_inputs(10_000 + seed * 1000 + i, int(n))) for i, n in enumerate(sizes)]


def _prove_alone(ctx, n, job):
    import whirlaway_amd as wa
    from whirlaway_amd.air import generate_witness
    rc, inputs, p16, p24 = job
    table = wa.AirTable(ctx, rc, n)
    w = generate_witness(table, inputs)
    proof = table.prove(wa.AirSettings(), p16, p24, w)
    w.close()
    table.close()
    return proof


def run_batched(args):
    """64 independent proofs of uneven sizes (log_n ~ U{10..16}): sequential plain
    single-proof path on one context vs. the engine with several slots.  Both include
    table setup, input upload, GPU trace generation and the proof."""
    import numpy as np
    import whirlaway_amd as wa
    from whirlaway_amd.engine import Engine
    reqs = _requests(args.seed, args.batch_jobs, 10, 16)
    ctx = wa.Context(0)
    for n, job in reqs:  # warm-up: pools and twiddle tables of every size
        _prove_alone(ctx, n, job)
    t0 = time.perf_counter()
    seq = [_prove_alone(ctx, n, job) for n, job in reqs]
    t_seq = time.perf_counter() - t0
    ctx.close()
    with Engine(0, n_slots=args.slots) as eng:
        for rnd in range(2):  # round 0 warms the slots' pools
            t0 = time.perf_counter()
            tickets = [eng.submit(n, *job) for n, job in reqs]
            got = [eng.wait(t)[0] for t in tickets]
            t_eng = time.perf_counter() - t0
        st = eng.stats()
    same = all(np.array_equal(a, b) for a, b in zip(seq, got))
    return {"jobs": len(reqs), "log_n": "U{10..16}", "slots": args.slots,
            "proofs_per_s": round(len(reqs) / t_eng, 1),
            "plain_sequential_proofs_per_s": round(len(reqs) / t_seq, 1),
            "speedup": round(t_seq / t_eng, 3), "peak_concurrent": int(st["peak_running"]),
            "identical_to_plain": bool(same)}


def _replay(eng, reqs, arrivals):
    """Open-loop replay: submit request i at arrivals[i] seconds after the start."""
    t0 = time.perf_counter()
    tickets = []
    for (n, job), at in zip(reqs, arrivals):
        while True:
            dt = t0 + at - time.perf_counter()
            if dt <= 0:
                break
            time.sleep(min(dt, 0.0005) if dt > 0.001 else 0)
        tickets.append(eng.submit(n, *job))
    res = [eng.wait(t) for t in tickets]
    wall = time.perf_counter() - t0
    return res, wall


def run_serving(args):
    """Poisson arrivals of proof requests (log_n ~ U{10..18}) replayed open-loop against
    the engine (several slots, continuous admission) and against the plain path (one
    slot: the single-proof path behind a FIFO), at each arrival rate of the sweep; the
    last rate is above what the plain path sustains.  Latency = submit to proof ready."""
    import numpy as np
    from whirlaway_amd.engine import Engine
    reqs = _requests(args.seed + 1, args.serve_requests, 10, 18)
    rates = [float(r) for r in args.serve_rates.split(",")]
    points = []
    same = True
    for name, slots in (("plain", 1), ("engine", args.slots)):
        with Engine(0, n_slots=slots) as eng:
            # warm-up: every slot proves each size once (pools, twiddle tables)
            by_n = {}
            for n, job in reqs:
                by_n.setdefault(n, job)
            for _ in range(slots):
                ts = [eng.submit(n, *job) for n, job in sorted(by_n.items())]
                for t in ts:
                    eng.wait(t)
            for i, rate in enumerate(rates):
                rng = np.random.default_rng(args.seed + 2 + i)
                arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
                arrivals -= arrivals[0]
                res, wall = _replay(eng, reqs, arrivals)
                lat = np.array([r[1]["latency_ms"] for r in res])
                if len(points) <= i:
                    points.append({"rate_per_s": rate})
                points[i][name] = {"proofs_per_s": round(len(reqs) / wall, 1),
                                   "p50_ms": round(float(np.percentile(lat, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat, 99)), 3),
                                   "mean_queue_ms": round(float(np.mean([r[1]["queued_ms"] for r in res])), 3)}
                if name == "plain":
                    points[i]["_proofs"] = [r[0] for r in res]
                else:
                    same = same and all(np.array_equal(a, r[0]) for a, r in zip(points[i].pop("_proofs"), res))
    sat = points[-1]
    e = sat["engine"]
    return {"requests": len(reqs), "log_n": "U{10..18}", "slots": args.slots,
            "rate_per_s": sat["rate_per_s"], "proofs_per_s": e["proofs_per_s"], "p50_ms": e["p50_ms"],
            "p99_ms": e["p99_ms"], "mean_queue_ms": e["mean_queue_ms"], "plain": sat["plain"],
            "sweep": points, "identical_to_plain": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=17)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-log-n", type=int, default=17)
    ap.add_argument("--pcie-steps", type=int, default=5,
                    help="extra proofs timed from a host-memory witness (PCIe-inclusive rate)")
    ap.add_argument("--batch-jobs", type=int, default=64, help="batched leg: proofs (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=256, help="serving leg: requests (0 = skip)")
    ap.add_argument("--serve-rates", default="120,400,1200",
                    help="serving leg: Poisson arrival rates per second, comma-separated; the last one "
                         "saturates the plain path and gives the headline serving figures")
    ap.add_argument("--slots", type=int, default=8, help="engine slots (concurrent proofs)")
    ap.add_argument("--no-micro", action="store_true", help="skip the kernel micro leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the generic-AIR and product-verifier leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (multi-GPU sharding is a design "
              "described in DESIGN.md, not a measured path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import whirlaway_amd as wa
    from whirlaway_amd.air import bench_inputs, download_witness, generate_witness

    n = args.log_n
    if args.copy_probe_mib:
        # a plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # give the scale factor for the kernels' HBM traffic
        import torch
        a = torch.ones(args.copy_probe_mib << 18, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    ctx = wa.Context(0)
    rc, inputs, p16, p24 = bench_inputs(args.seed, n)
    table = wa.AirTable(ctx, rc, n)
    wit = generate_witness(table, inputs)
    settings = wa.AirSettings()
    proof = None
    for _ in range(args.warmup):
        proof = table.prove(settings, p16, p24, wit)
    # the autotuner times the commitment kernels' launch variants on the first calls
    # of each shape (one candidate per call); untimed proofs until it has decided
    tune_proofs = 0
    while wa.autotune.exploring() and tune_proofs < 32:
        proof = table.prove(settings, p16, p24, wit)
        tune_proofs += 1

    # Untimed pass with HIP events around every kernel class: per-class device time
    # and the dominant class.  Events are stream packets on the proof's critical path,
    # so the timed steps below record them around the dominant class only.
    ctx.kernel_timing(list(MODELS))
    proof = table.prove(settings, p16, p24, wit)
    per_class = {k: ctx.kernel_timing_read(k) for k in MODELS}
    ctx.kernel_timing([])
    dom = max(per_class, key=lambda k: per_class[k][1])

    ctx.kernel_timing([dom])
    ctx.phase_times(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proof = table.prove(settings, p16, p24, wit)
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = (1 << n) / (ms / 1e3)
    dom_live = ctx.kernel_timing_read(dom)
    ctx.kernel_timing([])
    phases = ctx.phase_times(reset=True)
    _progress(f"{args.steps} timed proofs: {ms:.3f} ms each")
    # launch variants the autotuner chose for this proof's commitment shapes (during warm-up)
    tuned = {"proofs_after_warmup": tune_proofs,
             "choices": [{"kernel": e["kind"], "log_rows": e["log_size"], "width": e["width"],
                          "variant": e["variant"],
                          "best_us": None if e["best_us"] is None else round(e["best_us"], 2)}
                         for e in wa.autotune.entries()]}

    pcie = None
    if args.pcie_steps > 0:
        # same proof with the witness handed over in host memory (wl_air_prove): adds the
        # host-to-device copy of the 165 x 2^n trace to every proof
        host_w = download_witness(wit)
        proof_h = table.prove(settings, p16, p24, host_w)  # allocates the pinned staging once
        tp = time.perf_counter()
        for _ in range(args.pcie_steps):
            proof_h = table.prove(settings, p16, p24, host_w)
        tp = (time.perf_counter() - tp) / args.pcie_steps
        pcie = {"value": round((1 << n) / tp, 1), "ms_per_step": round(tp * 1e3, 4),
                "steps": args.pcie_steps, "same_proof": bool((proof_h == proof).all())}
        del host_w
        _progress(f"host-witness proofs: {tp * 1e3:.3f} ms each")

    name, bound, per_unit = MODELS[dom]
    launches, kms, work = dom_live
    avg_ms = kms / max(launches, 1)
    per_launch = work * per_unit / max(launches, 1)
    if bound == "hbm":
        achieved = per_launch / (avg_ms / 1e3) / 1e9
        peak, unit = HBM_PEAK_GBS, "GB/s"
    else:
        achieved = per_launch / (avg_ms / 1e3) / 1e12
        peak, unit = VALU_PEAK_TFMUL, "TFmul/s"
    traffic = None
    if n == 17 and dom in MEASURED_TRAFFIC_N17:
        rd, wr = MEASURED_TRAFFIC_N17[dom]
        traffic = round((rd + wr) * 2**20)
    roofline = {
        "bound": bound, "achieved": round(achieved, 4), "peak": round(peak, 4), "unit": unit,
        "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": name,
        "fmul_probe_TFmul": round(ctx.fmul_rate() / 1e12, 4),
        "avg_launch_ms": round(avg_ms, 5), "launches_per_step": launches / args.steps,
        # from the untimed all-class pass (one proof)
        "kernel_ms_per_step": {MODELS[k][0].split(" ")[0]: round(v[1], 4) for k, v in per_class.items()},
    }
    if dom == 1:
        perms_per_s = work / max(launches, 1) / (avg_ms / 1e3)
        roofline["valu_issue"] = round(perms_per_s * VALU_INSTR_PER_P24 / 64 * 4 / (1024 * 2.4e9), 4)

    cpu = None
    identical = None  # GPU proof == oracle proof of the same workload (None: not compared)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py  # the checker, timed here only as the reported CPU baseline
        b = oracle_py.Bench(args.cpu_log_n, args.seed)
        tc = time.perf_counter()
        cpu_proof = b.prove()
        tcpu = time.perf_counter() - tc
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round((1 << args.cpu_log_n) / tcpu, 1), "unit": "perms/s", "cores": cores,
               "kind": "port",
               "sample": f"one oracle proof of 2^{args.cpu_log_n} permutations (same seed and "
                         f"settings), {tcpu:.2f} s wall, OpenMP"}
        if args.cpu_log_n == n:
            identical = bool(cpu_proof.size == proof.size and (cpu_proof == proof).all())
        b.close()
        _progress(f"CPU baseline: {tcpu:.2f} s")

    micro = None if args.no_micro else run_micro(args)
    _progress("micro done")
    extra = None if args.no_extra else run_extra(args, ctx, table, rc, p16, p24, wit, proof)
    _progress("generic AIR and verifier done")
    batched = run_batched(args) if args.batch_jobs > 0 else None
    _progress("batched done")
    serving = run_serving(args) if args.serve_requests > 0 else None
    _progress("serving done")

    line = {
        "metric": "Poseidon2 perms proven/s (2^17 perms, 1 GPU); batched proofs/s + p50/p99 latency",
        "value": round(value, 1),
        "unit": "perms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (KoalaBear F_p, Montgomery) / EF8",
        "data": "synthetic (StdRng seed 0 inputs, GPU-generated Poseidon2 trace)",
        "config": {"workload": f"Poseidon2-16 KoalaBear AIR, 2^{n} permutations, one proof per step",
                   "log_n_rows": n, "security_bits": 128, "soundness": "CapacityBound", "folding": [7, 4],
                   "log_inv_rate": 1, "univariate_skips": 4, "rs_domain_initial_reduction": 5,
                   "proof_words": int(proof.size)},
        "identical_to_oracle": identical,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phase_ms_per_step": dict(zip(PHASES, [round(x / args.steps, 4) for x in phases])),
        "autotune": tuned,
        "pcie_inclusive": pcie,
        "micro": micro,
        "generic_air_and_verifier": extra,
        "batched": batched,
        "serving": serving,
    }
    wit.close()
    table.close()
    ctx.close()
    print(json.dumps(line))
    mismatch = [k for k, v in (("oracle", identical), ("pcie", pcie and pcie["same_proof"]),
                               ("batched", batched and batched["identical_to_plain"]),
                               ("serving", serving and serving["identical_to_plain"]),
                               ("generic_air", extra and extra["generic_air"]["identical_to_specialised"]),
                               ("verifier", extra and extra["verifier"]["accepted"]
                                and extra["verifier"]["batch"]["all_accepted"])) if v is False]
    if mismatch:
        print(f"bench.py: proofs differ ({', '.join(mismatch)})", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
