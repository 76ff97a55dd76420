#!/usr/bin/env python3
"""Benchmark: one equi-join per step on one MI355X, inputs resident in HBM.

Default workload (the headline configuration): phj_radix_bc on
|R| = 2^24, |S| = 2^28 tuples of 16 bytes {u64 key, u64 payload}, R keys a
seeded permutation of 1..2^24 (payload = key), S keys uniform over R's keys
(payload = 1-based row index).  One step = hj_join() = partition + join +
result read-back.  `value` is (|R|+|S|) tuples per second over the timed
steps (host wall clock, synchronised on both sides).

Also reported:
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average HIP-event duration in the timed steps, vs 8 TB/s;
                `traffic` is its measured HBM bytes per launch from the
                committed rocprofv3 --pmc summary, when one exists;
  cpu_baseline  the CPU oracle's restatement of the reference's DRAM-mode
                join (same algorithm, 11 radix bits, one pass, and two passes)
                timed on this host's cores on the full workload, 6 trials,
                the first dropped.

Prints exactly one JSON line on stdout; everything else goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hashjoin-scm_amd"))

METRIC = "(|R|+|S|) tuples/s + % HBM peak, 16M⋈256M; batched joins/s & p99 vs plain"
PEAK_GBPS = 8000.0
CONFIGS = {
    "C1": {"algo": "nphj_sc", "workload": "pkfk", "nr": 1 << 20, "ns": 1 << 24, "width": 16, "theta": 1.0},
    "C2": {"algo": "phj_radix_bc", "workload": "pkfk", "nr": 1 << 24, "ns": 1 << 28, "width": 16, "theta": 1.0},
    "C3": {"algo": "phj_radix_bc", "workload": "zipf", "nr": 1 << 24, "ns": 1 << 27, "width": 64, "theta": 1.0},
    # the reference's payload sweep (logs/payload-logs): row ids are positions
    "PAY16": {"algo": "ptr_phj_radix_bc", "workload": "payload_ref", "nr": 1 << 24, "ns": 1 << 28, "width": 16,
              "theta": 1.0},
    "PAY64": {"algo": "ptr_phj_radix_bc", "workload": "payload_ref", "nr": 1 << 24, "ns": 1 << 28, "width": 64,
              "theta": 1.0},
    "PAY512": {"algo": "ptr_phj_radix_bc", "workload": "payload_ref", "nr": 1 << 24, "ns": 1 << 28, "width": 512,
               "theta": 1.0},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, W, nr, ns, res=None):
    """Algorithmic HBM bytes one launch of `name` must move (no histogram
    passes, table zeroing or partial lines): a radix pass reads and writes
    every tuple once, the radix join reads every tuple's head once, an NPHJ
    build reads R and writes a 16-byte entry per tuple, a probe reads S and
    one 16-byte entry per probe.  A head-only read of a tuple wider than 64
    bytes counts one 64-byte HBM burst.  Radix joins of tuples wider than 16
    bytes partition the 16-byte heads: two passes read them in place in pass
    A, one pass partitions the heads that ptr_keyrid copies out.  Two passes over 16-byte tuples store
    a pass-A tile as 8-byte {key, row_id} tuples when they fit 32 bits
    (res["narrow_tuples"] tuples), of those a tile whose row ids are its
    positions + 1 as 4-byte words (res["positional_tuples"]), and pass B and
    the join then move 8-byte tuples when every tile was narrow
    (res["narrow"])."""
    N = nr + ns
    head = min(W, 64)
    # radix joins of tuples wider than 16 bytes partition their 16-byte heads
    # (copied out by ptr_keyrid), so the radix kernels move 16-byte tuples
    Wr = 16
    a_out = b_out = Wr * N  # pass A output, pass B output
    join_in = Wr * N
    if res is not None and res.get("passes") == 2:
        nt, nd = res.get("narrow_tuples", 0), res.get("positional_tuples", 0)
        a_out = 4 * nd + 8 * (nt - nd) + 16 * (N - nt)
        b_out = join_in = (8 if res.get("narrow") else 16) * N
    # two passes over wider tuples read the heads in place in pass A (one
    # burst per tuple); fewer passes copy them out first (ptr_keyrid)
    a_in = head * N if (W > 16 and res is not None and res.get("passes") == 2) else Wr * N
    return {
        "radix_scatter": 2 * Wr * N,
        "radix_sortchunk": 2 * Wr * N,
        "radix_tilesort": a_in + a_out,
        "radix_gathersort": a_out + b_out,
        "radix_join": join_in,
        # nphj_lp's narrow table (res["narrow"]): 8-byte entries
        "lp_build": (head + (8 if res is not None and res.get("narrow") else 16)) * nr,
        "sc_build": (head + 16) * nr,
        "lp_probe": (head + (8 if res is not None and res.get("narrow") else 16)) * ns,
        "sc_probe": (head + 16) * ns,
        # heads of wider tuples (radix and late materialisation): one burst
        # read and a 16-byte write per tuple; per pair the pair read and two
        # 8-byte row-id gathers
        "ptr_keyrid": (head + 16) * N if W > 16 else None,
        "ptr_retrieve": 32 * res["matched_cnt"] if res else None,
    }.get(name)


def load_traffic(kernel, workload_tag):
    p = os.path.join(ROOT, "profiles", f"{kernel}_{workload_tag}_traffic.json")
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("hbm_bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(cfg, algo):
    """The oracle's restatement of the reference's DRAM-mode join, timed on
    this host's cores as the reference times itself (src/main.c:429-459):
    the full workload, every core this process may run on (`nproc`; the GPU
    box gives a GPU's process 16), 6 trials with the first dropped, mean and
    min of total_time.  Radix joins run the shipped configuration (11 radix
    bits, one pass) and are also reported with two passes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    nr, ns = cfg["nr"], cfg["ns"]
    R, S = O.generate(cfg["workload"], nr, ns, cfg["width"], theta=cfg["theta"], threads=threads)
    base_algo = algo.replace("_bw_reg", "")

    def trials(passes):
        runs = []
        for _ in range(6):
            if base_algo.startswith("ptr_"):
                r = O.ptr_join(base_algo, R, S, cfg["width"], threads=threads, radix_bits=11, passes=passes)
            else:
                r = O.join(base_algo, R, S, cfg["width"], threads=threads, radix_bits=11, passes=passes)
            runs.append(r)
        kept = runs[1:]
        mean = sum(r["total_time"] for r in kept) / len(kept)
        return {"mean_s": mean, "min_s": min(r["total_time"] for r in kept),
                "buildpart_s": sum(r["buildpart_time"] for r in kept) / len(kept),
                "probejoin_s": sum(r["probejoin_time"] for r in kept) / len(kept),
                "tuples_per_s": (nr + ns) / mean, "matched_cnt": kept[0]["matched_cnt"],
                "check_sum": kept[0]["check_sum"]}

    radix = "radix" in base_algo
    one = trials(1)
    two = trials(2) if radix else None
    import platform
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    conf = "11 radix bits, 1 pass" if radix else "global table"
    out = {"value": one["tuples_per_s"], "unit": "tuples/s", "cores": threads, "kind": "port",
           "sample": f"{base_algo} restated (oracle/hjoracle.c), full workload {nr}x{ns} {cfg['width']}B "
                     f"{cfg['workload']}, {conf}, mean of trials 2-6 of 6, {threads} threads on "
                     f"{model or platform.machine()}",
           "one_pass" if radix else "trials": one}
    if two:
        out["two_pass"] = two
    return out


def tpch_pushdown_bytes(q, np_, nl, res):
    """Algorithmic HBM bytes of the two selection kernels of one query: every
    column a predicate tests, read for every row; the key and money columns
    read for the selected rows only; one 16-byte tuple written per output
    row.  Q14: p_partkey + p_type (33 B/part), l_shipdate (10 B/lineitem),
    l_partkey + l_extendedprice + l_discount (20 B/selected lineitem).  Q19:
    p_brand + p_size + p_container (24 B/part), p_partkey (8 B/selected
    part), l_quantity + l_shipmode + l_shipinstruct (39 B/lineitem), 20
    B/selected lineitem."""
    if q == 14:
        rd = 33 * np_ + 10 * nl + 20 * res["lineitem_selected"]
    else:
        rd = 24 * np_ + 8 * res["part_selected"] + 39 * nl + 20 * res["lineitem_selected"]
    return rd + 16 * (res["part_tuples"] + res["lineitem_tuples"])


def tpch_cpu_baseline(q, ctx, T):
    """The oracle's vectorised restatement of the query (numpy, one core) on
    the SF1 cardinalities, generated on the device and copied to the host; 3
    runs, first dropped."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tpch_oracle as TO
    np_, nl = T.cardinality(1)
    t = T.DeviceTables.generate(ctx, np_, nl, seed=12345)
    part, line = t.download()
    t.free()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        TO.query(q, part, line)
        times.append(time.perf_counter() - t0)
    el = sum(times[1:]) / 2
    return {"value": (np_ + nl) / el, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"Q{q} restated in numpy (oracle/tpch_oracle.py), SF1 cardinalities {np_} parts x {nl} "
                      f"lineitems, mean of runs 2-3 of 3, 1 thread"}


def run_tpch(args, hjmi):
    """TPC-H Q14 / Q19 end to end (src/tpch/tpch_query.c): one step = the
    selections, the joins on partkey and the sums, over columns resident in
    HBM at scale factor --sf.  `value` is (part + lineitem rows) per second."""
    import hjmi.tpch as T
    q = int(args.config[1:])
    algo = args.algo or "tpch_rdx_bc"
    ctx = hjmi.Context(0)
    info = ctx.device_info()
    np_, nl = T.cardinality(args.sf)
    t = T.DeviceTables.generate(ctx, np_, nl, seed=12345)
    log(f"TPC-H SF{args.sf}: {np_} parts, {nl} lineitems, {t.nbytes / 1e9:.1f} GB of columns in HBM")
    for _ in range(args.warmup):
        T.query(ctx, q, algo, t)
    ctx.sync()
    t0 = time.perf_counter()
    res = [T.query(ctx, q, algo, t) for _ in range(args.steps)]
    ctx.sync()
    elapsed = time.perf_counter() - t0
    keys = ("matched_cnt", "numerator_e4", "denominator_e4", "revenue_e4")
    if any(tuple(r[k] for k in keys) != tuple(res[0][k] for k in keys) for r in res):
        raise SystemExit("results differ between steps")
    r0 = res[0]
    push_ms = sum(r["pushdown_time"] for r in res) / len(res) * 1e3
    pbytes = tpch_pushdown_bytes(q, np_, nl, r0)
    achieved = pbytes / (push_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_GBPS, 4), "traffic": load_traffic("tpch_select", f"Q{q}_SF{args.sf}"),
                "kernel": "tpch_select (part + lineitem selections)", "kernel_avg_ms": round(push_ms, 4),
                "kernel_algo_bytes": pbytes,
                "phase_ms": {k: round(sum(r[k] for r in res) / len(res) * 1e3, 4)
                             for k in ("pushdown_time", "buildpart_time", "probejoin_time", "total_time")}}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = tpch_cpu_baseline(q, ctx, T)
        except Exception as e:  # reported, never fatal for the GPU number
            log(f"cpu baseline failed: {e}")
    answer = ({"promo_revenue": r0["promo_revenue"], "numerator": r0["numerator"], "denominator": r0["denominator"]}
              if q == 14 else {"revenue": r0["revenue"]})
    out = {
        "metric": f"TPC-H Q{q} rows/s (part + lineitem)", "value": (np_ + nl) * args.steps / elapsed,
        "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (TPC-H value domains, seeded on device)",
        "config": {"workload": f"Q{q} SF{args.sf}: {algo}, {np_} parts x {nl} lineitems",
                   "matched_cnt": r0["matched_cnt"], **answer, "part_selected": r0["part_selected"],
                   "lineitem_selected": r0["lineitem_selected"], "joins": r0["joins"],
                   "device": info["arch"], "cu": info["cu_count"]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    t.free()
    ctx.close()


def run_batched(args, hjmi):
    """C4: njoins independent PK-FK joins, |R_i| log-uniform in [2^lo, 2^hi],
    |S_i| = 16 |R_i|, 16-byte tuples, seeded per join.  Times one batched
    graph launch against the same joins as back-to-back plain hj_join calls;
    every per-join result must be identical between the two."""
    import random
    ctx = hjmi.Context(0)
    rng = random.Random(12345)
    sizes = [int(round(2 ** rng.uniform(args.batch_lo, args.batch_hi))) for _ in range(args.joins)]
    Rs, Ss = [], []
    for i, nr in enumerate(sizes):
        R, S = ctx.generate("pkfk", nr, 16 * nr, 16, seed=12345 + i)
        Rs.append(R)
        Ss.append(S)
    tuples = sum(17 * n for n in sizes)
    t0 = time.perf_counter()
    if not args.no_autotune:  # radix bits per size class for the plain joins
        ctx.autotune(16, 16, args.batch_lo, args.batch_hi)
    tune_s = time.perf_counter() - t0
    algo = args.algo or "phj_radix_bc"
    # plain: one synchronous hj_join per join
    plain_res = [ctx.join(algo, R, S) for R, S in zip(Rs, Ss)]  # warm-up
    plain_times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        plain_res = [ctx.join(algo, R, S) for R, S in zip(Rs, Ss)]
        plain_times.append(time.perf_counter() - t0)
    b = ctx.batch(algo, Rs, Ss, flags=args.batch_flags)
    out, info = b.run()  # warm-up
    batch_times, dev_ms = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out, info = b.run()
        batch_times.append(time.perf_counter() - t0)
        dev_ms.append(info["device_ms"])
    same = all(o == (p["matched_cnt"], p["check_sum"]) for o, p in zip(out, plain_res))
    if not same:
        raise SystemExit("batched results differ from plain joins")
    # per-kernel device times: the same batch launched without its graph,
    # each kernel between two events (HJ_F_KERNEL_TIMES), after the timed runs
    kernel_ms = {}
    if info.get("gathered"):
        bt = ctx.batch(algo, Rs, Ss, flags=args.batch_flags | hjmi.HJ_F_KERNEL_TIMES)
        bt.run()
        kout, kinfo = bt.run()
        bt.close()
        if kout != out:
            raise SystemExit("batch with kernel times differs from the graph's results")
        kernel_ms = {k: round(v, 3) for k, v in kinfo["kernel_ms"].items()}
        log("batch kernels (ms):", kernel_ms)
    tb, tp = min(batch_times), min(plain_times)
    # Roofline of the batch graph (one launch): the library's algorithmic
    # bytes of the run -- gathered batches: pass A reads 16 B per tuple and
    # writes each tile at its layout (4-byte positional words, 8-byte or
    # 16-byte tuples), the join reads those; scattered batches: the scatter
    # reads 16 and writes 8 (narrow) and the join reads 8, unpartitioned
    # joins read 16 -- over the graph's device time.
    algo_bytes = info["algo_bytes"]
    dev_s = min(dev_ms) * 1e-3
    achieved = algo_bytes / dev_s / 1e9
    res = {
        "metric": "batched joins/s (one hipGraph launch) vs back-to-back plain joins", "value": args.joins / tb,
        "unit": "joins/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": tb * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded on device)",
        "config": {"workload": f"C4: {args.joins} x {algo}, |R| log-uniform [2^{args.batch_lo}, 2^{args.batch_hi}], "
                               f"|S| = 16|R|, 16B pkfk", "tuples": tuples},
        "plain_joins_per_s": args.joins / tp, "speedup_vs_plain": tp / tb,
        "batch_tuples_per_s": tuples / tb, "plain_tuples_per_s": tuples / tp,
        "batch_device_ms": min(dev_ms), "identical_results": same, "autotune_s": tune_s,
        "batch_info": {k: v for k, v in info.items() if k != "kernel_ms"}, "batch_kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBPS, 4), "traffic": None, "kernel": "batch graph",
                     "algo_bytes": algo_bytes, "device_ms": round(dev_s * 1e3, 3)},
    }
    b.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def run_served(args, hjmi):
    """C5: Poisson arrivals of PK-FK joins, |R| log-uniform in [2^12, 2^22],
    |S| = U{4..64} |R|, 16-byte tuples; 30% of requests reuse one of 64
    resident build relations.  The continuous-batching server (admission by
    HBM footprint under --cap-gb) is compared with the unbatched plain path
    (one synchronous hj_join per request, first come first served) on the
    same arrival sequence.  Every request's result must equal its join run
    alone."""
    import random
    ctx = hjmi.Context(0)
    rng = random.Random(777)
    algo = args.algo or "phj_radix_bc"

    def rel_sizes():
        nr = int(round(2 ** rng.uniform(12, 22)))
        return nr, nr * rng.randint(4, 64)

    pool = []
    for i in range(args.pool):
        nr, ns = rel_sizes()
        pool.append(ctx.generate("pkfk", nr, ns, 16, seed=50000 + i))
    builds = []
    for i in range(64):
        nr, ns = rel_sizes()
        builds.append(ctx.generate("pkfk", nr, ns, 16, seed=90000 + i))
    if not args.no_autotune:
        ctx.autotune(16, 16, 12, 22)
    want = [ctx.join(algo, R, S) for R, S in pool] + [ctx.join(algo, R, S) for R, S in builds]
    want = [(w["matched_cnt"], w["check_sum"]) for w in want]
    srv = hjmi.Server(ctx, algo, hbm_cap_bytes=int(args.cap_gb * (1 << 30)), slots=args.slots,
                      max_batch=args.max_batch, solo_lanes=args.solo_lanes, solo_min_log2=args.solo_min_log2,
                      flags=args.serve_flags)
    # 16-byte bc servers batch gathered: every |R| <= 2^20 (2^22 without
    # lanes) in continuous batches, larger build sides in the solo lanes
    gathered = algo.endswith("bc") and not (args.serve_flags & hjmi.HJ_SERVE_SCATTER)
    if gathered:
        lanes = (f"{args.solo_lanes} solo lanes taking |R| > 2^20 while idle" if args.solo_lanes else "no solo lanes") + \
            ", gathered batches"
    else:
        lanes = f"{args.solo_lanes} solo lanes for |R| >= 2^{args.solo_min_log2}, scattered batches"
    bids = [srv.register_build(R) for R, _ in builds]
    rates = [int(x) for x in args.rates.split(",")]
    out_rates = []
    for rate in rates:
        arr, t = [], 0.0
        picks = []
        for _ in range(args.requests):
            t += rng.expovariate(rate)
            arr.append(t)
            picks.append(("b", rng.randrange(64)) if rng.random() < 0.3 else ("p", rng.randrange(len(pool))))
        # served
        ids = []
        t0 = time.perf_counter()
        for a, (kind, i) in zip(arr, picks):
            while time.perf_counter() - t0 < a:
                pass
            if kind == "b":
                ids.append(srv.submit(builds[i][1], build=bids[i]))
            else:
                ids.append(srv.submit(pool[i][1], R=pool[i][0]))
        srv.drain()
        t_end = time.perf_counter()
        lat, ok = [], True
        for rid, (kind, i) in zip(ids, picks):
            res, l = srv.poll(rid)
            lat.append(l)
            ok &= res == want[len(pool) + i if kind == "b" else i]
        lat.sort()
        served = {"joins_per_s": args.requests / (t_end - t0), "p50_ms": lat[len(lat) // 2],
                  "p99_ms": lat[int(len(lat) * 0.99) - 1], "identical_results": ok}
        # plain: one synchronous join at a time, in arrival order
        plat = []
        t0 = time.perf_counter()
        for a, (kind, i) in zip(arr, picks):
            while time.perf_counter() - t0 < a:
                pass
            R, S = (builds[i] if kind == "b" else pool[i])
            ctx.join(algo, R, S)
            plat.append((time.perf_counter() - t0 - a) * 1e3)
        t_end = time.perf_counter()
        plat.sort()
        plain = {"joins_per_s": args.requests / (t_end - t0), "p50_ms": plat[len(plat) // 2],
                 "p99_ms": plat[int(len(plat) * 0.99) - 1]}
        out_rates.append({"offered_per_s": rate, "served": served, "plain": plain})
        log(f"rate {rate}: served {served} plain {plain}")
    st = srv.stats()
    srv.close()
    best = max(out_rates, key=lambda r: r["served"]["joins_per_s"])
    res = {
        "metric": "served joins/s and p50/p99 latency (continuous batching) vs unbatched plain joins",
        "value": best["served"]["joins_per_s"], "unit": "joins/s", "n_gpus": 1, "steps": args.requests,
        "warmup": 0, "ms_per_step": 1e3 / best["served"]["joins_per_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded on device)",
        "config": {"workload": f"C5: Poisson arrivals, {algo}, |R| log-uniform [2^12, 2^22], |S| = U{{4..64}}|R|, "
                               f"30% reuse of 64 resident builds, cap {args.cap_gb} GB, {args.slots} slots, {lanes}",
                   "requests_per_rate": args.requests},
        "rates": out_rates, "server": st,
        "all_identical": all(r["served"]["identical_results"] for r in out_rates),
    }
    for R, S in pool + builds:
        ctx.free_relation(R)
        ctx.free_relation(S)
    ctx.close()
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C2", choices=sorted(CONFIGS) + ["C4", "C5", "Q14", "Q19"])
    ap.add_argument("--sf", type=int, default=100, choices=[1, 10, 100], help="Q14/Q19: TPC-H scale factor")
    ap.add_argument("--rates", default="1000,2000,5000,10000,20000", help="C5: offered joins/s")
    ap.add_argument("--requests", type=int, default=1000, help="C5: requests per rate")
    ap.add_argument("--pool", type=int, default=64, help="C5: distinct non-resident joins")
    ap.add_argument("--cap-gb", type=float, default=200.0, help="C5: admission cap (HBM GB)")
    ap.add_argument("--slots", type=int, default=3, help="C5: concurrent batches")
    ap.add_argument("--max-batch", type=int, default=64, help="C5: joins per launch")
    ap.add_argument("--solo-lanes", type=int, default=0,
                    help="C5: lanes running large joins alone (gathered servers: |R| > 2^20)")
    ap.add_argument("--serve-flags", type=int, default=0, help="C5: hj_serve_config_t flags (1 = HJ_SERVE_SCATTER)")
    ap.add_argument("--solo-min-log2", type=int, default=16, help="C5: |R| >= 2^x runs in a solo lane")
    ap.add_argument("--joins", type=int, default=2048, help="C4: joins per batch")
    ap.add_argument("--batch-lo", type=int, default=10, help="C4: log2 of the smallest |R|")
    ap.add_argument("--batch-hi", type=int, default=20, help="C4: log2 of the largest |R|")
    ap.add_argument("--batch-flags", type=int, default=0,
                    help="C4: hj_params_t flags of the batch (8 = HJ_F_BATCH_SCATTER: histogram + scatter path)")
    ap.add_argument("--algo", default=None)
    ap.add_argument("--theta", type=float, default=None, help="C3: Zipf exponent of the probe keys (default 1.0)")
    ap.add_argument("--part-variant", type=lambda x: int(x, 0), default=0,
                    help="hj_params_t.part_variant of the timed joins (tuning sweeps)")
    ap.add_argument("--skew-tuned", action="store_true",
                    help="tune the workload's size class on Zipf probe keys (hj_autotune_zipf) and join with "
                         "HJ_F_SKEW_HANDLING")
    ap.add_argument("--bits", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="C4/C5: skip the size-class autotuner (plain joins then use the default radix bits)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="instead of a join, time a plain device copy of MIB mebibytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py: this build measures one GPU; multi-GPU sharding is a design note "
            "(see the sharding paragraph of the design document), so --gpus must be 1")
        sys.exit(2)
    import hjmi

    if args.copy_calibration:
        ctx = hjmi.Context(0)
        nb = args.copy_calibration << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        for _ in range(args.warmup):
            ctx.copy_d2d(b, a, nb)
        ms = [ctx.copy_d2d(b, a, nb) for _ in range(args.steps)]
        avg = sum(ms) / len(ms)
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "avg_ms": avg,
                          "rw_GBps": 2 * nb / (avg * 1e-3) / 1e9, "launches": len(ms)}), flush=True)
        ctx.close()
        return

    if args.config == "C4":
        run_batched(args, hjmi)
        return
    if args.config == "C5":
        run_served(args, hjmi)
        return
    if args.config in ("Q14", "Q19"):
        run_tpch(args, hjmi)
        return

    cfg = dict(CONFIGS[args.config])
    if args.theta is not None:
        cfg["theta"] = args.theta
    algo = args.algo or cfg["algo"]
    nr, ns, W = cfg["nr"], cfg["ns"], cfg["width"]
    ctx = hjmi.Context(0)
    info = ctx.device_info()
    log(f"device {info}")
    R, S = ctx.generate(cfg["workload"], nr, ns, W, theta=cfg["theta"])
    flags = 0
    if args.skew_tuned:  # tune this class on Zipf probe keys, join with HJ_F_SKEW_HANDLING
        c = (nr - 1).bit_length()
        ctx.autotune_zipf(W, max(1, ns // nr), c, c, theta=cfg["theta"])
        flags = hjmi.HJ_F_SKEW_HANDLING
        log(f"skew-tuned configuration for |R| = {nr}: {ctx.tuned_skew_config(nr)}")
    for _ in range(args.warmup):
        ref = ctx.join(algo, R, S, radix_bits=args.bits, passes=args.passes, part_variant=args.part_variant,
                       flags=flags)
    ctx.sync()
    t0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        res.append(ctx.join(algo, R, S, radix_bits=args.bits, passes=args.passes, part_variant=args.part_variant,
                            kernel_times=True, flags=flags))
    ctx.sync()
    elapsed = time.perf_counter() - t0
    for r in res:
        if (r["matched_cnt"], r["check_sum"]) != (res[0]["matched_cnt"], res[0]["check_sum"]):
            raise SystemExit("results differ between steps")
    value = (nr + ns) * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3

    # per-kernel device time over the timed steps
    tot, cnt = {}, {}
    for r in res:
        for k, lst in r["kernel_ms"].items():
            tot[k] = tot.get(k, 0.0) + sum(lst)
            cnt[k] = cnt.get(k, 0) + len(lst)
    dom = max((k for k in tot if kernel_bytes(k, W, nr, ns, res[0])), key=lambda k: tot[k])
    avg_ms = tot[dom] / cnt[dom]
    achieved = kernel_bytes(dom, W, nr, ns, res[0]) / (avg_ms * 1e-3) / 1e9
    job_s = sum(r["total_time"] for r in res) / len(res)
    job_gbps = res[0]["algo_bytes"] / job_s / 1e9
    tag = f"{args.config}_{algo}"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_GBPS, 4), "traffic": load_traffic(dom, tag),
                "kernel": dom, "kernel_avg_ms": round(avg_ms, 4),
                "kernel_algo_bytes": kernel_bytes(dom, W, nr, ns, res[0]),
                "narrow_tuples": res[0].get("narrow_tuples", 0), "narrow": res[0].get("narrow", 0),
                "positional_tuples": res[0].get("positional_tuples", 0),
                "job_algo_bytes": res[0]["algo_bytes"], "job_device_ms": round(job_s * 1e3, 4),
                "job_frac": round(job_gbps / PEAK_GBPS, 4),
                "kernel_ms_per_step": {k: round(v / args.steps, 4) for k, v in sorted(tot.items())}}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(cfg, algo)
        except Exception as e:  # reported, never fatal for the GPU number
            log(f"cpu baseline failed: {e}")
    out = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded on device)",
        "config": {"workload": f"{args.config}: {algo} |R|={nr} |S|={ns} {W}B {cfg['workload']}"
                               + (f" theta={cfg['theta']}" if cfg["workload"] == "zipf" else ""),
                   "radix_bits": res[0]["radix_bits"], "passes": res[0]["passes"],
                   "matched_cnt": res[0]["matched_cnt"], "check_sum": res[0]["check_sum"],
                   "device": info["arch"], "cu": info["cu_count"]},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
