"""Batched (C4) and served (C5) joins at their configured workloads, every
result checked against the oracle, not against another GPU path.

C4: 2048 PK-FK joins in one batch, |R_i| log-uniform in [2^10, 2^20],
|S_i| = 16 |R_i|, 16-byte tuples.  C5: Poisson arrivals of PK-FK joins with
|R| log-uniform in [2^12, 2^22] and |S| = U{4..64} |R|, 30% of them on 64
resident builds, admission under a 200 GB cap.

The device generator is byte-identical to the oracle's host generator
(test_gpu_parity.py::test_device_generator_matches_host), so each join's
expected (count, checksum) is the oracle's PK-FK closed form over its probe
side generated on the host with the same seed, one join at a time, so host
memory stays bounded.  Contract: matched_cnt = number of equal-key pairs,
check_sum = sum of r.row_id + s.row_id over them mod 2^64
(src/algo/phj.c:555-557)."""
import time

import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture()
def fresh():
    """A context of its own: the configured-size runs take the library's
    default radix rule, whatever size-class tuning earlier tests stored in
    the session context."""
    c = hjmi.Context(0)
    yield c
    c.close()


def _want(nr, ns, seed):
    _, hS = O.generate("pkfk", nr, ns, 16, seed=seed, threads=16)
    w = O.pkfk_closed_form(hS, 16)
    return w["matched_cnt"], w["check_sum"]


def test_c4_batch_at_configured_size_vs_oracle(fresh):
    ctx = fresh
    rng = np.random.default_rng(12345)
    sizes = [int(round(2 ** rng.uniform(10, 20))) for _ in range(2048)]
    Rs, Ss = [], []
    for i, nr in enumerate(sizes):
        R, S = ctx.generate("pkfk", nr, 16 * nr, 16, seed=70000 + i)
        Rs.append(R)
        Ss.append(S)
    print(f"C4: 2048 joins, {sum(17 * n for n in sizes)} tuples on the device", flush=True)
    got = {}
    for algo in ("phj_radix_bc", "phj_radix_lp"):
        b = ctx.batch(algo, Rs, Ss)
        res, info = b.run()
        again, _ = b.run()  # graph replay
        b.close()
        assert again == res, algo
        assert info["njoins"] == 2048 and info["n_fallback"] == 0, (algo, info)
        got[algo] = res
        print(f"C4 {algo}: {info['device_ms']:.1f} ms on the device", flush=True)
    for R, S in zip(Rs, Ss):
        ctx.free_relation(R)
        ctx.free_relation(S)
    bad = []
    for i, nr in enumerate(sizes):
        w = _want(nr, 16 * nr, 70000 + i)
        for algo, res in got.items():
            if res[i] != w:
                bad.append((algo, i, nr, res[i], w))
    assert not bad, bad[:5]


def test_c5_served_at_configured_size_vs_oracle(fresh):
    ctx = fresh
    rng = np.random.default_rng(777)

    def rel_sizes():
        nr = int(round(2 ** rng.uniform(12, 22)))
        return nr, nr * int(rng.integers(4, 65))

    pool = [(*rel_sizes(), 50000 + i) for i in range(64)]
    builds = [(*rel_sizes(), 90000 + i) for i in range(64)]
    dev_pool = [ctx.generate("pkfk", nr, ns, 16, seed=sd) for nr, ns, sd in pool]
    dev_builds = [ctx.generate("pkfk", nr, ns, 16, seed=sd) for nr, ns, sd in builds]
    tuples = sum(nr + ns for nr, ns, _ in pool + builds)
    print(f"C5: 64 + 64 resident relation pairs, {tuples} tuples on the device", flush=True)
    srv = hjmi.Server(ctx, "phj_radix_bc", hbm_cap_bytes=200 << 30, slots=3, max_batch=64, solo_lanes=2,
                      solo_min_log2=16)
    bids = [srv.register_build(R) for R, _ in dev_builds]
    # Poisson arrivals at 5000 joins/s offered, 30% on the resident builds
    n_req, rate = 1000, 5000.0
    gaps = rng.exponential(1.0 / rate, n_req)
    picks = [("b", int(rng.integers(64))) if rng.random() < 0.3 else ("p", int(rng.integers(64)))
             for _ in range(n_req)]
    ids = []
    t0, t = time.perf_counter(), 0.0
    for gap, (kind, i) in zip(gaps, picks):
        t += gap
        while time.perf_counter() - t0 < t:
            pass
        if kind == "b":
            ids.append(srv.submit(dev_builds[i][1], build=bids[i]))
        else:
            ids.append(srv.submit(dev_pool[i][1], R=dev_pool[i][0]))
    srv.drain()
    results = [srv.poll(rid) for rid in ids]
    st = srv.stats()
    srv.close()
    print(f"C5: {n_req} requests in {time.perf_counter() - t0:.2f} s, {st['batches']} batches, "
          f"{st['solo_completed']} in solo lanes", flush=True)
    assert st["completed"] == n_req and st["rejected"] == 0
    for R, S in dev_pool + dev_builds:
        ctx.free_relation(R)
        ctx.free_relation(S)
    want_p = [_want(*p) for p in pool]
    want_b = [_want(*b) for b in builds]
    bad = [(k, i, res) for (res, _), (k, i) in zip(results, picks)
           if res != (want_b[i] if k == "b" else want_p[i])]
    assert not bad, bad[:5]
