"""GPU tests of serving (hj_serve_*): continuous batching over several
streams; every request's result equals the same join run alone, whether it
ran alone or in a mix with others; resident builds; admission rejection."""
import time

import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu


def _pool(ctx, n, lo, hi, seed, wl="pkfk"):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nr = int(round(2 ** rng.uniform(lo, hi)))
        ns = nr * int(rng.integers(4, 65))
        out.append(ctx.generate(wl, nr, ns, 16, seed=seed * 1000 + i, key_range=max(1, nr // 2)))
    return out


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp"])
def test_serve_results_equal_alone_and_in_mix(ctx, algo):
    pool = _pool(ctx, 24, 6, 17, 7)
    builds = _pool(ctx, 4, 10, 16, 8)
    alone = [ctx.join(algo, R, S) for R, S in pool] + [ctx.join(algo, R, S) for R, S in builds]
    want = [(a["matched_cnt"], a["check_sum"]) for a in alone]
    srv = hjmi.Server(ctx, algo, hbm_cap_bytes=6 << 30, slots=3, max_batch=8)
    bids = [srv.register_build(R) for R, _ in builds]
    rng = np.random.default_rng(1)
    reqs = []
    # a burst of arrivals so batches form and overlap, then a trickle
    for k in range(120):
        i = int(rng.integers(0, len(pool) + len(builds)))
        if i < len(pool):
            rid = srv.submit(pool[i][1], R=pool[i][0])
        else:
            b = i - len(pool)
            rid = srv.submit(builds[b][1], build=bids[b])
        reqs.append((rid, i))
        if k >= 100:
            time.sleep(0.002)
    srv.drain()
    for rid, i in reqs:
        res, lat = srv.poll(rid)
        assert res == want[i], (rid, i)
        assert lat >= 0
    st = srv.stats()
    assert st["completed"] == 120 and st["rejected"] == 0
    assert st["batches"] >= 2 and st["max_batch_joins"] >= 2
    assert st["n_builds"] == 4 and st["resident_bytes"] > 0
    # the same request alone, on an idle server
    rid = srv.submit(pool[3][1], R=pool[3][0])
    srv.drain()
    assert srv.poll(rid)[0] == want[3]
    srv.close()
    for R, S in pool + builds:
        ctx.free_relation(R)
        ctx.free_relation(S)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp"])
def test_serve_solo_lanes_in_mix(ctx, algo):
    """lp servers send requests with |R| >= 2^14 to solo lanes (full
    two-pass engine on their own streams) and batch the small ones; bc
    servers batch every request up to |R| = 2^22 (gathered batches) and send
    only larger build sides to the lanes.  Every result equals the join run
    alone."""
    pool = _pool(ctx, 10, 8, 13, 21) + _pool(ctx, 6, 14, 18, 22)
    builds = _pool(ctx, 2, 12, 16, 23)
    big = ctx.generate("pkfk", (1 << 22) + 1, 1 << 24, 16, seed=24)  # above the gathered limit
    alone = [ctx.join(algo, R, S) for R, S in pool] + [ctx.join(algo, R, S) for R, S in builds]
    want = [(a["matched_cnt"], a["check_sum"]) for a in alone]
    a_big = ctx.join(algo, *big)
    srv = hjmi.Server(ctx, algo, hbm_cap_bytes=8 << 30, slots=2, max_batch=8, solo_lanes=2, solo_min_log2=14)
    bids = [srv.register_build(R) for R, _ in builds]
    rng = np.random.default_rng(3)
    reqs = []
    for k in range(60):
        i = int(rng.integers(0, len(pool) + len(builds)))
        if i < len(pool):
            rid = srv.submit(pool[i][1], R=pool[i][0])
        else:
            rid = srv.submit(builds[i - len(pool)][1], build=bids[i - len(pool)])
        reqs.append((rid, i))
        if k == 30:
            rid_big = srv.submit(big[1], R=big[0])
    srv.drain()
    for rid, i in reqs:
        assert srv.poll(rid)[0] == want[i], (rid, i)
    assert srv.poll(rid_big)[0] == (a_big["matched_cnt"], a_big["check_sum"])
    st = srv.stats()
    assert st["completed"] == 61 and st["solo_completed"] > 0 and st["batches"] > 0
    if algo == "phj_radix_bc":
        assert st["solo_completed"] == 1  # only the build side above 2^22
    srv.close()
    for R, S in pool + builds + [big]:
        ctx.free_relation(R)
        ctx.free_relation(S)


def test_serve_lp_duplicates_exact(ctx):
    hR, hS = O.generate("random", 5000, 40000, 16, seed=5, key_range=1500)
    want = O.aggregate(hR, hS, 16, lp=True)
    R, S = ctx.upload(hR, 16), ctx.upload(hS, 16)
    srv = hjmi.Server(ctx, "phj_radix_lp", hbm_cap_bytes=1 << 30, slots=2, max_batch=4)
    ids = [srv.submit(S, R=R) for _ in range(5)]
    srv.drain()
    for rid in ids:
        assert srv.poll(rid)[0] == (want["matched_cnt"], want["check_sum"])
    srv.close()
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_serve_lp_scatter_flag_is_exact(ctx):
    """An lp server told not to gather (HJ_SERVE_SCATTER) joins every
    request through the exact first-match path."""
    pool = _pool(ctx, 6, 6, 14, 31, wl="random")
    want = [(a["matched_cnt"], a["check_sum"]) for a in (ctx.join("phj_radix_lp", R, S) for R, S in pool)]
    srv = hjmi.Server(ctx, "phj_radix_lp", hbm_cap_bytes=2 << 30, slots=2, max_batch=4, flags=hjmi.HJ_SERVE_SCATTER)
    ids = [(srv.submit(S, R=R), i) for i, (R, S) in enumerate(pool) for _ in range(2)]
    srv.drain()
    for rid, i in ids:
        assert srv.poll(rid)[0] == want[i], (rid, i)
    srv.close()
    for R, S in pool:
        ctx.free_relation(R)
        ctx.free_relation(S)


def test_serve_admission_rejects_oversized(ctx):
    srv = hjmi.Server(ctx, "phj_radix_bc", hbm_cap_bytes=64 << 20, slots=2, max_batch=4)
    R, S = ctx.generate("pkfk", 1 << 16, 1 << 21, 16)  # 32 MiB probe side > one 32 MiB arena
    with pytest.raises(hjmi.HJError) as e:
        srv.submit(S, R=R)
    assert e.value.code == hjmi.HJ_EREJECT
    r2, s2 = ctx.generate("pkfk", 1000, 16000, 16)
    rid = srv.submit(s2, R=r2)
    srv.drain()
    got = srv.poll(rid)[0]
    alone = ctx.join("phj_radix_bc", r2, s2)
    assert got == (alone["matched_cnt"], alone["check_sum"])
    assert srv.stats()["rejected"] == 1
    srv.close()
    for x in (R, S, r2, s2):
        ctx.free_relation(x)


def test_serve_rejects_bad_config(ctx):
    with pytest.raises(hjmi.HJError):
        hjmi.Server(ctx, "nphj_sc")
    with pytest.raises(hjmi.HJError):
        hjmi.Server(ctx, "phj_radix_bc", slots=0)
    with pytest.raises(hjmi.HJError):
        hjmi.Server(ctx, "phj_radix_bc", solo_lanes=5)
