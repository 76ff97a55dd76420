"""GPU tests of batched joins (hj_batch_*): every join's result in the batch
equals the same join run alone through hj_join, and the oracle, bit for bit;
graph replays see updated inputs; LP batches take the exact path per join."""
import math

import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu


def _sizes(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [int(round(2 ** rng.uniform(lo, hi))) for _ in range(n)]


def _make(ctx, n, lo, hi, ratio=16, wl="pkfk", width=16, seed=1):
    Rs, Ss = [], []
    for i, nr in enumerate(_sizes(n, lo, hi, seed)):
        R, S = ctx.generate(wl, nr, nr * ratio, width, seed=1000 + i, key_range=max(1, nr // 2))
        Rs.append(R)
        Ss.append(S)
    return Rs, Ss


def _free(ctx, *lists):
    for lst in lists:
        for r in lst:
            ctx.free_relation(r)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp"])
def test_batch_equals_individual_joins(ctx, algo):
    Rs, Ss = _make(ctx, 60, 4, 16)
    b = ctx.batch(algo, Rs, Ss)
    got, info = b.run()
    assert info["n_unpartitioned"] > 0 and info["n_partitioned"] > 0
    for i, (R, S) in enumerate(zip(Rs, Ss)):
        alone = ctx.join(algo, R, S)
        assert got[i] == (alone["matched_cnt"], alone["check_sum"]), i
        assert got[i][0] == S.tuple_num  # PK-FK: every probe tuple matches once
    again, _ = b.run()  # graph replay
    assert again == got
    b.close()
    _free(ctx, Rs, Ss)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp"])
def test_batch_duplicates_misses_vs_oracle(ctx, algo):
    """Random keys (duplicates and misses): LP joins with duplicate build
    keys are recomputed exactly per join."""
    sizes = [3, 100, 3000, 9000, 40000]
    Rs, Ss, want = [], [], []
    for i, nr in enumerate(sizes):
        hR, hS = O.generate("random", nr, nr * 5, 16, seed=50 + i, key_range=max(1, nr // 3))
        want.append(O.aggregate(hR, hS, 16, lp=algo.endswith("lp")))
        Rs.append(ctx.upload(hR, 16))
        Ss.append(ctx.upload(hS, 16))
    b = ctx.batch(algo, Rs, Ss)
    got, info = b.run()
    for i in range(len(sizes)):
        assert got[i] == (want[i]["matched_cnt"], want[i]["check_sum"]), (i, sizes[i])
    if algo.endswith("lp"):
        assert info["n_fallback"] >= 1
    b.close()
    _free(ctx, Rs, Ss)


@pytest.mark.parametrize("width,flags", [(64, 0), (16, hjmi.HJ_F_BATCH_SCATTER)])
def test_lp_batch_outside_the_gathered_path_is_exact(ctx, width, flags):
    """lp batches that do not run gathered (tuples wider than 16 bytes, or
    HJ_F_BATCH_SCATTER) join every pair through the exact first-match path:
    results equal the joins run alone and the oracle's."""
    sizes = [5, 2000, 30000]
    Rs, Ss, want = [], [], []
    for i, nr in enumerate(sizes):
        hR, hS = O.generate("random", nr, nr * 4, width, seed=70 + i, key_range=max(1, nr // 2))
        want.append(O.aggregate(hR, hS, width, lp=True))
        Rs.append(ctx.upload(hR, width))
        Ss.append(ctx.upload(hS, width))
    b = ctx.batch("phj_radix_lp", Rs, Ss, flags=flags)
    got, info = b.run()
    assert info["gathered"] == 0 and info["n_fallback"] == len(sizes)
    for i in range(len(sizes)):
        assert got[i] == (want[i]["matched_cnt"], want[i]["check_sum"]), i
        alone = ctx.join("phj_radix_lp", Rs[i], Ss[i])
        assert got[i] == (alone["matched_cnt"], alone["check_sum"]), i
    again, _ = b.run()
    assert again == got
    b.close()
    _free(ctx, Rs, Ss)


def test_batch_sees_new_inputs_on_replay(ctx):
    Rs, Ss = _make(ctx, 8, 8, 14, seed=3)
    b = ctx.batch("phj_radix_bc", Rs, Ss)
    first, _ = b.run()
    # overwrite join 5's probe side in place with another seeded relation
    nr, ns = Rs[5].tuple_num, Ss[5].tuple_num
    R2, S2 = ctx.generate("random", nr, ns, 16, seed=99, key_range=nr)
    ctx.copy_d2d(Ss[5].ptr, S2.ptr, S2.nbytes)
    second, _ = b.run()
    want = ctx.join("phj_radix_bc", Rs[5], Ss[5])
    assert second[5] == (want["matched_cnt"], want["check_sum"])
    assert second[5] != first[5]
    assert second[:5] == first[:5] and second[6:] == first[6:]
    b.close()
    _free(ctx, Rs, Ss, [R2, S2])


def test_batch_empty_and_wide(ctx):
    empty = hjmi.DeviceRelation(0, 0, 64)
    Rs, Ss = _make(ctx, 5, 6, 13, width=64, wl="zipf", seed=4)
    Rs.insert(2, empty)
    Ss.insert(2, Ss[0])
    b = ctx.batch("phj_radix_bc", Rs, Ss)
    got, _ = b.run()
    assert got[2] == (0, 0)
    for i in (0, 1, 3, 4, 5):
        alone = ctx.join("phj_radix_bc", Rs[i], Ss[i])
        assert got[i] == (alone["matched_cnt"], alone["check_sum"])
    b.close()
    _free(ctx, [r for r in Rs if r.ptr], [s for k, s in enumerate(Ss) if k != 2])


def test_batch_switches_between_narrow_and_wide_partitions(ctx):
    """The partition layout is decided on the device at every replay: 8-byte
    tuples while every key and row id fits 32 bits, 16-byte ones once any
    join of the batch holds a wider key.  Results stay those of the joins
    run alone, and of the oracle for the join made wide."""
    Rs, Ss = _make(ctx, 12, 8, 15, seed=6)
    b = ctx.batch("phj_radix_bc", Rs, Ss)
    first, info = b.run()
    assert info["n_partitioned"] > 0
    for i, (R, S) in enumerate(zip(Rs, Ss)):
        alone = ctx.join("phj_radix_bc", R, S)
        assert first[i] == (alone["matched_cnt"], alone["check_sum"]), i
    # join j gets keys above 2^32 (same tuple counts), written in place
    j = max(range(12), key=lambda i: Rs[i].tuple_num)
    nr, ns = Rs[j].tuple_num, Ss[j].tuple_num
    hR, hS = O.generate("random", nr, ns, 16, seed=77, key_range=max(1, nr // 2))
    for h in (hR, hS):
        h.view(np.uint64).reshape(-1, 2)[:, 0] += np.uint64(1 << 33)
    want = O.aggregate(hR, hS, 16)
    dR, dS = ctx.upload(hR, 16), ctx.upload(hS, 16)
    ctx.copy_d2d(Rs[j].ptr, dR.ptr, dR.nbytes)
    ctx.copy_d2d(Ss[j].ptr, dS.ptr, dS.nbytes)
    second, _ = b.run()
    assert second[j] == (want["matched_cnt"], want["check_sum"])
    assert second[:j] == first[:j] and second[j + 1:] == first[j + 1:]
    b.close()
    _free(ctx, Rs, Ss, [dR, dS])


def test_batch_rejects_bad_requests(ctx):
    R, S = ctx.generate("pkfk", 100, 1000, 16)
    with pytest.raises(hjmi.HJError):
        ctx.batch("nphj_sc", [R], [S])
    W, Z = ctx.generate("pkfk", 100, 1000, 64)
    with pytest.raises(hjmi.HJError):
        ctx.batch("phj_radix_bc", [R, W], [S, Z])
    _free(ctx, [R, S, W, Z])


def test_autotune_classes_and_parity(ctx):
    ctx.autotune(16, 16, 10, 16)
    for c in range(10, 17):
        bits, passes = ctx.tuned_config(1 << c)
        assert bits <= c and passes in (0, 1, 2)
        # gathered batches: the 6144-tuple rule or one bit either side
        rule = 0 if (1 << c) <= 6144 else ((1 << c) + 6143) // 6144 - 1
        rule = rule.bit_length()
        assert abs(ctx.tuned_batch_bits(1 << c) - rule) <= 1, c
        b1 = ctx.tuned_pass1_bits(1 << c)
        assert b1 == 0 or (passes == 2 and 1 <= b1 < bits)
    R, S = ctx.generate("pkfk", 1 << 16, 1 << 20, 16)
    hR, hS = O.generate("pkfk", 1 << 16, 1 << 20, 16)
    want = O.pkfk_closed_form(hS, 16)
    got = ctx.join("phj_radix_bc", R, S)
    assert (got["matched_cnt"], got["check_sum"]) == (want["matched_cnt"], want["check_sum"])
    _free(ctx, [R, S])
