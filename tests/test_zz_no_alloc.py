"""hj_join_noalloc (nphj_sc / nphj_lp with the library's fixed per-GPU
scratch) through the C ABI: its results equal hj_join's and the oracle's on
two calls in a row, the second with a smaller input, and sizes beyond the
fixed region are rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu


def _pair(r):
    return r["matched_cnt"], r["check_sum"]


@pytest.mark.parametrize("algo", ["nphj_lp", "nphj_sc"])
@pytest.mark.parametrize("width", [16, 64])
def test_noalloc_equals_join_two_calls(ctx, algo, width):
    # the largest build side each algorithm's region takes, then a smaller one
    big = (1 << 21) if algo == "nphj_lp" else (1 << 19)
    sizes = [(big, 4 * big), (3001, 50021)]
    for i, (nr, ns) in enumerate(sizes):
        print(f"noalloc {algo} W={width} call {i}: |R|={nr} |S|={ns}", flush=True)
        R, S = ctx.generate("pkfk", nr, ns, width, seed=71 + i)
        got = ctx.join_noalloc(algo, R, S)
        ref = ctx.join(algo, R, S)
        assert _pair(got) == _pair(ref), (algo, width, i, got, ref)
        want = O.pkfk_closed_form(ctx.download(S), width)
        assert _pair(got) == (want["matched_cnt"], want["check_sum"]), (algo, width, i)
        assert got["fallback"] == 0
        ctx.free_relation(R)
        ctx.free_relation(S)


@pytest.mark.parametrize("algo", ["nphj_lp", "nphj_sc"])
def test_noalloc_duplicates_and_misses(ctx, algo):
    """Duplicate build keys (nphj_lp's fix-up indices come from the fixed
    region too) and probe misses, then a smaller PK-FK call."""
    hR, hS = O.generate("random", 1 << 16, 1 << 19, 16, seed=73, key_range=40000)
    want = O.aggregate(hR, hS, 16, lp=algo.endswith("lp"))
    dR, dS = ctx.upload(hR, 16), ctx.upload(hS, 16)
    got = ctx.join_noalloc(algo, dR, dS)
    assert _pair(got) == (want["matched_cnt"], want["check_sum"])
    assert _pair(got) == _pair(ctx.join(algo, dR, dS))
    if algo == "nphj_lp":
        assert got["fallback"] & hjmi.HJ_FB_LP_DUPLICATES
    ctx.free_relation(dR)
    ctx.free_relation(dS)
    R, S = ctx.generate("pkfk", 1000, 9000, 16, seed=74)
    want = O.pkfk_closed_form(ctx.download(S), 16)
    assert _pair(ctx.join_noalloc(algo, R, S)) == (want["matched_cnt"], want["check_sum"])
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_noalloc_rejects_what_does_not_fit(ctx):
    lib = hjmi.load_library()
    small_r, small_s = ctx.generate("pkfk", 256, 1024, 16, seed=75)
    # build sides beyond the fixed region: only the descriptor is read
    for algo, nr in (("nphj_lp", (1 << 21) + 1), ("nphj_sc", (1 << 19) + 1)):
        R = hjmi.Relation(C.c_void_p(small_r.ptr), nr, 16, hjmi.HJ_LOC_DEVICE)
        S = small_s.as_c()
        res = hjmi.Result()
        rc = lib.hj_join_noalloc(ctx.h, algo.encode(), C.byref(R), C.byref(S), None, C.byref(res))
        assert rc == hjmi.HJ_EINVAL, (algo, rc)
        assert b"scratch bytes" in lib.hj_last_error(ctx.h)
    # host-resident inputs, radix algorithms and unknown flags are refused
    host = np.zeros(16 * 4, dtype=np.uint8)
    Rh = hjmi.Relation(host.ctypes.data_as(C.c_void_p), 4, 16, hjmi.HJ_LOC_HOST)
    res = hjmi.Result()
    assert lib.hj_join_noalloc(ctx.h, b"nphj_lp", C.byref(Rh), C.byref(small_s.as_c()), None,
                               C.byref(res)) == hjmi.HJ_EINVAL
    assert lib.hj_join_noalloc(ctx.h, b"phj_radix_bc", C.byref(small_r.as_c()), C.byref(small_s.as_c()), None,
                               C.byref(res)) == hjmi.HJ_EINVAL
    prm = hjmi.Params(0, 0, 0, hjmi.HJ_F_NO_NARROW, 0, 0)
    assert lib.hj_join_noalloc(ctx.h, b"nphj_sc", C.byref(small_r.as_c()), C.byref(small_s.as_c()), C.byref(prm),
                               C.byref(res)) == hjmi.HJ_EINVAL
    # the context still joins afterwards
    want = O.pkfk_closed_form(ctx.download(small_s), 16)
    assert _pair(ctx.join_noalloc("nphj_sc", small_r, small_s)) == (want["matched_cnt"], want["check_sum"])
    ctx.free_relation(small_r)
    ctx.free_relation(small_s)
