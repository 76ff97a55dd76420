"""GPU parity: every join through the C ABI (libhjmi.so) against the CPU
oracle on the same seeded inputs, bit-exact count and 64-bit checksum.

Small and medium sizes are checked against the restated reference joins and
the sort-based aggregation; full benchmark and reference sizes are checked
against closed forms (the reference's logged golden vectors)."""
import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu

ALGOS = list(hjmi.ALGOS)


def _check(got, want, what=""):
    assert (got["matched_cnt"], got["check_sum"]) == (want["matched_cnt"], want["check_sum"]), (what, got, want)


def _dev(ctx, arr, width):
    return ctx.upload(arr, width)


# ----------------------------------------------------------------- generator
@pytest.mark.parametrize("wl,width,theta", [("pkfk", 16, 1.0), ("pkfk_ref", 16, 1.0), ("payload_ref", 32, 1.0),
                                            ("zipf", 64, 1.5), ("zipf", 16, 0.5), ("random", 128, 1.0)])
def test_device_generator_matches_host(ctx, wl, width, theta):
    nr, ns = 3001, 20011
    R, S = ctx.generate(wl, nr, ns, width, seed=4242, theta=theta, key_range=2000)
    hr, hs = O.generate(wl, nr, ns, width, seed=4242, theta=theta, key_range=2000)
    assert np.array_equal(ctx.download(R), hr)
    assert np.array_equal(ctx.download(S), hs)
    ctx.free_relation(R)
    ctx.free_relation(S)


# ------------------------------------------------------- small/medium parity
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("wl,nr,ns,width", [("pkfk", 1 << 12, 1 << 16, 16), ("pkfk_ref", 5000, 80000, 16),
                                           ("payload_ref", 1 << 14, 1 << 18, 32), ("zipf", 1 << 13, 1 << 17, 64)])
def test_join_matches_oracle(ctx, algo, wl, nr, ns, width):
    R, S = O.generate(wl, nr, ns, width, theta=1.0)
    want = O.join(algo, R, S, width, threads=4, radix_bits=6, passes=1)
    _check(want, O.aggregate(R, S, width, lp=algo.endswith("lp")), "oracle self-check")
    dR, dS = _dev(ctx, R, width), _dev(ctx, S, width)
    _check(ctx.join(algo, dR, dS), want, algo)
    ctx.free_relation(dR)
    ctx.free_relation(dS)


RADIX = ["phj_radix_bc", "phj_radix_lp", "phj_radix_sc", "phj_radix_hr", "phj_radix_hro"]


@pytest.mark.parametrize("algo", RADIX)
@pytest.mark.parametrize("bits,passes", [(0, 0), (1, 1), (4, 1), (8, 1), (11, 1), (12, 1), (2, 2), (9, 2),
                                         (12, 2), (14, 2), (16, 2), (20, 2)])
def test_radix_configurations(ctx, algo, bits, passes):
    nr, ns = 1 << 16, 1 << 20
    R, S = O.generate("pkfk", nr, ns, 16, seed=11)
    want = O.pkfk_closed_form(S, 16)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    got = ctx.join(algo, dR, dS, radix_bits=bits, passes=passes)
    _check(got, want, (bits, passes))
    # unique build keys: one pass (or none) joins oversized lp partitions
    # fill by fill; two passes re-run them exactly when a partition exceeds
    # the LDS table
    if (nr >> got["radix_bits"]) <= 4096 or got["passes"] <= 1:
        assert got["fallback"] == 0
    elif algo.endswith("lp"):
        assert got["fallback"] == hjmi.HJ_FB_LP_OVERSIZE
    ctx.free_relation(dR)
    ctx.free_relation(dS)


@pytest.mark.parametrize("bits,passes", [(0, 0), (1, 1), (2, 1)])
def test_lp_fills_and_cross_fill_duplicates(ctx, bits, passes):
    """lp partitions larger than one LDS table are joined fill by fill: with
    unique build keys the result is exact with no re-run; a key repeated in a
    later fill of the same partition (here the last R tuple repeats the key
    of an early one, same radix digit) is caught and the join recomputed, so
    the first build row of each key still wins."""
    nr, ns = 1 << 15, 1 << 18
    R, S = O.generate("pkfk", nr, ns, 16, seed=41)
    want = O.aggregate(R, S, 16, lp=True)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    got = ctx.join("phj_radix_lp", dR, dS, radix_bits=bits, passes=passes)
    _check(got, want, ("unique", bits))
    assert got["fallback"] == 0, got
    R2 = R.copy()
    w = R2.view(np.uint64)
    w[2 * (nr - 1)] = w[2 * 5]  # last R key := key of R[5]: same digit, a later fill
    want2 = O.aggregate(R2, S, 16, lp=True)
    dR2 = _dev(ctx, R2, 16)
    got2 = ctx.join("phj_radix_lp", dR2, dS, radix_bits=bits, passes=passes)
    _check(got2, want2, ("duplicate", bits))
    assert got2["fallback"] & hjmi.HJ_FB_LP_DUPLICATES, got2
    for d in (dR, dS, dR2):
        ctx.free_relation(d)


@pytest.mark.parametrize("bits", [10, 13])
def test_lp_narrow_two_pass_duplicates(ctx, bits):
    """Two passes over narrow tuples (keys and row ids below 2^32) join lp
    partitions with the compact 8-byte table: unique keys need no re-run;
    one repeated build key is caught when its partition's slots are claimed
    and the join is recomputed, so the first build row of each key wins."""
    nr, ns = 1 << 16, 1 << 20
    R, S = O.generate("pkfk", nr, ns, 16, seed=43)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    got = ctx.join("phj_radix_lp", dR, dS, radix_bits=bits, passes=2)
    _check(got, O.aggregate(R, S, 16, lp=True), ("unique", bits))
    assert got["narrow"] == 1 and got["fallback"] == 0, got
    R2 = R.copy()
    w = R2.view(np.uint64)
    w[2 * (nr - 1)] = w[2 * 5]  # the last R key := the key of R[5]
    dR2 = _dev(ctx, R2, 16)
    got2 = ctx.join("phj_radix_lp", dR2, dS, radix_bits=bits, passes=2)
    _check(got2, O.aggregate(R2, S, 16, lp=True), ("duplicate", bits))
    assert got2["fallback"] & hjmi.HJ_FB_LP_DUPLICATES, got2
    for d in (dR, dS, dR2):
        ctx.free_relation(d)


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("width", [16, 64])
def test_duplicates_and_misses(ctx, algo, width):
    """Keys uniform in [1, 3000] on both sides: duplicate build keys, misses.
    bc/sc count every pair; lp keeps the first build row of each key."""
    R, S = O.generate("random", 9000, 50000, width, key_range=3000, seed=5)
    want = O.aggregate(R, S, width, lp=algo.endswith("lp"))
    dR, dS = _dev(ctx, R, width), _dev(ctx, S, width)
    for bits, passes in ((0, 0), (3, 1), (6, 2)):
        got = ctx.join(algo, dR, dS, radix_bits=bits, passes=passes)
        _check(got, want, (algo, bits, passes))
        if algo.endswith("lp"):
            assert got["fallback"] & hjmi.HJ_FB_LP_DUPLICATES
    ctx.free_relation(dR)
    ctx.free_relation(dS)


def _tuples(keys, pays, width=16):
    a = np.zeros((len(keys), width // 8), dtype=np.uint64)
    a[:, 0] = np.array(keys, dtype=np.uint64)
    a[:, 1] = np.array(pays, dtype=np.uint64)
    return a.view(np.uint8).reshape(-1)


@pytest.mark.parametrize("algo", ALGOS)
def test_extreme_keys(ctx, algo):
    """Key 0, key 2^64-1 (the LP empty marker), 2^63; checksums that wrap."""
    big = (1 << 64) - 1
    rng = np.random.default_rng(1)
    rk = [0, big, 5, big, 1 << 63] + list(rng.integers(1, 1 << 62, 500, dtype=np.uint64))
    rp = [1, 2, 3, 4, big] + list(rng.integers(0, 1 << 63, 500, dtype=np.uint64))
    sk = [big, 0, 7, 1 << 63, big] + list(rng.choice(np.array(rk, dtype=np.uint64), 4000))
    sp = [10, 20, 30, 40, big] + list(rng.integers(0, 1 << 63, 4000, dtype=np.uint64))
    R, S = _tuples(rk, rp), _tuples(sk, sp)
    want = O.bruteforce(R, S, 16, lp=algo.endswith("lp"))
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    for bits, passes in ((0, 0), (2, 1), (7, 2)):
        _check(ctx.join(algo, dR, dS, radix_bits=bits, passes=passes), want, (algo, bits))
    ctx.free_relation(dR)
    ctx.free_relation(dS)


@pytest.mark.parametrize("case", ["narrow", "key_ffffffff", "pay_2_32", "probe_wide_keys"])
def test_lp_narrow_table_boundaries(ctx, case):
    """nphj_lp keeps an 8-byte {key, payload} table while every build key is
    below 2^32 - 1 and every payload fits 32 bits, and rebuilds the 16-byte
    table when one tuple does not: a single such tuple among many, keys at
    the boundary, duplicates (the first build row wins), probe keys above
    32 bits against a narrow table."""
    rng = np.random.default_rng(17)
    rk = list(rng.integers(1, 40000, 30000, dtype=np.uint64)) + [0xFFFFFFFE, 0xFFFFFFFE]
    rp = list(rng.integers(0, 1 << 32, 30000, dtype=np.uint64)) + [7, 8]
    if case == "key_ffffffff":
        rk[123] = 0xFFFFFFFF
    elif case == "pay_2_32":
        rp[456] = 1 << 32
    sk = list(rng.choice(np.array(rk, dtype=np.uint64), 60000)) + [0xFFFFFFFE, 0xFFFFFFFF]
    if case == "probe_wide_keys":
        sk += [(1 << 32) + int(k) for k in rk[:500]] + [(1 << 64) - 1]
    sp = list(rng.integers(0, 1 << 63, len(sk), dtype=np.uint64))
    R, S = _tuples(rk, rp), _tuples(sk, sp)
    want = O.aggregate(R, S, 16, lp=True)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    got = ctx.join("nphj_lp", dR, dS)
    _check(got, want, case)
    assert got["fallback"] & hjmi.HJ_FB_LP_DUPLICATES
    ctx.free_relation(dR)
    ctx.free_relation(dS)


@pytest.mark.parametrize("algo", ALGOS)
def test_empty_and_tiny_relations(ctx, algo):
    one = _tuples([9], [1])
    many = _tuples([9, 8, 9, 1], [2, 3, 4, 5])
    empty = np.zeros(0, dtype=np.uint8)
    for a, b in ((empty, many), (many, empty), (empty, empty), (one, many), (many, one)):
        want = O.bruteforce(a, b, 16, lp=algo.endswith("lp"))
        dA, dB = _dev(ctx, a, 16), _dev(ctx, b, 16)
        for bits, passes in ((0, 0), (1, 1), (4, 2)):
            _check(ctx.join(algo, dA, dB, radix_bits=bits, passes=passes), want, (algo, len(a), len(b)))
        ctx.free_relation(dA)
        ctx.free_relation(dB)


@pytest.mark.parametrize("algo", ["nphj_sc", "nphj_lp"] + RADIX)
def test_one_hot_build_key(ctx, algo):
    """Every build tuple has the same key: one bucket chain of |R| entries
    (SC overflow pool grows and rebuilds; radix bc walks an LDS chain over
    several build chunks; lp falls back to its exact path)."""
    nr, ns = 20000, 3000
    R = _tuples([77] * nr, list(range(nr)))
    S = _tuples([77, 78] * (ns // 2), list(range(ns)))
    want = O.aggregate(R, S, 16, lp=algo.endswith("lp"))
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    _check(ctx.join(algo, dR, dS), want, algo)
    ctx.free_relation(dR)
    ctx.free_relation(dS)


def test_host_resident_relations(ctx):
    R, S = O.generate("pkfk", 1 << 12, 1 << 16, 16)
    want = O.pkfk_closed_form(S, 16)
    for algo in ALGOS:
        _check(ctx.join_host(algo, R, S, 16), want, algo)


@pytest.mark.parametrize("algo", ["phj_radix_sc", "phj_radix_hr", "phj_radix_hro"])
@pytest.mark.parametrize("width", [16, 32, 64])
def test_partition_tables_on_collisions(ctx, algo, width):
    """Keys sharing their low 16 bits fill one partition and a few buckets:
    sc overflow chains, hr/hro ranges over many tuples, several LDS fills."""
    rng = np.random.default_rng(width)
    base = rng.integers(1, 1 << 40, 300, dtype=np.uint64) << np.uint64(16)
    rk = np.concatenate([base, base[:100], rng.integers(1, 1 << 50, 5000, dtype=np.uint64)])
    sk = np.concatenate([rng.choice(rk, 60000), rng.integers(1, 1 << 50, 5000, dtype=np.uint64)])
    R = np.zeros((len(rk), width // 8), dtype=np.uint64)
    S = np.zeros((len(sk), width // 8), dtype=np.uint64)
    R[:, 0], R[:, 1] = rk, rng.integers(0, 1 << 63, len(rk), dtype=np.uint64)
    S[:, 0], S[:, 1] = sk, rng.integers(0, 1 << 63, len(sk), dtype=np.uint64)
    R, S = R.view(np.uint8).reshape(-1), S.view(np.uint8).reshape(-1)
    want = O.aggregate(R, S, width)
    dR, dS = _dev(ctx, R, width), _dev(ctx, S, width)
    for bits, passes in ((0, 0), (4, 1), (12, 1), (8, 2), (16, 2)):
        _check(ctx.join(algo, dR, dS, radix_bits=bits, passes=passes), want, (algo, bits, passes))
    ctx.free_relation(dR)
    ctx.free_relation(dS)


def test_hro_rejects_wide_tuples(ctx):
    R, S = ctx.generate("pkfk", 64, 256, 128)
    with pytest.raises(hjmi.HJError) as e:
        ctx.join("phj_radix_hro", R, S, radix_bits=2, passes=1)
    assert e.value.code == hjmi.HJ_EINVAL
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_errors_are_status_codes(ctx):
    R, S = ctx.generate("pkfk", 64, 256, 16)
    with pytest.raises(hjmi.HJError) as e:
        ctx.join("phj_radix_zz", R, S)
    assert e.value.code == hjmi.HJ_EINVAL
    with pytest.raises(hjmi.HJError) as e:
        ctx.join("phj_radix_bc", R, S, radix_bits=21)
    assert e.value.code == hjmi.HJ_EINVAL
    bad = hjmi.DeviceRelation(S.ptr, S.tuple_num, 24)
    with pytest.raises(hjmi.HJError):
        ctx.join("nphj_sc", R, bad)
    # the context still works after errors
    assert ctx.join("nphj_sc", R, S)["matched_cnt"] == 256
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_repeated_joins_reuse_scratch(ctx):
    R, S = ctx.generate("pkfk", 1 << 15, 1 << 19, 16)
    first = ctx.join("phj_radix_bc", R, S)
    for algo in ALGOS * 2:
        got = ctx.join(algo, R, S)
        _check(got, first, algo)
    ctx.free_relation(R)
    ctx.free_relation(S)


@pytest.mark.parametrize("passes", [1, 2])
def test_kernel_times_cover_the_join(ctx, passes):
    """kernel_times=True: the same result, every phase named, and the
    phases' device times sum to no more than the join's own (the two-pass
    phases are chained: one event between two, the last closing with the
    join)."""
    R, S = ctx.generate("pkfk", 1 << 20, 1 << 24, 16, seed=31)
    bits = 13 if passes == 2 else 10
    want = ctx.join("phj_radix_bc", R, S, radix_bits=bits, passes=passes)
    for _ in range(2):
        got = ctx.join("phj_radix_bc", R, S, radix_bits=bits, passes=passes, kernel_times=True)
        assert got["passes"] == passes
        _check(got, want, "phj_radix_bc")
        names = set(got["kernel_ms"])
        need = {"radix_tilesort", "radix_scan", "radix_plan", "radix_gathersort", "radix_join"} if passes == 2 \
            else {"radix_plan", "radix_join"}
        assert need <= names, names
        total = sum(sum(v) for v in got["kernel_ms"].values())
        assert all(x >= 0 for v in got["kernel_ms"].values() for x in v)
        assert total <= got["total_time"] * 1e3 * 1.0001 + 1e-3, (total, got["total_time"])
        if passes == 2:
            # the chained regions span the join, from pass A to the join's end
            assert total >= 0.9 * got["total_time"] * 1e3, (total, got["total_time"])
    ctx.free_relation(R)
    ctx.free_relation(S)


ASYMM = ["phj_asymm_radix_bc", "phj_asymm_radix_sc", "phj_asymm_radix_hr", "phj_asymm_radix_hro"]


@pytest.mark.parametrize("algo", ASYMM)
def test_asymmetric_first_pass_bits(ctx, algo):
    """phj_asymm_radix_*: NUM_RADIX_BITS_PASS1 first-pass bits (default 11),
    the rest in the second pass; a first pass of all the bits is one pass.
    Every split gives the oracle's result on duplicates and misses."""
    R, S = O.generate("random", 30000, 200000, 16, key_range=20000, seed=9)
    want = O.aggregate(R, S, 16)
    _check(O.join(algo, R, S, 16, threads=4, radix_bits=14, radix_bits_pass1=9), want, "oracle split")
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    for b1 in (2, 5, 7, 9, 11, 12):
        got = ctx.join(algo, dR, dS, radix_bits=14, radix_bits_pass1=b1)
        _check(got, want, (algo, b1))
        assert got["passes"] == 2 and got["radix_bits"] == 14
    got = ctx.join(algo, dR, dS, radix_bits=13)  # default first pass: 11 bits
    _check(got, want, (algo, "default"))
    assert got["passes"] == 2
    got = ctx.join(algo, dR, dS, radix_bits=9)  # 11 >= 9: one pass of 9 bits
    _check(got, want, (algo, "one pass"))
    assert got["passes"] == 1
    ctx.free_relation(dR)
    ctx.free_relation(dS)


def test_first_pass_bits_on_plain_radix(ctx):
    R, S = O.generate("pkfk", 1 << 16, 1 << 20, 16, seed=3)
    want = O.pkfk_closed_form(S, 16)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    for b1 in (1, 3, 8, 12):
        _check(ctx.join("phj_radix_bc", dR, dS, radix_bits=13, passes=2, radix_bits_pass1=b1), want, b1)
    with pytest.raises(hjmi.HJError) as e:
        ctx.join("phj_radix_bc", dR, dS, radix_bits=13, passes=2, radix_bits_pass1=13)
    assert e.value.code == hjmi.HJ_EINVAL
    ctx.free_relation(dR)
    ctx.free_relation(dS)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp", "phj_radix_hr", "ptr_phj_radix_bc"])
def test_narrow_second_pass_matches_wide(ctx, algo):
    """Two passes over 16-byte tuples whose keys and row ids fit 32 bits
    write 8-byte tuples from pass B on; one wide value anywhere (here the
    last probe tuple's row id, or a build key) keeps 16-byte tuples.  Every
    variant equals the oracle."""
    nr, ns = 1 << 15, 1 << 19
    R, S = O.generate("pkfk", nr, ns, 16, seed=21)
    cases = [("narrow", R, S)]
    S2 = S.copy()
    S2.view(np.uint64)[-1] = np.uint64(1 << 40)  # last S tuple: row id above 32 bits
    R2 = R.copy()
    R2.view(np.uint64)[2 * 100] += np.uint64(1 << 33)  # one build key above 32 bits
    cases += [("wide row id", R, S2), ("wide key", R2, S)]
    for name, hR, hS in cases:
        if algo.startswith("ptr_"):
            want = O.ptr_direct_address(hR, hS, 16) if name == "narrow" else None
            if want is None:
                continue
        else:
            want = O.aggregate(hR, hS, 16, lp=algo.endswith("lp"))
        dR, dS = _dev(ctx, hR, 16), _dev(ctx, hS, 16)
        for bits in (12, 14):
            a = ctx.join(algo, dR, dS, radix_bits=bits, passes=2)
            b = ctx.join(algo, dR, dS, radix_bits=bits, passes=2, flags=hjmi.HJ_F_NO_NARROW)
            _check(a, want, (name, bits))
            _check(b, want, (name, bits, "no narrow"))
        ctx.free_relation(dR)
        ctx.free_relation(dS)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp", "phj_radix_hr", "ptr_phj_radix_bc"])
def test_positional_first_pass_tiles(ctx, algo):
    """Pass A stores a narrow tile whose row ids are its relation positions
    + 1 as 4-byte {key >> first digit, index} words and pass B restores key
    and row id; any other tile keeps 8 or 16 bytes.  The PK-FK generator
    (as the reference's, src/gen.c:198-283) gives S row ids = positions + 1
    and R row ids = keys, so S's tiles qualify and R's do not.  Cases: as
    generated; S rows shuffled (no tile qualifies); one S row id off by one;
    one probe key too large for the word (>= 2^28 at these widths, < 2^32);
    one build key above 32 bits (a wide tile, so the 16-byte second pass
    decodes S's words).  Each equals the oracle and the same join with
    HJ_F_NO_POSITIONAL, and reports how many tuples went as words."""
    nr, ns = 1 << 15, 1 << 19
    R, S = O.generate("pkfk", nr, ns, 16, seed=23)
    rng = np.random.default_rng(5)
    Ssh = S.reshape(ns, 16)[rng.permutation(ns)].reshape(-1).copy()
    Soff = S.copy()
    Soff.view(np.uint64)[2 * 77777 + 1] += np.uint64(1)  # one row id off its position
    Sbig = S.copy()
    Sbig.view(np.uint64)[2 * 300] = np.uint64((1 << 29) + 5)  # narrow, but no room beside the index
    Rwide = R.copy()
    Rwide.view(np.uint64)[2 * 9000] += np.uint64(1 << 34)  # one wide build key
    one_tile = (ns - 8192, ns - 1)  # every S tile but the one holding the changed tuple (tiles of <= 8192)
    cases = [("as generated", R, S, (ns, ns)), ("shuffled S", R, Ssh, (0, 0)), ("S row id off", R, Soff, one_tile),
             ("big probe key", R, Sbig, one_tile), ("wide build key", Rwide, S, (ns, ns))]
    for name, hR, hS, (lo, hi) in cases:
        if algo.startswith("ptr_"):
            if name != "as generated":
                continue
            want = O.ptr_direct_address(hR, hS, 16)
        else:
            want = O.aggregate(hR, hS, 16, lp=algo.endswith("lp"))
        dR, dS = _dev(ctx, hR, 16), _dev(ctx, hS, 16)
        for bits in (12, 14):
            a = ctx.join(algo, dR, dS, radix_bits=bits, passes=2)
            b = ctx.join(algo, dR, dS, radix_bits=bits, passes=2, flags=hjmi.HJ_F_NO_POSITIONAL)
            _check(a, want, (name, bits))
            _check(b, want, (name, bits, "no positional"))
            assert b["positional_tuples"] == 0, (name, b)
            assert lo <= a["positional_tuples"] <= hi, (name, bits, a["positional_tuples"])
            assert a["narrow"] == (0 if name == "wide build key" else 1), (name, a)
        ctx.free_relation(dR)
        ctx.free_relation(dS)


def test_positional_tiles_of_wide_tuple_heads(ctx):
    """64-byte tuples: pass A reads the 16-byte heads in place and stores
    positional words for them too (S's tiles)."""
    nr, ns = 1 << 14, 1 << 18
    R, S = O.generate("pkfk", nr, ns, 64, seed=29)
    want = O.aggregate(R, S, 64)
    dR, dS = _dev(ctx, R, 64), _dev(ctx, S, 64)
    for bits in (11, 13):
        got = ctx.join("phj_radix_bc", dR, dS, radix_bits=bits, passes=2)
        _check(got, want, bits)
        assert got["positional_tuples"] == ns, got
    ctx.free_relation(dR)
    ctx.free_relation(dS)


@pytest.mark.parametrize("algo", ["phj_radix_bc", "phj_radix_lp"])
def test_tuning_variants_match_oracle(ctx, algo):
    """Every kernel shape a part_variant selects (one-pass tile shapes and
    write-combining lines; two-pass first-pass tiles, second-pass
    workgroups per CU and tile, 8192-tuple first-pass tiles, probe tuples per work item) gives the oracle's result, on narrow tuples
    (PK-FK data: keys below 2^16, row ids below 2^20, so every pass-A tile
    and pass B store 8 bytes) and on wide ones (one build key above 32 bits).
    Build keys are unique, so lp joins run in LDS (no exact-path re-run) and
    the result depends on the kernel shape under test."""
    nr, ns = 1 << 16, 1 << 20
    R, S = O.generate("pkfk", nr, ns, 16, seed=31)
    Rw = R.copy()
    Rw.view(np.uint64)[2 * 7] += np.uint64(1 << 35)  # one wide build key
    for tag, hR in (("narrow", R), ("wide", Rw)):
        want = O.aggregate(hR, S, 16, lp=algo.endswith("lp"))
        dR, dS = _dev(ctx, hR, 16), _dev(ctx, S, 16)

        def run(**kw):
            got = ctx.join(algo, dR, dS, **kw)
            _check(got, want, (tag, kw))
            assert got["fallback"] == 0, (tag, kw, got)
            if kw.get("passes") == 2:
                assert got["narrow"] == (1 if tag == "narrow" else 0), (tag, kw, got)
            return got

        one = [v | (ln << 4) for v in range(6) for ln in (0, 2, 3)] + [1 << 20, 2 << 20, 3 << 20]
        for pv in one:
            run(radix_bits=9, passes=1, part_variant=pv)
        two = [a << 12 for a in range(4)] + [g << 14 for g in range(3)] + [1 << 16, (1 << 16) | (1 << 18),
                                                                           1 << 18, (1 << 18) | (1 << 14),
                                                                           1 << 20, 2 << 20, 3 << 20]
        for pv in two:
            for bits in (11, 14):
                run(radix_bits=bits, passes=2, part_variant=pv)
        ctx.free_relation(dR)
        ctx.free_relation(dS)


@pytest.mark.parametrize("narrow", [True, False])
def test_second_pass_4096_tiles_at_12_bits(ctx, narrow):
    """The 4096-tuple second-pass tile (part_variant bits 18-19 = 1) at its
    largest LDS footprint: 20 radix bits (the library's limit) as 8 + 12, so
    pass B sorts by a 12-bit digit (4096 counters beside the 4096-tuple
    tile)."""
    nr, ns = 1 << 17, 1 << 20
    R, S = O.generate("pkfk", nr, ns, 16, seed=33)
    if not narrow:
        R = R.copy()
        R.view(np.uint64)[2 * 11] += np.uint64(1 << 36)
    want = O.aggregate(R, S, 16)
    dR, dS = _dev(ctx, R, 16), _dev(ctx, S, 16)
    for pv in (0, 1 << 18):
        got = ctx.join("phj_radix_bc", dR, dS, radix_bits=20, passes=2, radix_bits_pass1=8, part_variant=pv)
        _check(got, want, (narrow, hex(pv)))
        assert got["narrow"] == int(narrow), got
    ctx.free_relation(dR)
    ctx.free_relation(dS)


def test_autotune_zipf_and_skew_handling():
    """hj_autotune_zipf times the radix candidates on Zipf-skewed probe sides
    and then the probe-range scale; joins with HJ_F_SKEW_HANDLING take those
    choices and give the oracle's result (as do joins without the flag)."""
    c = hjmi.Context(0)
    try:
        c.autotune_zipf(16, 8, 12, 15, theta=1.5)
        for k in range(12, 16):
            t = c.tuned_skew_config(1 << k)
            assert t["radix_bits"] <= k and t["passes"] in (0, 1, 2) and t["probe_scale"] in (0, 1, 2, 3), (k, t)
            assert t["radix_bits_pass1"] == 0 or (t["passes"] == 2 and 1 <= t["radix_bits_pass1"] < t["radix_bits"])
        for theta, nr in ((1.5, 1 << 15), (1.0, 1 << 14), (0.5, 40000)):
            R, S = O.generate("zipf", nr, 8 * nr, 16, seed=57, theta=theta)
            want = O.aggregate(R, S, 16)
            dR, dS = _dev(c, R, 16), _dev(c, S, 16)
            for flags in (hjmi.HJ_F_SKEW_HANDLING, 0):
                _check(c.join("phj_radix_bc", dR, dS, flags=flags), want, (theta, nr, flags))
            c.free_relation(dR)
            c.free_relation(dS)
    finally:
        c.close()
