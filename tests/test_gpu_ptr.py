"""GPU parity of the late-materialising joins (ptr_*): the device joins the
{key, row_id} heads, writes one row-id pair per match and retrieves
check_sum = sum R[r_row_id-1].row_id + S[s_row_id-1].row_id.  Checked
against the payload closed form (row_id = position, the reference's ptr
workload), the oracle's restated ptr joins and the literal nested-loop
pipeline on data whose row ids point anywhere, plus error handling."""
import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu

PTR = list(hjmi.PTR_ALGOS)


def _pair(got):
    return got["matched_cnt"], got["check_sum"]


def _pointer_relations(seed, nr, ns, width, key_range):
    R, S = O.generate("random", nr, ns, width, seed=seed, key_range=key_range)
    rng = np.random.default_rng(seed)
    for rel, n in ((R, nr), (S, ns)):
        w = rel.view(np.uint64).reshape(n, width // 8)
        w[:, 1] = rng.integers(1, n + 1, size=n, dtype=np.uint64)
    return R, S


@pytest.mark.parametrize("width", [16, 32, 64, 128, 256, 512])
def test_ptr_payload_closed_form(ctx, width):
    nr, ns = 1 << 14, 1 << 18
    R, S = ctx.generate("payload_ref", nr, ns, width)
    want = O.payload_ref_closed_form(nr, ns)
    for algo in PTR:
        assert _pair(ctx.join(algo, R, S)) == (want["matched_cnt"], want["check_sum"]), (width, algo)
    for bits, passes in ((0, 0), (6, 1), (12, 2)):
        got = ctx.join("ptr_phj_radix_bc", R, S, radix_bits=bits, passes=passes)
        assert _pair(got) == (want["matched_cnt"], want["check_sum"]), (width, bits, passes)
    ctx.free_relation(R)
    ctx.free_relation(S)


@pytest.mark.parametrize("width", [16, 64])
def test_ptr_pkfk_direct_address(ctx, width):
    """Uniform pk-fk (R.row_id = key, a permutation): the retrieve follows
    R.row_id to another row, which the O(N) direct-address oracle mirrors."""
    nr, ns = 1 << 16, 1 << 20
    hR, hS = O.generate("pkfk", nr, ns, width, seed=77)
    want = O.ptr_direct_address(hR, hS, width)
    R, S = ctx.upload(hR, width), ctx.upload(hS, width)
    for algo in PTR:
        assert _pair(ctx.join(algo, R, S)) == (want["matched_cnt"], want["check_sum"]), algo
    ctx.free_relation(R)
    ctx.free_relation(S)


@pytest.mark.parametrize("seed,nr,ns,kr,width", [(1, 300, 2000, 100, 16), (2, 1000, 3000, 5000, 32),
                                                  (3, 64, 64, 8, 16), (4, 700, 900, 700, 64),
                                                  (5, 5000, 40000, 1500, 16), (6, 3000, 20000, 2500, 512)])
def test_ptr_duplicates_and_misses_match_literal_pipeline(ctx, seed, nr, ns, kr, width):
    hR, hS = _pointer_relations(seed, nr, ns, width, kr)
    want_all = O.ptr_bruteforce(hR, hS, width)
    want_lp = O.ptr_bruteforce(hR, hS, width, lp=True)
    R, S = ctx.upload(hR, width), ctx.upload(hS, width)
    for algo in PTR:
        want = want_lp if algo.endswith("_lp") else want_all
        for bits, passes in ((0, 0), (3, 1), (8, 2)):
            got = ctx.join(algo, R, S, radix_bits=bits, passes=passes)
            assert _pair(got) == (want["matched_cnt"], want["check_sum"]), (algo, bits, passes)
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_ptr_pair_buffer_grows_past_probe_side():
    """Many duplicates: matches far exceed |S|, so the first join overflows
    the pair buffer (sized |S| in a fresh context) and runs again with the
    exact count."""
    hR, hS = _pointer_relations(11, 4000, 4000, 16, 20)
    want_all = O.ptr_bruteforce(hR, hS, 16)
    want_lp = O.ptr_bruteforce(hR, hS, 16, lp=True)
    assert want_all["matched_cnt"] > 100 * 4000
    for algo in ("ptr_phj_radix_bc", "ptr_nphj_sc", "ptr_phj_radix_hr", "ptr_phj_radix_lp"):
        with hjmi.Context(0) as c:  # fresh context: pair buffer starts at |S|
            R, S = c.upload(hR, 16), c.upload(hS, 16)
            want = want_lp if algo.endswith("_lp") else want_all
            assert _pair(c.join(algo, R, S)) == (want["matched_cnt"], want["check_sum"]), algo
            assert _pair(c.join(algo, R, S)) == (want["matched_cnt"], want["check_sum"]), algo


def test_ptr_rejects_out_of_range_row_ids(ctx):
    """Random payloads are not row positions: the retrieve reports the bad
    pairs as HJ_EINVAL (the reference reads out of bounds); the context stays
    usable."""
    R, S = ctx.generate("random", 2000, 20000, 16, key_range=1000)
    for algo in ("ptr_phj_radix_bc", "ptr_nphj_lp"):
        with pytest.raises(hjmi.HJError) as e:
            ctx.join(algo, R, S)
        assert e.value.code == hjmi.HJ_EINVAL and "row-id pairs" in str(e.value)
    R2, S2 = ctx.generate("payload_ref", 1000, 16000, 16)
    want = O.payload_ref_closed_form(1000, 16000)
    assert _pair(ctx.join("ptr_phj_radix_bc", R2, S2)) == (want["matched_cnt"], want["check_sum"])
    for x in (R, S, R2, S2):
        ctx.free_relation(x)


def test_ptr_empty_and_host_relations(ctx):
    hR, hS = O.generate("payload_ref", 1000, 16000, 64)
    want = O.payload_ref_closed_form(1000, 16000)
    for algo in ("ptr_phj_radix_sc", "ptr_nphj_sc"):
        assert _pair(ctx.join_host(algo, hR, hS, 64)) == (want["matched_cnt"], want["check_sum"]), algo
    empty = np.zeros(0, dtype=np.uint8)
    for algo in PTR:
        assert _pair(ctx.join_host(algo, empty, hS, 64)) == (0, 0), algo
        assert _pair(ctx.join_host(algo, hR, empty, 64)) == (0, 0), algo


def test_plain_joins_on_512_byte_tuples(ctx):
    """The payload sweep's widest tuples through the plain joins too (hro
    takes at most 64 bytes); a one-pass request runs as two tile sorts."""
    nr, ns = 1 << 12, 1 << 16
    R, S = ctx.generate("payload_ref", nr, ns, 512)
    want = O.payload_ref_closed_form(nr, ns)
    for algo in hjmi.ALGOS:
        if algo.endswith("hro"):
            with pytest.raises(hjmi.HJError) as e:
                ctx.join(algo, R, S)
            assert e.value.code == hjmi.HJ_EINVAL
            continue
        assert _pair(ctx.join(algo, R, S)) == (want["matched_cnt"], want["check_sum"]), algo
    for bits, passes in ((0, 0), (1, 1), (6, 1), (6, 2), (10, 2)):
        got = ctx.join("phj_radix_bc", R, S, radix_bits=bits, passes=passes)
        assert _pair(got) == (want["matched_cnt"], want["check_sum"]), (bits, passes)
    hR, hS = _pointer_relations(8, 2000, 9000, 512, 700)
    for algo in ("phj_radix_bc", "phj_radix_lp", "nphj_sc", "phj_radix_hr"):
        w = O.aggregate(hR, hS, 512, lp=algo.endswith("lp"))
        dR, dS = ctx.upload(hR, 512), ctx.upload(hS, 512)
        assert _pair(ctx.join(algo, dR, dS, radix_bits=6, passes=2)) == (w["matched_cnt"], w["check_sum"]), algo
        ctx.free_relation(dR)
        ctx.free_relation(dS)
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_ptr_kernel_times_name_the_phases(ctx):
    R, S = ctx.generate("payload_ref", 1 << 12, 1 << 16, 64)
    got = ctx.join("ptr_phj_radix_bc", R, S, kernel_times=True)
    names = set(got["kernel_ms"])
    assert {"ptr_keyrid", "ptr_retrieve"} <= names, names
    assert got["total_time"] >= got["buildpart_time"] + got["probejoin_time"] - 1e-9
    ctx.free_relation(R)
    ctx.free_relation(S)
