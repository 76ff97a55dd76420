"""CPU tests of the reference-dataset support: the glibc-compatible
generator behind hj_ref_generate, the data configuration parser, and the
logged dupfk / Zipf golden vectors that only the reference's own draws
reproduce (src/gen.c, src/inc/eth_genzipf.c; logs/dupfk-logs,
logs/skew-logs)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import hjmi
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_logs.json")
M64 = (1 << 64) - 1


def _logged(workload):
    for r in json.load(open(GOLD)):
        if r["workload"] == workload:
            return r["matched_cnt"], r["check_sum"]
    raise KeyError(workload)


def _words(rel, width):
    return rel.view(np.uint64).reshape(-1, width // 8)[:, :2]


@pytest.mark.parametrize("seed", [12345, 0, 1, 42, 2147483647])
def test_ref_rand_is_glibc_rand(seed):
    """hj_ref_rand restates glibc random(); compare with this C library's
    own srand()/rand() for 100k draws."""
    libc = C.CDLL("libc.so.6")
    libc.srand(C.c_uint(seed))
    want = np.array([libc.rand() for _ in range(100_000)], dtype=np.int32)
    assert np.array_equal(hjmi.ref_rand(seed, 100_000), want)


CASES = [("pkfk", 1 << 10, 1 << 14, 16, 0.0), ("pkfk", 1 << 10, 1 << 13, 64, 0.0), ("dupfk", 1 << 10, 1 << 14, 16, 0.0),
         ("dupfk", 1 << 12, 1 << 14, 32, 0.0), ("zipf", 1 << 10, 1 << 14, 16, 1.05),
         ("zipf", 1 << 11, 1 << 14, 64, 1.75), ("payload", 1 << 10, 1 << 14, 16, 0.0),
         ("payload", 1 << 10, 1 << 12, 512, 0.0)]


@pytest.mark.parametrize("kind,nr,ns,width,zf", CASES)
def test_ref_generate_matches_glibc_driven_restatement(kind, nr, ns, width, zf):
    """Byte-for-byte against the oracle's restatement of gen.c on glibc
    rand(), shuffles included (the reference's RUN_MINI sizes)."""
    sp = hjmi.RefSpec(hjmi.__dict__["HJ_REF_" + kind.upper()], width, nr, ns, zf, 12345, 1)
    R, S = hjmi.ref_generate(sp)
    oR, oS = O.ref_generate(kind, nr, ns, width, zipf_factor=zf, exact=True, threads=3)
    assert np.array_equal(R, oR) and np.array_equal(S, oS)
    # without the shuffles: the same tuples in another order (payload row
    # ids are positions, so there only the key and row-id columns agree as
    # multisets -- the join checksum is the same either way)
    sp.exact_order = 0
    uR, uS = hjmi.ref_generate(sp)
    for a, b in ((uR, R), (uS, S)):
        wa, wb = _words(a, width), _words(b, width)
        for col in (0, 1):
            assert np.array_equal(np.sort(wa[:, col]), np.sort(wb[:, col]))
        if kind != "payload":
            assert np.array_equal(np.sort(np.ascontiguousarray(wa).view("V16").ravel()),
                                  np.sort(np.ascontiguousarray(wb).view("V16").ravel()))


def test_ref_spec_names_follow_reference_configs():
    sp = hjmi.ref_spec("uniform_dupfk_E")
    assert (sp.kind, sp.r_num, sp.s_num, sp.tuple_size, sp.seed) == (hjmi.HJ_REF_DUPFK, 1 << 26, 1 << 28, 16, 12345)
    sp = hjmi.ref_spec("zipfian_1.750_A")
    assert (sp.kind, sp.r_num, sp.zipf_factor) == (hjmi.HJ_REF_ZIPF, 1 << 24, 1.75)
    sp = hjmi.ref_spec("payload_pkfk_512")
    assert (sp.kind, sp.tuple_size) == (hjmi.HJ_REF_PAYLOAD, 512)
    with pytest.raises(hjmi.HJError):
        hjmi.ref_spec("uniform_pkfk_Z")


def _dupfk_from_r(nr, m):
    """count / checksum of R join S when S is m copies of R (gen_fk_from_pk):
    count = m * sum_k c(k)^2, checksum = 2m * sum_k c(k) * rowids(k)."""
    sp = hjmi.ref_spec({1 << 24: "uniform_dupfk_D", 1 << 26: "uniform_dupfk_E"}[nr])
    sp.s_num = nr  # R does not depend on |S|; one copy keeps the host buffer small
    R, _ = hjmi.ref_generate(sp)
    w = _words(R, 16)
    keys, rids = w[:, 0].astype(np.int64), w[:, 1]
    c = np.bincount(keys, minlength=nr + 1).astype(np.uint64)
    # per-key row-id sums stay below 2^40 (a few row ids < 2^26 per key), so
    # float64 weights add them exactly
    rsum = np.bincount(keys, weights=rids.astype(np.float64), minlength=nr + 1)
    assert rsum.max() < 2.0 ** 52
    rsum = rsum.astype(np.uint64)
    with np.errstate(over="ignore"):
        count = int((np.uint64(m) * c * c).sum(dtype=np.uint64)) & M64
        check = int((np.uint64(2 * m) * c * rsum).sum(dtype=np.uint64)) & M64
    return count, check


@pytest.mark.parametrize("name,nr", [("uniform_dupfk_D", 1 << 24), ("uniform_dupfk_E", 1 << 26)])
def test_dupfk_logged_vectors(name, nr):
    """logs/dupfk-logs: D 534841984 / 8973526820846944, E 528480568 /
    35467655496411728 -- fixed by glibc's draws, reproduced here."""
    assert _dupfk_from_r(nr, (1 << 28) // nr) == _logged(name)


def test_zipf_logged_vector_1_050():
    """logs/skew-logs zipfian_1.050_A: 268435456 / 38229679500325349.  R is
    pk_AR (row_id = key), so checksum = sum of S keys + sum of S row ids."""
    sp = hjmi.ref_spec("zipfian_1.050_A")
    R, S = hjmi.ref_generate(sp)
    w = _words(S, 16)
    got = (w.shape[0], (int(w[:, 0].sum(dtype=np.uint64)) + int(w[:, 1].sum(dtype=np.uint64))) & M64)
    assert got == _logged("zipfian_1.050_A")
    rw = _words(R, 16)
    assert np.array_equal(np.sort(rw[:, 0]), np.arange(1, (1 << 24) + 1, dtype=np.uint64))
    assert np.array_equal(rw[:, 0], rw[:, 1])


CFG = """# reference-style data configuration
r_path_suffix = "uniform/dup_DR.bin";
s_path_suffix = "uniform/fk_DS.bin";   // probe side

r_tuple_num = 16777216;
s_tuple_num = 268435456;
r_tuple_size = 16;
s_tuple_size = 16;
/* skew and
   selectivity */
theta = 0.25;
selectivity = 1.0;
density = 1.0;
sparsity = 1.0;

min_key = 1;
max_key = 16777216;

workload_name = "uniform_dupfk_D";
"""


def test_data_cfg_parser(tmp_path):
    p = tmp_path / "dupfk_D.cfg"
    p.write_text(CFG)
    c = hjmi.read_data_cfg(str(p))
    assert c["r_path_suffix"] == "uniform/dup_DR.bin" and c["s_path_suffix"] == "uniform/fk_DS.bin"
    assert (c["r_tuple_num"], c["s_tuple_num"], c["r_tuple_size"], c["s_tuple_size"]) == (1 << 24, 1 << 28, 16, 16)
    assert (c["theta"], c["min_key"], c["max_key"], c["workload_name"]) == (0.25, 1, 1 << 24, "uniform_dupfk_D")
    sp = hjmi.ref_spec(c["workload_name"], c["r_tuple_size"])
    assert (sp.r_num, sp.s_num) == (c["r_tuple_num"], c["s_tuple_num"])
    bad = tmp_path / "bad.cfg"
    bad.write_text("r_tuple_num 5;\n")
    with pytest.raises(hjmi.HJError):
        hjmi.read_data_cfg(str(bad))
    with pytest.raises(hjmi.HJError):
        hjmi.read_data_cfg(str(tmp_path / "missing.cfg"))
