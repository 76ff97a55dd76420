"""GPU parity of TPC-H Q14 / Q19 (src/tpch/tpch_query.c): selections on the
device, the join on partkey by each of the reference's TPC-H algorithms,
exact revenue sums.  Checked against the hand-computed answers of the .tbl
fixture and against the CPU oracle on adversarial and generated tables."""
import json
import os

import numpy as np
import pytest

import hjmi
import hjmi.tpch as T
import tpch_oracle as TO

pytestmark = pytest.mark.gpu

MINI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tpch_mini")
KEYS = {14: ("matched_cnt", "numerator_e4", "denominator_e4", "part_selected", "lineitem_selected"),
        19: ("matched_cnt", "revenue_e4", "part_selected", "lineitem_selected")}


def _sub(d, keys):
    return {k: d[k] for k in keys}


def _nasty(seed, np_, nl, key_range):
    from test_tpch import _nasty_tables
    return _nasty_tables(seed, np_, nl, key_range)


def test_fixture_known_answers_every_algo(ctx):
    part = T.read_tbl(os.path.join(MINI, "part.tbl"), "part")
    line = T.read_tbl(os.path.join(MINI, "lineitem.tbl"), "lineitem")
    want = json.load(open(os.path.join(MINI, "expected.json")))
    t = T.DeviceTables.upload(ctx, part, line)
    for algo in T.TPCH_ALGOS:
        for q in (14, 19):
            got = T.query(ctx, q, algo, t)
            assert _sub(got, want[f"q{q}"]) == want[f"q{q}"], (algo, q)
        got = T.query(ctx, 14, algo, t)
        assert got["numerator"] == pytest.approx(81435.97, rel=1e-15)
        assert got["denominator"] == pytest.approx(124021.57, rel=1e-15)
        assert got["promo_revenue"] == pytest.approx(100 * 81435.97 / 124021.57, rel=1e-15)
        assert got["joins"] == 2 and T.query(ctx, 19, algo, t)["joins"] == 3
    t.free()


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_adversarial_tables_match_oracle(ctx, seed):
    """Duplicate part keys (many-to-many pairs), embedded NULs, full-width
    strings, near-miss literals, boundary dates and quantities."""
    part, line = _nasty(seed, 3000, 40000, 800)
    t = T.DeviceTables.upload(ctx, part, line)
    for q in (14, 19):
        want = _sub(TO.query(q, part, line), KEYS[q])
        assert want["matched_cnt"] > 0
        for algo in T.TPCH_ALGOS:
            assert _sub(T.query(ctx, q, algo, t), KEYS[q]) == want, (q, algo)
    t.free()


@pytest.mark.parametrize("part_rows,line_rows", [(20000, 600000), (200000, 6001215)])
def test_generated_tables_match_oracle(ctx, part_rows, line_rows):
    """TPC-H-like columns generated on the device (the SF1 cardinalities in
    the second case), downloaded and queried by the oracle."""
    t = T.DeviceTables.generate(ctx, part_rows, line_rows, seed=77)
    part, line = t.download()
    assert np.array_equal(part["partkey"], np.arange(1, part_rows + 1, dtype=np.uint64))
    assert line["partkey"].min() >= 1 and line["partkey"].max() <= part_rows
    assert set(np.unique(line["discount"])) <= set(range(11))
    assert line["quantity"].min() >= 1 and line["quantity"].max() <= 50
    d = np.sort(line["shipdate"])
    assert d[0] >= b"1992-01-02" and d[-1] <= b"1998-12-31"
    assert np.array_equal(line["extendedprice"] % line["quantity"], np.zeros(line_rows, dtype=np.int64))
    for q in (14, 19):
        want = TO.query(q, part, line)
        assert want["matched_cnt"] > 0
        for algo in T.TPCH_ALGOS:
            got = T.query(ctx, q, algo, t)
            assert _sub(got, KEYS[q]) == _sub(want, KEYS[q]), (q, algo)
        if q == 14:
            assert got["promo_revenue"] == pytest.approx(want["promo_revenue"], rel=1e-12)
            assert got["denominator"] == pytest.approx(want["denominator_double"], rel=1e-9)
        else:
            assert got["revenue"] == pytest.approx(want["revenue_double"], rel=1e-9)
    t.free()


def test_generator_is_seeded(ctx):
    a = T.DeviceTables.generate(ctx, 1000, 20000, seed=5)
    b = T.DeviceTables.generate(ctx, 1000, 20000, seed=5)
    c = T.DeviceTables.generate(ctx, 1000, 20000, seed=6)
    (pa, la), (pb, lb), (pc, lc) = a.download(), b.download(), c.download()
    for k in la:
        assert np.array_equal(la[k], lb[k]), k
    assert not np.array_equal(la["partkey"], lc["partkey"])
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    for x in (a, b, c):
        x.free()


def test_bin_files_through_the_device(ctx, tmp_path):
    """.tbl -> the reference's packed .bin -> columns -> HBM -> query."""
    want = json.load(open(os.path.join(MINI, "expected.json")))
    cols = {}
    for table in ("part", "lineitem"):
        b = str(tmp_path / f"{table}.bin")
        T.tbl2bin(table, os.path.join(MINI, f"{table}.tbl"), b)
        cols[table] = T.read_bin(b, table)
    t = T.DeviceTables.upload(ctx, cols["part"], cols["lineitem"])
    assert _sub(T.query(ctx, 19, "tpch_rdx_bc", t), want["q19"]) == want["q19"]
    t.free()


def test_empty_tables_and_errors(ctx):
    part, line = TO.read_tbl(os.path.join(MINI, "part.tbl"), "part"), TO.read_tbl(
        os.path.join(MINI, "lineitem.tbl"), "lineitem")
    empty_p, empty_l = T.empty("part", 0), T.empty("lineitem", 0)
    for p, l in ((empty_p, line), (part, empty_l), (empty_p, empty_l)):
        t = T.DeviceTables.upload(ctx, p, l)
        for algo in ("tpch_nphj_sc", "tpch_rdx_hr"):
            for q in (14, 19):
                got = T.query(ctx, q, algo, t)
                assert got["matched_cnt"] == 0 and got["revenue_e4"] == 0 and got["numerator_e4"] == 0
        t.free()
    t = T.DeviceTables.upload(ctx, part, line)
    with pytest.raises(hjmi.HJError) as e:
        T.query(ctx, 14, "tpch_nphj_lp", t)
    assert e.value.code == hjmi.HJ_EINVAL
    with pytest.raises(hjmi.HJError):
        T.query(ctx, 6, "tpch_rdx_bc", t)
    assert T.query(ctx, 14, "tpch_rdx_bc", t)["matched_cnt"] == 8  # still usable
    t.free()


def test_cli_prints_the_reference_lines(tmp_path):
    """hjmain --tpch=Q (the drop-in for bin/tpch_query) on the fixture, from
    .tbl files and from the reference's .bin files."""
    import re
    import shutil
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hashjoin-scm_amd", "bin",
                       "hjmain")
    for t in ("part", "lineitem"):
        shutil.copy(os.path.join(MINI, f"{t}.tbl"), tmp_path / f"{t}.tbl")
    subprocess.run([cli, f"--tbl2bin={tmp_path}"], check=True, capture_output=True)
    for src in (f"--tbl-dir={tmp_path}", f"--tpch-bin-dir={tmp_path}"):
        out = subprocess.run([cli, "--tpch=14", "--algo=tpch_rdx_hr", src], capture_output=True, text=True,
                             check=True).stdout
        assert re.search(r"matched_cnt: 8\tpromo_revenue: 65\.66\tnumerator: 81435\.97\tdenominator: 124021\.57", out), out
        out = subprocess.run([cli, "--tpch=19", "--algo=tpch_nphj_sc", src, "--trials=2"], capture_output=True,
                             text=True, check=True).stdout
        assert out.count("matched_cnt: 5\trevenue: 64534.85") == 2, out
    out = subprocess.run([cli, "--tpch=14", "--sf=1"], capture_output=True, text=True, check=True).stdout
    assert "lineitem_num: 6001215" in out and "promo_revenue:" in out
