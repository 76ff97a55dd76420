"""CPU tests of the TPC-H Q14 / Q19 consumer (src/tpch/): the oracle against
hand-computed answers and a literal nested-loop restatement, the library's
.tbl and packed-.bin readers (host code, no GPU) against the oracle's reader,
and the reference's algorithm names and cardinalities."""
import json
import os

import numpy as np
import pytest

import hjmi
import hjmi.tpch as T
import tpch_oracle as TO

MINI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tpch_mini")


def _mini():
    return (TO.read_tbl(os.path.join(MINI, "part.tbl"), "part"),
            TO.read_tbl(os.path.join(MINI, "lineitem.tbl"), "lineitem"))


def _expected():
    return json.load(open(os.path.join(MINI, "expected.json")))


def _sub(d, keys):
    return {k: d[k] for k in keys}


@pytest.mark.parametrize("q", [14, 19])
def test_oracle_known_answers(q):
    """The hand-computed answers of the fixture (every filter boundary: first
    and last September day, quantities 10/11/20/30/31, sizes 5/6, AIR REG vs
    REG AIR, a 10-byte container without NUL, a lineitem without part)."""
    part, line = _mini()
    want = _expected()[f"q{q}"]
    assert _sub(TO.query(q, part, line), want) == want
    assert _sub(TO.bruteforce(q, part, line), want) == want


def test_oracle_double_sums_are_the_exact_ones():
    part, line = _mini()
    got = TO.q14(part, line)
    assert got["numerator_double"] == pytest.approx(81435.97, rel=1e-14)
    assert got["denominator_double"] == pytest.approx(124021.57, rel=1e-14)
    assert got["promo_revenue"] == pytest.approx(100 * 81435.97 / 124021.57, rel=1e-14)
    assert TO.q19(part, line)["revenue_double"] == pytest.approx(64534.85, rel=1e-14)


def _nasty_tables(seed, np_, nl, key_range):
    """Small tables drawn to hit every predicate edge: duplicate part keys,
    keys without parts, strings of full width, embedded NULs, near-miss
    literals and boundary quantities and dates."""
    rng = np.random.default_rng(seed)
    brands = [b"Brand#12", b"Brand#23", b"Brand#34", b"Brand#1", b"Brand#123", b"Brand#34x", b"Brand#", b"Brand#43"]
    conts = [b"SM CASE", b"SM BOX", b"SM PACK", b"SM PKG", b"MED BAG", b"MED BOX", b"MED PKG", b"MED PACK",
             b"LG CASE", b"LG BOX", b"LG PACK", b"LG PKG", b"JUMBO DRUM", b"SM CASE\0XY", b"SM CASEX", b"LG"]
    types = [b"PROMO BRUSHED TIN", b"PROMOTION", b"PROM", b"STANDARD BURNISHED NICKEL", b"promo tin", b"PROMO"]
    dates = [b"1995-08-31", b"1995-09-01", b"1995-09-15", b"1995-09-30", b"1995-10-01", b"1995-09\0\0\0",
             b"1995-09-", b"1996-09-10", b"1994-12-31", b"1995-09-99"]
    instr = [b"DELIVER IN PERSON", b"DELIVER IN PERSONS", b"COLLECT COD", b"NONE", b"DELIVER IN PERSON\0X"]
    modes = [b"AIR", b"AIR REG", b"REG AIR", b"AIRX", b"AIR\0REG", b"RAIL", b"MAIL"]
    part = {"partkey": rng.integers(1, key_range + 1, np_).astype(np.uint64),
            "brand": np.array([brands[i] for i in rng.integers(0, len(brands), np_)], dtype="S10"),
            "type": np.array([types[i] for i in rng.integers(0, len(types), np_)], dtype="S25"),
            "size": rng.integers(-1, 20, np_).astype(np.int32),
            "container": np.array([conts[i] for i in rng.integers(0, len(conts), np_)], dtype="S10")}
    line = {"partkey": rng.integers(1, key_range + 8, nl).astype(np.uint64),
            "quantity": rng.integers(0, 33, nl).astype(np.int32),
            "extendedprice": rng.integers(0, 10_000_000, nl).astype(np.int64),
            "discount": rng.integers(0, 101, nl).astype(np.int32),
            "shipdate": np.array([dates[i] for i in rng.integers(0, len(dates), nl)], dtype="S10"),
            "shipinstruct": np.array([instr[i] for i in rng.integers(0, len(instr), nl)], dtype="S25"),
            "shipmode": np.array([modes[i] for i in rng.integers(0, len(modes), nl)], dtype="S10")}
    return part, line


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("q", [14, 19])
def test_oracle_matches_literal_probe_loop(seed, q):
    part, line = _nasty_tables(seed, 150, 1500, 60)
    got, want = TO.query(q, part, line), TO.bruteforce(q, part, line)
    assert _sub(got, want) == want
    assert want["matched_cnt"] > 0


def test_planted_mistakes_change_the_known_answers(monkeypatch):
    """A plausible slip in the oracle moves a known answer: the specification's
    'REG AIR' in place of the reference's 'AIR REG', or a date bound off by
    one day."""
    part, line = _mini()
    want = _expected()
    monkeypatch.setattr(TO, "Q19_MODES", (b"AIR", b"REG AIR"))
    assert TO.q19(part, line)["revenue_e4"] != want["q19"]["revenue_e4"]
    monkeypatch.undo()
    orig = TO.q14_lineitem_mask
    monkeypatch.setattr(TO, "q14_lineitem_mask",
                        lambda li: (TO._cstr(li["shipdate"]) > b"1995-09-01") & (TO._cstr(li["shipdate"]) < b"1995-10-01"))
    assert TO.q14(part, line)["denominator_e4"] != want["q14"]["denominator_e4"]
    monkeypatch.setattr(TO, "q14_lineitem_mask", orig)
    assert TO.q14(part, line)["denominator_e4"] == want["q14"]["denominator_e4"]


def _same_cols(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k


def test_library_tbl_reader_matches_oracle_reader(tmp_path):
    for table in ("part", "lineitem"):
        path = os.path.join(MINI, f"{table}.tbl")
        _same_cols(T.read_tbl(path, table), TO.read_tbl(path, table))
    part, line = _nasty_tables(7, 300, 4000, 100)
    # rows written as dbgen writes them (no NULs inside a field)
    for cols in (part, line):
        for k, v in cols.items():
            if v.dtype.kind == "S":
                cols[k] = np.array([x.split(b"\0")[0] for x in v], dtype=v.dtype)
    line["discount"] = np.minimum(line["discount"], 99).astype(np.int32)
    line["shipdate"] = np.array([d if len(d) == 10 else b"1995-09-05" for d in line["shipdate"]], dtype="S10")
    for table, cols in (("part", part), ("lineitem", line)):
        p = str(tmp_path / f"{table}.tbl")
        TO.write_tbl(p, table, cols)
        got = T.read_tbl(p, table)
        _same_cols(got, TO.read_tbl(p, table))
        _same_cols(got, cols)


def test_tbl_reader_rejects_malformed_rows(tmp_path):
    bad = tmp_path / "lineitem.tbl"
    bad.write_text("1|2|3|1|5.50|100.00|0.05|0.02|N|O|1995-09-01|1995-09-10|1995-09-12|NONE|AIR|c|\n")
    with pytest.raises(hjmi.HJError):  # fractional quantity
        T.read_tbl(str(bad), "lineitem")
    bad.write_text("1|2|3|1|5|100.00|0.05|0.02|N|O|1995-9-1|1995-09-10|1995-09-12|NONE|AIR|c|\n")
    with pytest.raises(hjmi.HJError):  # short date
        T.read_tbl(str(bad), "lineitem")
    bad.write_text("x|name|mfgr|Brand#12|PROMO|5|SM CASE|1.00|c|\n")
    with pytest.raises(hjmi.HJError):
        T.read_tbl(str(bad), "part")
    with pytest.raises(hjmi.HJError):
        T.read_tbl(str(tmp_path / "missing.tbl"), "part")


def test_tbl2bin_writes_the_reference_rows(tmp_path):
    """Packed part_t (168 bytes) / lineitem_t (175 bytes) rows as
    tpch_tbl2bin.c writes them, read back into the same columns."""
    for table, width in (("part", 168), ("lineitem", 175)):
        src = os.path.join(MINI, f"{table}.tbl")
        dst = str(tmp_path / f"{table}.bin")
        n = T.tbl2bin(table, src, dst)
        assert os.path.getsize(dst) == n * width
        _same_cols(T.read_bin(dst, table), TO.read_tbl(src, table))
    raw = np.fromfile(str(tmp_path / "part.bin"), dtype=np.uint8).reshape(-1, 168)
    assert raw[0, :8].view(np.uint64)[0] == 1
    assert bytes(raw[0, 88:96]) == b"Brand#12"  # 8 + p_name[55] + p_mfgr[25]
    assert raw[0, 127:137].tobytes().rstrip(b"\0") == b"SM CASE"  # after p_type[25] and p_size
    assert raw[0, 137:145].view(np.float64)[0] == 901.0
    rawl = np.fromfile(str(tmp_path / "lineitem.bin"), dtype=np.uint8).reshape(-1, 175)
    assert rawl[1, 32:64].view(np.float64).tolist() == [11.0, 9911.0, 0.10, 0.06]
    assert bytes(rawl[1, 66:76]) == b"1995-08-31"
    assert bytes(rawl[1, 96:113]) == b"DELIVER IN PERSON" and bytes(rawl[1, 121:128]) == b"AIR REG"


def test_reference_algo_names_and_cardinalities():
    assert T.algo_names() == list(T.TPCH_ALGOS)
    assert T.cardinality(1) == (200000, 6001215)
    assert T.cardinality(10) == (2000000, 59986052)
    assert T.cardinality(100) == (20000000, 600037902)
    with pytest.raises(hjmi.HJError):
        T.cardinality(3)


CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hashjoin-scm_amd", "bin", "hjmain")


def test_cli_tbl2bin_matches_library(tmp_path):
    """hjmain --tbl2bin=DIR converts DIR/part.tbl and DIR/lineitem.tbl as
    bin/tpch_tbl2bin does (host only, no GPU)."""
    import shutil
    import subprocess
    for t in ("part", "lineitem"):
        shutil.copy(os.path.join(MINI, f"{t}.tbl"), tmp_path / f"{t}.tbl")
    out = subprocess.run([CLI, f"--tbl2bin={tmp_path}"], capture_output=True, text=True, check=True).stdout
    assert "part_num: 6" in out and "lineitem_num: 12" in out
    for t in ("part", "lineitem"):
        T.tbl2bin(t, os.path.join(MINI, f"{t}.tbl"), str(tmp_path / f"{t}.lib.bin"))
        assert (tmp_path / f"{t}.bin").read_bytes() == (tmp_path / f"{t}.lib.bin").read_bytes()


def test_cli_tpch_rejects_unknown_algo_like_reference():
    import subprocess
    p = subprocess.run([CLI, "--tpch=14", "--algo=tpch_rdx_lp", "--sf=1"], capture_output=True, text=True)
    assert p.returncode == 1 and "Join algorithm named `tpch_rdx_lp' does not exist!" in p.stdout
