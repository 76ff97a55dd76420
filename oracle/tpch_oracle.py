"""TEST INFRASTRUCTURE ONLY -- the CPU oracle for TPC-H Q14 and Q19.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module, and only as the checker; the product never does.

Restates the reference's end-to-end consumer over host columns (numpy
arrays; strings are fixed-width ``S<w>`` arrays holding the reference's
NUL-padded char fields, src/tpch/inc/schema.h:37-67):

* ``q14`` / ``q19``: the selections part_filter / lineitem_filter /
  match_filter (src/tpch/inc/tpch_utils.c:44-160) vectorised, the join on
  partkey as a count of all matching pairs, and the sums of
  src/tpch/algo/tpch_nphj.c:302-319 -- exactly, as integers in units of 1e-4
  (cents * (100 - discount hundredths)), and also as the reference adds them,
  in doubles.
* ``bruteforce``: the same query as the reference's probe loop runs it, one
  (part, lineitem) pair at a time with the row predicates written out as
  strncmp calls; pure Python, for tiny tables.  It pins the vectorised form.
* ``read_tbl``: a pure-Python reader of dbgen .tbl rows into the same
  columns (checks the library's reader).
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1

PART_COLS = {"partkey": np.uint64, "brand": "S10", "type": "S25", "size": np.int32, "container": "S10"}
LINEITEM_COLS = {"partkey": np.uint64, "quantity": np.int32, "extendedprice": np.int64, "discount": np.int32,
                 "shipdate": "S10", "shipinstruct": "S25", "shipmode": "S10"}

Q19_GROUPS = (  # (brand, max size, containers, quantity range): tpch_utils.c:44-82, :95-108
    (12, 5, (b"SM CASE", b"SM BOX", b"SM PACK", b"SM PKG"), (1, 11)),
    (23, 10, (b"MED BAG", b"MED BOX", b"MED PKG", b"MED PACK"), (10, 20)),
    (34, 15, (b"LG CASE", b"LG BOX", b"LG PACK", b"LG PKG"), (20, 30)),
)
Q19_MODES = (b"AIR", b"AIR REG")  # the reference's literals (tpch_utils.c:99-101)


# ------------------------------------------------------------------ vectorised
def _cstr(col: np.ndarray) -> np.ndarray:
    """The strings as C reads them: everything after the first NUL zeroed."""
    w = col.dtype.itemsize
    b = np.ascontiguousarray(col).view(np.uint8).reshape(-1, w).copy()
    b[np.cumsum(b == 0, axis=1) > 0] = 0
    return b.view(f"S{w}").ravel()


def _fast_atoi(col: np.ndarray, start: int) -> np.ndarray:
    """fast_atoi(field + start) (src/inc/utils.h:185-191), within the field."""
    w = col.dtype.itemsize
    b = np.ascontiguousarray(col).view(np.uint8).reshape(-1, w)[:, start:].astype(np.int64)
    v = np.zeros(b.shape[0], dtype=np.int64)
    live = np.ones(b.shape[0], dtype=bool)
    for j in range(b.shape[1]):
        d = (b[:, j] >= 48) & (b[:, j] <= 57)
        live &= d
        v = np.where(live, v * 10 + (b[:, j] - 48), v)
    return v


def revenue_e4(lineitem) -> np.ndarray:
    """l_extendedprice * (1 - l_discount) in units of 1e-4, exactly."""
    return (lineitem["extendedprice"].astype(np.int64) * (100 - lineitem["discount"].astype(np.int64))).astype(
        np.uint64)


def q14_lineitem_mask(lineitem) -> np.ndarray:
    d8 = _cstr(lineitem["shipdate"]).astype("S8")
    return (d8 >= b"1995-09-") & (d8 < b"1995-10-")


def q14_promo_mask(part) -> np.ndarray:
    return np.char.startswith(_cstr(part["type"]), b"PROMO")


def q19_part_group(part) -> np.ndarray:
    """-1, or the group 0..2 whose brand / size / container a part has."""
    brand = _fast_atoi(part["brand"], 6)
    size = part["size"].astype(np.int64)
    ct = _cstr(part["container"])
    g = np.full(len(size), -1, dtype=np.int64)
    for i, (b, smax, cts, _) in enumerate(Q19_GROUPS):
        m = (brand == b) & (size >= 1) & (size <= smax) & np.isin(ct, np.array(cts, dtype="S10"))
        g = np.where((g < 0) & m, i, g)
    return g


def q19_lineitem_common(lineitem) -> np.ndarray:
    return (_cstr(lineitem["shipinstruct"]) == b"DELIVER IN PERSON") & np.isin(
        _cstr(lineitem["shipmode"]), np.array(Q19_MODES, dtype="S10"))


def _pair_sums(rkeys, lkeys, lrev):
    """(pairs, sum of lrev over pairs) for all pairs rkeys[r] == lkeys[l]."""
    if len(rkeys) == 0 or len(lkeys) == 0:
        return 0, 0, 0.0
    uk, cnt = np.unique(rkeys, return_counts=True)
    pos = np.minimum(np.searchsorted(uk, lkeys), len(uk) - 1)
    c = np.where(uk[pos] == lkeys, cnt[pos], 0).astype(np.uint64)
    with np.errstate(over="ignore"):
        s = int((c * lrev).sum(dtype=np.uint64))
    return int(c.sum(dtype=np.uint64)), s & M64, c


def _double_sum(lineitem, sel, weight):
    """The reference's floating sum: l_extendedprice * (1 - l_discount)."""
    price = lineitem["extendedprice"][sel].astype(np.float64) / 100.0
    disc = lineitem["discount"][sel].astype(np.float64) / 100.0
    return float(np.sum(price * (1.0 - disc) * weight))


def q14(part, lineitem) -> dict:
    sel = q14_lineitem_mask(lineitem)
    lk, lrev = lineitem["partkey"][sel], revenue_e4(lineitem)[sel]
    promo = q14_promo_mask(part)
    n_all, den, c_all = _pair_sums(part["partkey"], lk, lrev)
    _, num, c_promo = _pair_sums(part["partkey"][promo], lk, lrev)
    out = {"matched_cnt": n_all, "numerator_e4": num, "denominator_e4": den,
           "part_selected": len(part["partkey"]), "lineitem_selected": int(sel.sum())}
    out["promo_revenue"] = 100.0 * num / den if den else 0.0
    w_all = np.asarray(c_all, dtype=np.float64) if n_all else 0.0
    w_promo = np.asarray(c_promo, dtype=np.float64) if num else 0.0
    out["denominator_double"] = _double_sum(lineitem, sel, w_all) if n_all else 0.0
    out["numerator_double"] = _double_sum(lineitem, sel, w_promo) if num else 0.0
    return out


def q19(part, lineitem) -> dict:
    g = q19_part_group(part)
    common = q19_lineitem_common(lineitem)
    q = lineitem["quantity"].astype(np.int64)
    rev = revenue_e4(lineitem)
    matched, total, dsum = 0, 0, 0.0
    any_q = np.zeros(len(q), dtype=bool)
    for i, (_, _, _, (lo, hi)) in enumerate(Q19_GROUPS):
        sel = common & (q >= lo) & (q <= hi)
        any_q |= sel
        n, s, c = _pair_sums(part["partkey"][g == i], lineitem["partkey"][sel], rev[sel])
        matched += n
        total = (total + s) & M64
        if n:
            dsum += _double_sum(lineitem, sel, np.asarray(c, dtype=np.float64))
    return {"matched_cnt": matched, "revenue_e4": total, "revenue_double": dsum,
            "part_selected": int((g >= 0).sum()), "lineitem_selected": int(any_q.sum())}


def query(q: int, part, lineitem) -> dict:
    return q14(part, lineitem) if q == 14 else q19(part, lineitem)


# ------------------------------------------------------------------ literal
def strncmp(a: bytes, b: bytes, n: int) -> int:
    """C strncmp on a fixed-width field `a` and a literal `b`."""
    for i in range(n):
        ca = a[i] if i < len(a) else 0
        cb = b[i] if i < len(b) else 0
        if ca != cb:
            return ca - cb
        if ca == 0:
            return 0
    return 0


def _field(col, i) -> bytes:
    w = col.dtype.itemsize
    return bytes(np.ascontiguousarray(col).view(np.uint8)[i * w:(i + 1) * w])


def _atoi(s: bytes) -> int:
    v = 0
    for ch in s:
        if not 48 <= ch <= 57:
            break
        v = v * 10 + ch - 48
    return v


def _part_group_literal(brand, size, ct):
    b = _atoi(brand[6:])
    for i, (gb, smax, cts, _) in enumerate(Q19_GROUPS):
        if b == gb and 1 <= size <= smax and any(strncmp(ct, c, 10) == 0 for c in cts):
            return i
    return -1


def bruteforce(q: int, part, lineitem) -> dict:
    """tpch_nphj.c's probe loop (:297-319) over every (part, lineitem) pair,
    with the filters of tpch_utils.c written out; exact integer sums."""
    np_, nl = len(part["partkey"]), len(lineitem["partkey"])
    lsel = []
    for j in range(nl):
        if q == 14:
            d = _field(lineitem["shipdate"], j)
            ok = strncmp(d, b"1995-09-01", 8) >= 0 and strncmp(d, b"1995-10-01", 8) < 0
        else:
            qty = int(lineitem["quantity"][j])
            ok = (strncmp(_field(lineitem["shipinstruct"], j), b"DELIVER IN PERSON", 25) == 0
                  and (strncmp(_field(lineitem["shipmode"], j), b"AIR", 10) == 0
                       or strncmp(_field(lineitem["shipmode"], j), b"AIR REG", 10) == 0)
                  and (1 <= qty <= 11 or 10 <= qty <= 20 or 20 <= qty <= 30))
        if ok:
            lsel.append(j)
    psel = []
    for i in range(np_):
        if q == 14 or _part_group_literal(_field(part["brand"], i), int(part["size"][i]),
                                          _field(part["container"], i)) >= 0:
            psel.append(i)
    matched = num = den = rev = 0
    for i in psel:
        pk = int(part["partkey"][i])
        promo = strncmp(_field(part["type"], i), b"PROMO", 5) == 0
        grp = -1 if q == 14 else _part_group_literal(_field(part["brand"], i), int(part["size"][i]),
                                                     _field(part["container"], i))
        for j in lsel:
            if int(lineitem["partkey"][j]) != pk:
                continue
            r = int(lineitem["extendedprice"][j]) * (100 - int(lineitem["discount"][j]))
            if q == 14:
                matched += 1
                den += r
                num += r if promo else 0
            else:
                qty = int(lineitem["quantity"][j])
                lo, hi = Q19_GROUPS[grp][3]
                if lo <= qty <= hi:  # match_filter: the part's own group
                    matched += 1
                    rev += r
    if q == 14:
        return {"matched_cnt": matched, "numerator_e4": num & M64, "denominator_e4": den & M64,
                "part_selected": len(psel), "lineitem_selected": len(lsel)}
    return {"matched_cnt": matched, "revenue_e4": rev & M64, "part_selected": len(psel),
            "lineitem_selected": len(lsel)}


# ------------------------------------------------------------------ files
def _money(s: str) -> int:
    whole, _, frac = s.partition(".")
    neg = whole.startswith("-")
    v = int(whole.lstrip("-") or "0") * 100 + int((frac + "00")[:2])
    return -v if neg else v


def read_tbl(path: str, table: str) -> dict:
    """dbgen .tbl rows into columns (pure Python)."""
    rows = [ln.rstrip("\n").split("|") for ln in open(path, encoding="latin-1") if ln.strip()]
    if table == "part":
        return {"partkey": np.array([int(r[0]) for r in rows], dtype=np.uint64),
                "brand": np.array([r[3].encode("latin-1") for r in rows], dtype="S10"),
                "type": np.array([r[4].encode("latin-1") for r in rows], dtype="S25"),
                "size": np.array([int(r[5]) for r in rows], dtype=np.int32),
                "container": np.array([r[6].encode("latin-1") for r in rows], dtype="S10")}
    return {"partkey": np.array([int(r[1]) for r in rows], dtype=np.uint64),
            "quantity": np.array([_money(r[4]) // 100 for r in rows], dtype=np.int32),
            "extendedprice": np.array([_money(r[5]) for r in rows], dtype=np.int64),
            "discount": np.array([_money(r[6]) for r in rows], dtype=np.int32),
            "shipdate": np.array([r[10].encode("latin-1") for r in rows], dtype="S10"),
            "shipinstruct": np.array([r[13].encode("latin-1") for r in rows], dtype="S25"),
            "shipmode": np.array([r[14].encode("latin-1") for r in rows], dtype="S10")}


def write_tbl(path: str, table: str, cols: dict):
    """Columns back to dbgen-style rows (the fields the queries do not read
    are filled with fixed placeholders)."""
    def s(col, i):
        return _field(col, i).split(b"\0")[0].decode("latin-1")

    with open(path, "w", encoding="latin-1") as f:
        n = len(cols["partkey"])
        for i in range(n):
            if table == "part":
                f.write(f"{int(cols['partkey'][i])}|name {i}|Manufacturer#1|{s(cols['brand'], i)}|"
                        f"{s(cols['type'], i)}|{int(cols['size'][i])}|{s(cols['container'], i)}|901.00|"
                        f"comment|\n")
            else:
                p, d = int(cols["extendedprice"][i]), int(cols["discount"][i])
                f.write(f"{i + 1}|{int(cols['partkey'][i])}|7|1|{int(cols['quantity'][i])}.00|"
                        f"{p // 100}.{p % 100:02d}|{d // 100}.{d % 100:02d}|0.04|N|O|{s(cols['shipdate'], i)}|"
                        f"1996-01-01|1996-01-02|{s(cols['shipinstruct'], i)}|{s(cols['shipmode'], i)}|c|\n")
