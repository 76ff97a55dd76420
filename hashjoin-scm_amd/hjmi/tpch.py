"""TPC-H Q14 and Q19 through libhjmi.so: host mirror of the reference's
tpch_query driver (src/tpch/tpch_query.c).

``query(ctx, 14, "tpch_rdx_bc", tables)`` runs one query over columns resident
in HBM and returns what tpch_spj prints (matched_cnt, promo_revenue /
numerator / denominator for Q14, revenue for Q19, tpch_query.c:739-745) plus
the exact integer sums (units of 1e-4) and the phase times (pushdown,
buildpart, probejoin).  Host columns are dicts of numpy arrays; strings are
fixed-width ``S<w>`` arrays, i.e. the reference's NUL-padded char fields.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import HJ_OK, HJError, load_library

PART_COLS = (("partkey", np.uint64, 1), ("brand", np.uint8, 10), ("type", np.uint8, 25), ("size", np.int32, 1),
             ("container", np.uint8, 10))
LINEITEM_COLS = (("partkey", np.uint64, 1), ("quantity", np.int32, 1), ("extendedprice", np.int64, 1),
                 ("discount", np.int32, 1), ("shipdate", np.uint8, 10), ("shipinstruct", np.uint8, 25),
                 ("shipmode", np.uint8, 10))
TPCH_ALGOS = ("tpch_nphj_sc", "tpch_rdx_bc", "tpch_rdx_hr", "tpch_rdx_bw_reg_bc", "tpch_rdx_bw_reg_hr")


class Part(C.Structure):
    _fields_ = [("rows", C.c_uint64)] + [(n, C.c_void_p) for n, _, _ in PART_COLS]


class Lineitem(C.Structure):
    _fields_ = [("rows", C.c_uint64)] + [(n, C.c_void_p) for n, _, _ in LINEITEM_COLS]


class TpchResult(C.Structure):
    _fields_ = [("matched_cnt", C.c_uint64), ("numerator_e4", C.c_uint64), ("denominator_e4", C.c_uint64),
                ("revenue_e4", C.c_uint64), ("promo_revenue", C.c_double), ("numerator", C.c_double),
                ("denominator", C.c_double), ("revenue", C.c_double), ("part_selected", C.c_uint64),
                ("lineitem_selected", C.c_uint64), ("part_tuples", C.c_uint64), ("lineitem_tuples", C.c_uint64),
                ("joins", C.c_uint32), ("reserved", C.c_uint32),
                ("pushdown_time", C.c_double), ("buildpart_time", C.c_double), ("probejoin_time", C.c_double),
                ("total_time", C.c_double)]


_SIGS = None


def _lib():
    global _SIGS
    lib = load_library()
    if _SIGS is None:
        P, U64, I = C.c_void_p, C.c_uint64, C.c_int
        _SIGS = {
            "hj_tpch_algo_count": (I, []),
            "hj_tpch_algo_name": (C.c_char_p, [I]),
            "hj_tpch_cardinality": (I, [I, C.POINTER(U64), C.POINTER(U64)]),
            "hj_tpch_generate": (I, [P, U64, C.POINTER(Part), C.POINTER(Lineitem)]),
            "hj_tpch_query": (I, [P, I, C.c_char_p, C.POINTER(Part), C.POINTER(Lineitem), C.POINTER(TpchResult)]),
            "hj_tpch_tbl_rows": (I, [C.c_char_p, C.POINTER(U64)]),
            "hj_tpch_tbl_read_part": (I, [C.c_char_p, C.POINTER(Part)]),
            "hj_tpch_tbl_read_lineitem": (I, [C.c_char_p, C.POINTER(Lineitem)]),
            "hj_tpch_tbl2bin": (I, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(U64)]),
            "hj_tpch_bin_read_part": (I, [C.c_char_p, C.POINTER(Part)]),
            "hj_tpch_bin_read_lineitem": (I, [C.c_char_p, C.POINTER(Lineitem)]),
        }
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
    return lib


def exported_symbols():
    return ["hj_tpch_algo_count", "hj_tpch_algo_name", "hj_tpch_cardinality", "hj_tpch_generate", "hj_tpch_query",
            "hj_tpch_tbl_rows", "hj_tpch_tbl_read_part", "hj_tpch_tbl_read_lineitem", "hj_tpch_tbl2bin",
            "hj_tpch_bin_read_part", "hj_tpch_bin_read_lineitem"]


def _check(rc: int, what: str):
    if rc != HJ_OK:
        raise HJError(rc, what)


def _schema(table: str):
    if table not in ("part", "lineitem"):
        raise ValueError(f"table must be 'part' or 'lineitem', not {table!r}")
    return (PART_COLS, Part) if table == "part" else (LINEITEM_COLS, Lineitem)


def algo_names():
    lib = _lib()
    return [lib.hj_tpch_algo_name(i).decode() for i in range(lib.hj_tpch_algo_count())]


def cardinality(sf: int):
    p, l = C.c_uint64(), C.c_uint64()
    _check(_lib().hj_tpch_cardinality(sf, C.byref(p), C.byref(l)), f"hj_tpch_cardinality({sf})")
    return p.value, l.value


# ------------------------------------------------------------------ host columns
def empty(table: str, rows: int) -> dict:
    """Host columns of `rows` rows (strings as fixed-width S<w> arrays)."""
    cols, _ = _schema(table)
    return {n: (np.zeros(rows, dtype=f"S{w}") if dt is np.uint8 else np.zeros(rows, dtype=dt)) for n, dt, w in cols}


def _host_struct(table: str, cols: dict):
    schema, S = _schema(table)
    rows = len(cols["partkey"])
    st = S(rows)
    keep = []
    for n, dt, w in schema:
        a = cols[n]
        want = np.dtype(f"S{w}") if dt is np.uint8 else np.dtype(dt)
        if a.dtype != want or len(a) != rows or not a.flags.c_contiguous:
            raise ValueError(f"column {n}: need {rows} contiguous {want}")
        keep.append(a)
        setattr(st, n, a.ctypes.data if rows else None)
    return st, keep


def read_tbl(path: str, table: str) -> dict:
    """A dbgen .tbl file into host columns."""
    lib = _lib()
    n = C.c_uint64()
    _check(lib.hj_tpch_tbl_rows(path.encode(), C.byref(n)), f"hj_tpch_tbl_rows({path})")
    cols = empty(table, n.value)
    st, _ = _host_struct(table, cols)
    fn = lib.hj_tpch_tbl_read_part if table == "part" else lib.hj_tpch_tbl_read_lineitem
    _check(fn(path.encode(), C.byref(st)), f"read {table} .tbl {path}")
    return cols


def tbl2bin(table: str, tbl_path: str, bin_path: str) -> int:
    """The reference's tbl2bin: a .tbl file into packed part_t / lineitem_t rows."""
    _schema(table)
    n = C.c_uint64()
    _check(_lib().hj_tpch_tbl2bin(table.encode(), tbl_path.encode(), bin_path.encode(), C.byref(n)),
           f"hj_tpch_tbl2bin({table})")
    return n.value


BIN_ROW_BYTES = {"part": 168, "lineitem": 175}


def read_bin(path: str, table: str) -> dict:
    """The reference's packed part.bin / lineitem.bin into host columns."""
    import os
    _schema(table)
    size = os.path.getsize(path)
    if size % BIN_ROW_BYTES[table]:
        raise HJError(-1, f"{path}: {size} bytes is not a whole number of {table} rows")
    cols = empty(table, size // BIN_ROW_BYTES[table])
    st, _ = _host_struct(table, cols)
    lib = _lib()
    fn = lib.hj_tpch_bin_read_part if table == "part" else lib.hj_tpch_bin_read_lineitem
    _check(fn(path.encode(), C.byref(st)), f"read {table} .bin {path}")
    return cols


# ------------------------------------------------------------------ device tables
class DeviceTables:
    """part and lineitem columns resident in HBM on one Context."""

    def __init__(self, ctx, part_rows: int, lineitem_rows: int):
        self.ctx = ctx
        self.part, self.lineitem = Part(part_rows), Lineitem(lineitem_rows)
        self._ptrs = []
        for st, schema, rows in ((self.part, PART_COLS, part_rows), (self.lineitem, LINEITEM_COLS, lineitem_rows)):
            for n, dt, w in schema:
                p = ctx.alloc(max(rows * np.dtype(dt).itemsize * w, 16))
                self._ptrs.append(p)
                setattr(st, n, p)

    @classmethod
    def generate(cls, ctx, part_rows: int, lineitem_rows: int, seed: int = 12345) -> "DeviceTables":
        t = cls(ctx, part_rows, lineitem_rows)
        ctx._check(_lib().hj_tpch_generate(ctx.h, seed, C.byref(t.part), C.byref(t.lineitem)), "hj_tpch_generate")
        return t

    @classmethod
    def from_sf(cls, ctx, sf: int, seed: int = 12345) -> "DeviceTables":
        return cls.generate(ctx, *cardinality(sf), seed=seed)

    @classmethod
    def upload(cls, ctx, part: dict, lineitem: dict) -> "DeviceTables":
        t = cls(ctx, len(part["partkey"]), len(lineitem["partkey"]))
        for st, cols, table in ((t.part, part, "part"), (t.lineitem, lineitem, "lineitem")):
            hs, _ = _host_struct(table, cols)
            for n, dt, w in _schema(table)[0]:
                a = cols[n]
                if a.nbytes:
                    ctx._check(ctx.lib.hj_copy_h2d(ctx.h, C.c_void_p(getattr(st, n)), a.ctypes.data_as(C.c_void_p),
                                                   a.nbytes), "hj_copy_h2d")
        return t

    def download(self):
        out = []
        for st, table in ((self.part, "part"), (self.lineitem, "lineitem")):
            cols = empty(table, st.rows)
            for n, dt, w in _schema(table)[0]:
                a = cols[n]
                if a.nbytes:
                    self.ctx._check(self.ctx.lib.hj_copy_d2h(self.ctx.h, a.ctypes.data_as(C.c_void_p),
                                                             C.c_void_p(getattr(st, n)), a.nbytes), "hj_copy_d2h")
            out.append(cols)
        return tuple(out)

    @property
    def nbytes(self) -> int:
        tot = 0
        for st, schema in ((self.part, PART_COLS), (self.lineitem, LINEITEM_COLS)):
            tot += sum(st.rows * np.dtype(dt).itemsize * w for _, dt, w in schema)
        return tot

    def free(self):
        for p in self._ptrs:
            self.ctx.free(p)
        self._ptrs = []


def result_dict(r: TpchResult) -> dict:
    return {name: getattr(r, name) for name, _ in TpchResult._fields_ if name != "reserved"}


def query(ctx, q: int, algo: str, tables: DeviceTables) -> dict:
    """Run TPC-H Q14 or Q19 with one of TPCH_ALGOS over device tables."""
    _lib()
    res = TpchResult()
    ctx._check(ctx.lib.hj_tpch_query(ctx.h, q, algo.encode(), C.byref(tables.part), C.byref(tables.lineitem),
                                     C.byref(res)), f"hj_tpch_query(Q{q}, {algo})")
    return result_dict(res)
