"""TEST INFRASTRUCTURE ONLY -- the CPU oracle for the equi-join hot path.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module, and only as the checker; the product never does.

* ``liboracle.so`` (hjoracle.c): restatements of the reference's DRAM-mode
  joins (src/algo/nphj.c, src/algo/phj.c) plus sort-based aggregation and a
  brute-force nested loop.
* numpy closed forms for the generated PK-FK workloads (derived from the
  reference's generator, src/gen.c:174-283, and checked against the
  reference's logged results in tests/golden/reference_logs.json).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(DIR, "liboracle.so")
M64 = (1 << 64) - 1


class OrResult(C.Structure):
    _fields_ = [("matched_cnt", C.c_uint64), ("check_sum", C.c_uint64), ("buildpart_time", C.c_double),
                ("probejoin_time", C.c_double), ("total_time", C.c_double)]


class GenSpec(C.Structure):
    _fields_ = [("workload", C.c_uint32), ("width", C.c_uint32), ("seed", C.c_uint64), ("r_num", C.c_uint64),
                ("s_num", C.c_uint64), ("key_range", C.c_uint64), ("theta", C.c_double)]


WORKLOADS = {"pkfk": 0, "pkfk_ref": 1, "payload_ref": 2, "zipf": 3, "random": 4}
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
        L.or_generate.argtypes = [C.POINTER(GenSpec), P, P, I]
        L.or_generate.restype = None
        for name in ("or_nphj_sc", "or_nphj_lp"):
            getattr(L, name).argtypes = [P, U64, P, U64, U32, I, C.POINTER(OrResult)]
            getattr(L, name).restype = I
        L.or_phj_radix.argtypes = [P, U64, P, U64, U32, U32, U32, I, I, C.POINTER(OrResult)]
        L.or_phj_radix.restype = I
        L.or_phj_asymm_radix.argtypes = [P, U64, P, U64, U32, U32, U32, I, I, C.POINTER(OrResult)]
        L.or_phj_asymm_radix.restype = I
        L.or_radix_partition.argtypes = [P, U64, U32, U32, U32, I, P, P]
        L.or_radix_partition.restype = I
        for name in ("or_aggregate", "or_bruteforce", "or_ptr_bruteforce"):
            getattr(L, name).argtypes = [P, U64, P, U64, U32, I, C.POINTER(OrResult)]
            getattr(L, name).restype = I
        L.or_ptr_join.argtypes = [I, I, U32, U32, P, U64, P, U64, U32, I, C.POINTER(OrResult)]
        L.or_ptr_join.restype = I
        L.or_ref_generate.argtypes = [I, U64, U64, U32, C.c_double, I, I, I, P, P]
        L.or_ref_generate.restype = I
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def generate(workload="pkfk", r_num=1 << 10, s_num=1 << 14, width=16, seed=12345, theta=1.0, key_range=0,
             threads=8):
    """Host copy of the generated relations (uint8 arrays, row-major tuples)."""
    wl = WORKLOADS[workload] if isinstance(workload, str) else int(workload)
    spec = GenSpec(wl, width, seed, r_num, s_num, key_range or r_num, theta)
    r = np.empty(r_num * width, dtype=np.uint8)
    s = np.empty(s_num * width, dtype=np.uint8)
    lib().or_generate(C.byref(spec), _ptr(r), _ptr(s), threads)
    return r, s


def _res(r: OrResult) -> dict:
    return {"matched_cnt": r.matched_cnt, "check_sum": r.check_sum, "buildpart_time": r.buildpart_time,
            "probejoin_time": r.probejoin_time, "total_time": r.total_time}


# per-partition join of the radix joins (phj.c:2515-2535): bc, lp, sc, hr, hro
RADIX_KINDS = {"phj_radix_bc": 0, "phj_radix_lp": 1, "phj_radix_sc": 2, "phj_radix_hr": 3, "phj_radix_hro": 4}


def join(algo: str, R: np.ndarray, S: np.ndarray, width: int, threads: int = 8, radix_bits: int = 11,
         passes: int = 1, radix_bits_pass1: int = 11) -> dict:
    """Run the restated reference join `algo` on host tuples (asymmetric
    names take radix_bits_pass1 first-pass bits of radix_bits)."""
    out = OrResult()
    nr, ns = R.nbytes // width, S.nbytes // width
    L = lib()
    algo = algo.replace("_bw_reg", "")
    if algo.startswith("phj_asymm_radix_"):
        kind = RADIX_KINDS["phj_radix_" + algo[len("phj_asymm_radix_"):]]
        rc = L.or_phj_asymm_radix(_ptr(R), nr, _ptr(S), ns, width, radix_bits, radix_bits_pass1, kind, threads,
                                  C.byref(out))
        if rc:
            raise RuntimeError(f"oracle {algo} failed: {rc}")
        return _res(out)
    if algo == "nphj_sc":
        rc = L.or_nphj_sc(_ptr(R), nr, _ptr(S), ns, width, threads, C.byref(out))
    elif algo == "nphj_lp":
        rc = L.or_nphj_lp(_ptr(R), nr, _ptr(S), ns, width, threads, C.byref(out))
    elif algo in RADIX_KINDS:
        rc = L.or_phj_radix(_ptr(R), nr, _ptr(S), ns, width, radix_bits, passes, RADIX_KINDS[algo],
                            threads, C.byref(out))
    else:
        raise ValueError(algo)
    if rc:
        raise RuntimeError(f"oracle {algo} failed: {rc}")
    return _res(out)


class RowIdError(ValueError):
    """A matched row id lies outside 1..|relation| (the reference's retrieve
    would read out of bounds, src/inc/utils.c:712-723)."""


# late-materialising joins (src/main.c:87-122): inner engine and partition kind
PTR_ALGOS = ["ptr_nphj_sc", "ptr_nphj_lp", "ptr_phj_radix_bc", "ptr_phj_radix_sc", "ptr_phj_radix_lp",
             "ptr_phj_radix_hr", "ptr_phj_radix_bw_reg_bc", "ptr_phj_radix_bw_reg_sc", "ptr_phj_radix_bw_reg_lp",
             "ptr_phj_radix_bw_reg_hr"]


def ptr_join(algo: str, R: np.ndarray, S: np.ndarray, width: int, threads: int = 8, radix_bits: int = 11,
             passes: int = 1) -> dict:
    """Late materialisation: pair count and retrieve checksum of `algo`
    (a ptr_* name) restated on host tuples."""
    inner = algo.replace("_bw_reg", "")
    if not inner.startswith("ptr_"):
        raise ValueError(algo)
    inner = inner[4:]
    if inner == "nphj_sc":
        a, kind = 0, 0
    elif inner == "nphj_lp":
        a, kind = 1, 0
    elif inner in RADIX_KINDS and inner != "phj_radix_hro":
        a, kind = 2, RADIX_KINDS[inner]
    else:
        raise ValueError(algo)
    out = OrResult()
    rc = lib().or_ptr_join(a, kind, radix_bits, passes, _ptr(R), R.nbytes // width, _ptr(S), S.nbytes // width,
                           width, threads, C.byref(out))
    if rc == -1:
        raise RowIdError(f"{algo}: a matched row id lies outside its relation")
    if rc:
        raise RuntimeError(f"oracle {algo} failed: {rc}")
    return _res(out)


def ptr_bruteforce(R: np.ndarray, S: np.ndarray, width: int, lp: bool = False) -> dict:
    """The literal late-materialisation pipeline on small inputs: nested-loop
    pair emission, then the retrieve of src/inc/utils.c:712-723."""
    out = OrResult()
    rc = lib().or_ptr_bruteforce(_ptr(R), R.nbytes // width, _ptr(S), S.nbytes // width, width, int(lp),
                                 C.byref(out))
    if rc == -1:
        raise RowIdError("a matched row id lies outside its relation")
    if rc:
        raise RuntimeError(f"ptr_bruteforce failed: {rc}")
    return _res(out)


def aggregate(R: np.ndarray, S: np.ndarray, width: int, lp: bool = False) -> dict:
    out = OrResult()
    rc = lib().or_aggregate(_ptr(R), R.nbytes // width, _ptr(S), S.nbytes // width, width, int(lp), C.byref(out))
    if rc:
        raise RuntimeError("aggregate failed")
    return _res(out)


def bruteforce(R: np.ndarray, S: np.ndarray, width: int, lp: bool = False) -> dict:
    out = OrResult()
    lib().or_bruteforce(_ptr(R), R.nbytes // width, _ptr(S), S.nbytes // width, width, int(lp), C.byref(out))
    return _res(out)


def radix_partition(rel: np.ndarray, width: int, shift: int, bits: int, threads: int = 4):
    n = rel.nbytes // width
    out = np.empty_like(rel)
    hist = np.zeros(1 << bits, dtype=np.uint64)
    lib().or_radix_partition(_ptr(rel), n, width, shift, bits, threads, _ptr(out), _ptr(hist))
    return out, hist


# --------------------------------------------------------------- closed forms
def keys_payloads(rel: np.ndarray, width: int):
    w = rel.view(np.uint64).reshape(-1, width // 8)
    return w[:, 0], w[:, 1]


def pkfk_closed_form(S: np.ndarray, width: int) -> dict:
    """PK-FK with unique R keys, R.payload = key, every S key present in R:
    count = |S|, checksum = sum over S of (key + payload) mod 2^64."""
    k, p = keys_payloads(S, width)
    return {"matched_cnt": int(k.size), "check_sum": (int(k.sum(dtype=np.uint64)) + int(p.sum(dtype=np.uint64))) & M64}


def pkfk_ref_closed_form(r_num: int, s_num: int) -> dict:
    """Reference-style PK-FK (src/gen.c:199-223): S = |S|/|R| copies of R's
    keys, row_id = key on both sides: checksum = |S| * (|R| + 1) mod 2^64."""
    return {"matched_cnt": s_num, "check_sum": (s_num * (r_num + 1)) & M64}


def payload_ref_closed_form(r_num: int, s_num: int) -> dict:
    """Reference payload-style PK-FK (src/gen.c:226-283): row_id = 1-based
    position on both sides: checksum = |S| * (|S| + |R| + 2) / 2 mod 2^64."""
    return {"matched_cnt": s_num, "check_sum": (s_num * (s_num + r_num + 2) // 2) & M64}


def direct_address(R: np.ndarray, S: np.ndarray, width: int) -> dict:
    """R keys a permutation of 1..|R|: P[key] = payload, then
    count = |S|, checksum = sum_s (P[s.key] + s.payload)."""
    rk, rp = keys_payloads(R, width)
    sk, sp = keys_payloads(S, width)
    P = np.zeros(rk.size + 1, dtype=np.uint64)
    P[rk.astype(np.int64)] = rp
    with np.errstate(over="ignore"):
        s = int(P[sk.astype(np.int64)].sum(dtype=np.uint64)) + int(sp.sum(dtype=np.uint64))
    return {"matched_cnt": int(sk.size), "check_sum": s & M64}


def ptr_direct_address(R: np.ndarray, S: np.ndarray, width: int) -> dict:
    """Late materialisation when R keys are unique and every S key is in R:
    the pair of s is (r.row_id, s.row_id) with r the R row of s.key, and the
    retrieve adds R[r.row_id - 1].row_id + S[s.row_id - 1].row_id.  O(N)."""
    rk, rp = keys_payloads(R, width)
    sk, sp = keys_payloads(S, width)
    order = np.argsort(rk, kind="stable")
    ks = rk[order]
    idx = np.searchsorted(ks, sk)
    if np.any(idx >= ks.size) or np.any(ks[np.minimum(idx, ks.size - 1)] != sk):
        raise ValueError("an S key is missing from R")
    r_rid = rp[order[idx]]
    if np.any(r_rid - 1 >= rp.size) or np.any(sp - 1 >= sp.size):
        raise RowIdError("a matched row id lies outside its relation")
    with np.errstate(over="ignore"):
        s = int(rp[(r_rid - 1).astype(np.int64)].sum(dtype=np.uint64)) + int(
            sp[(sp - 1).astype(np.int64)].sum(dtype=np.uint64))
    return {"matched_cnt": int(sk.size), "check_sum": s & M64}


REF_KINDS = {"pkfk": 0, "dupfk": 1, "zipf": 2, "payload": 3}


def ref_generate(kind: str, r_num: int, s_num: int, width: int = 16, zipf_factor: float = 0.0, seed: int = 12345,
                 exact: bool = True, threads: int = 8):
    """The reference's generator (src/gen.c) driven by this C library's own
    srand()/rand(): host tuples R, S."""
    r = np.empty(r_num * width, dtype=np.uint8)
    s = np.empty(s_num * width, dtype=np.uint8)
    rc = lib().or_ref_generate(REF_KINDS[kind], r_num, s_num, width, zipf_factor, seed, int(exact), threads,
                               _ptr(r), _ptr(s))
    if rc:
        raise ValueError(f"or_ref_generate({kind}, {r_num}, {s_num}, {width}) failed: {rc}")
    return r, s
