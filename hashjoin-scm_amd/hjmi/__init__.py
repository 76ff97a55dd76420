"""Python binding of libhjmi.so, the MI355X equi-join library.

Host-side mirror of the reference's join entry points (src/main.c:66-143):
``Context.join(algo, R, S)`` runs one join selected by the reference's
algorithm name and returns ``matched_cnt``, ``check_sum`` and the phase times
(``buildpart_time``, ``probejoin_time``, ``total_time``) that the reference
prints (src/algo/nphj.c:881, src/algo/phj.c:2206, src/main.c:439-445).

The library is loaded with ctypes from ``hashjoin-scm_amd/libhjmi.so``.  There
is no CPU fallback: if the shared library or a GPU is missing, constructing a
``Context`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libhjmi.so")

HJ_OK, HJ_EINVAL, HJ_ENOMEM, HJ_EHIP, HJ_EREJECT, HJ_ENODEV, HJ_EBUSY = 0, -1, -2, -3, -4, -5, -6
HJ_LOC_DEVICE, HJ_LOC_HOST = 0, 1
HJ_F_KERNEL_TIMES = 1
HJ_F_NO_NARROW = 2
HJ_F_NO_POSITIONAL = 4
HJ_F_BATCH_SCATTER = 8
HJ_F_SKEW_HANDLING = 16
HJ_SERVE_SCATTER = 1
HJ_FB_LP_DUPLICATES, HJ_FB_LP_OVERSIZE, HJ_FB_SC_REBUILD = 1, 2, 4

WL_PKFK_UNIFORM, WL_PKFK_REF, WL_PAYLOAD_REF, WL_ZIPF, WL_RANDOM = 0, 1, 2, 3, 4
WORKLOADS = {"pkfk": WL_PKFK_UNIFORM, "pkfk_ref": WL_PKFK_REF, "payload_ref": WL_PAYLOAD_REF,
             "zipf": WL_ZIPF, "random": WL_RANDOM}

ALGOS = ("nphj_sc", "nphj_lp", "phj_radix_bc", "phj_radix_lp", "phj_radix_sc", "phj_radix_hr", "phj_radix_hro",
         "phj_radix_bw_reg_bc", "phj_radix_bw_reg_lp", "phj_radix_bw_reg_sc", "phj_radix_bw_reg_hr",
         "phj_radix_bw_reg_hro", "phj_asymm_radix_bc", "phj_asymm_radix_sc", "phj_asymm_radix_hr",
         "phj_asymm_radix_hro", "phj_asymm_radix_bw_reg_bc", "phj_asymm_radix_bw_reg_sc",
         "phj_asymm_radix_bw_reg_hr", "phj_asymm_radix_bw_reg_hro")
# late materialisation (src/algo/ptr_*.c): the join emits row-id pairs and a
# retrieve pass computes check_sum = sum R[r_row_id-1].row_id + S[s_row_id-1].row_id
PTR_ALGOS = ("ptr_nphj_sc", "ptr_nphj_lp", "ptr_phj_radix_bc", "ptr_phj_radix_sc", "ptr_phj_radix_lp",
             "ptr_phj_radix_hr", "ptr_phj_radix_bw_reg_bc", "ptr_phj_radix_bw_reg_sc", "ptr_phj_radix_bw_reg_lp",
             "ptr_phj_radix_bw_reg_hr")
MAX_KTIMES = 16


class Relation(C.Structure):
    _fields_ = [("tuples", C.c_void_p), ("tuple_num", C.c_uint64), ("tuple_size", C.c_uint32),
                ("location", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [("radix_bits", C.c_uint32), ("passes", C.c_uint32), ("lp_slots_log2", C.c_uint32),
                ("flags", C.c_uint32), ("part_variant", C.c_uint32), ("radix_bits_pass1", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class Result(C.Structure):
    _fields_ = [("matched_cnt", C.c_uint64), ("check_sum", C.c_uint64), ("buildpart_time", C.c_double),
                ("probejoin_time", C.c_double), ("total_time", C.c_double), ("algo_bytes", C.c_uint64),
                ("radix_bits", C.c_uint32), ("passes", C.c_uint32), ("fallback", C.c_uint32),
                ("n_ktimes", C.c_uint32), ("ktime_ms", C.c_double * MAX_KTIMES),
                ("ktime_id", C.c_uint32 * MAX_KTIMES), ("narrow_tuples", C.c_uint64), ("narrow", C.c_uint32),
                ("reserved", C.c_uint32), ("positional_tuples", C.c_uint64)]


class GenSpec(C.Structure):
    _fields_ = [("workload", C.c_uint32), ("width", C.c_uint32), ("seed", C.c_uint64), ("r_num", C.c_uint64),
                ("s_num", C.c_uint64), ("key_range", C.c_uint64), ("theta", C.c_double)]


class DeviceInfo(C.Structure):
    _fields_ = [("cu_count", C.c_int), ("lds_per_block", C.c_int), ("lds_per_cu", C.c_int),
                ("warp_size", C.c_int), ("hbm_bytes", C.c_uint64), ("clock_khz", C.c_int),
                ("arch", C.c_char * 64)]


class JoinResult(C.Structure):
    _fields_ = [("matched_cnt", C.c_uint64), ("check_sum", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("fallback_ms", C.c_double), ("njoins", C.c_uint32),
                ("n_unpartitioned", C.c_uint32), ("n_partitioned", C.c_uint32), ("n_fallback", C.c_uint32),
                ("work_items", C.c_uint64), ("tuples", C.c_uint64), ("scratch_bytes", C.c_uint64),
                ("algo_bytes", C.c_uint64), ("gathered", C.c_uint32), ("n_ktimes", C.c_uint32),
                ("ktime_ms", C.c_double * 16), ("ktime_id", C.c_uint32 * 16)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_cap_bytes", C.c_uint64), ("slots", C.c_uint32), ("max_batch", C.c_uint32),
                ("solo_lanes", C.c_uint32), ("solo_min_log2", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("batches", C.c_uint64), ("max_batch_joins", C.c_uint64), ("max_busy_slots", C.c_uint32),
                ("n_builds", C.c_uint32), ("resident_bytes", C.c_uint64), ("arena_bytes", C.c_uint64),
                ("peak_admitted_bytes", C.c_uint64), ("solo_completed", C.c_uint64)]


class RefSpec(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("tuple_size", C.c_uint32), ("r_num", C.c_uint64), ("s_num", C.c_uint64),
                ("zipf_factor", C.c_double), ("seed", C.c_int32), ("exact_order", C.c_uint32)]


class DataCfg(C.Structure):
    _fields_ = [("r_path_suffix", C.c_char * 256), ("s_path_suffix", C.c_char * 256),
                ("workload_name", C.c_char * 128), ("r_tuple_num", C.c_uint64), ("s_tuple_num", C.c_uint64),
                ("r_tuple_size", C.c_uint32), ("s_tuple_size", C.c_uint32), ("theta", C.c_double),
                ("selectivity", C.c_double), ("density", C.c_double), ("sparsity", C.c_double),
                ("min_key", C.c_int64), ("max_key", C.c_int64)]


HJ_REF_PKFK, HJ_REF_DUPFK, HJ_REF_ZIPF, HJ_REF_PAYLOAD = 0, 1, 2, 3


class HJError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libhjmi.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build() or "
                          f"`make -C hashjoin-scm_amd`")
    lib = C.CDLL(path)
    P, U64, U32, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    sig = {
        "hj_ctx_create": (I, [I, C.POINTER(P)]),
        "hj_ctx_destroy": (I, [P]),
        "hj_strerror": (C.c_char_p, [I]),
        "hj_last_error": (C.c_char_p, [P]),
        "hj_algo_count": (I, []),
        "hj_algo_name": (C.c_char_p, [I]),
        "hj_kernel_name": (C.c_char_p, [U32]),
        "hj_join": (I, [P, C.c_char_p, C.POINTER(Relation), C.POINTER(Relation), C.POINTER(Params),
                        C.POINTER(Result)]),
        "hj_join_noalloc": (I, [P, C.c_char_p, C.POINTER(Relation), C.POINTER(Relation), C.POINTER(Params),
                                C.POINTER(Result)]),
        "hj_dev_alloc": (I, [P, U64, C.POINTER(P)]),
        "hj_dev_free": (I, [P, P]),
        "hj_copy_h2d": (I, [P, P, P, U64]),
        "hj_copy_d2h": (I, [P, P, P, U64]),
        "hj_sync": (I, [P]),
        "hj_generate": (I, [P, C.POINTER(GenSpec), P, P]),
        "hj_device_info": (I, [P, C.POINTER(DeviceInfo)]),
        "hj_copy_d2d": (I, [P, P, P, U64, C.POINTER(C.c_double)]),
        "hj_batch_create": (I, [P, C.c_char_p, U32, C.POINTER(Relation), C.POINTER(Relation), C.POINTER(Params),
                                C.POINTER(P)]),
        "hj_batch_run": (I, [P, C.POINTER(JoinResult), C.POINTER(BatchInfo)]),
        "hj_batch_destroy": (I, [P]),
        "hj_autotune": (I, [P, U32, U32, U32, U32]),
        "hj_tuned_config": (I, [P, U64, C.POINTER(U32), C.POINTER(U32)]),
        "hj_tuned_pass1_bits": (I, [P, U64, C.POINTER(U32)]),
        "hj_autotune_zipf": (I, [P, U32, U32, U32, U32, C.c_double]),
        "hj_tuned_batch_bits": (I, [P, U64, C.POINTER(U32)]),
        "hj_tuned_skew_config": (I, [P, U64, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]),
        "hj_serve_open": (I, [P, C.c_char_p, C.POINTER(ServeConfig), C.POINTER(P)]),
        "hj_serve_register_build": (I, [P, C.POINTER(Relation), C.POINTER(U32)]),
        "hj_serve_submit": (I, [P, C.c_int32, C.POINTER(Relation), C.POINTER(Relation), C.POINTER(U64)]),
        "hj_serve_poll": (I, [P, U64, C.POINTER(JoinResult), C.POINTER(C.c_double)]),
        "hj_serve_drain": (I, [P]),
        "hj_serve_stats": (I, [P, C.POINTER(ServeStats)]),
        "hj_serve_close": (I, [P]),
        "hj_ref_spec_for": (I, [C.c_char_p, U32, C.POINTER(RefSpec)]),
        "hj_ref_generate": (I, [C.POINTER(RefSpec), P, P]),
        "hj_ref_rand": (I, [C.c_int32, U64, P]),
        "hj_data_cfg_read": (I, [C.c_char_p, C.POINTER(DataCfg)]),
        "hj_bin_read": (I, [P, C.c_char_p, U64, U32, P]),
        "hj_bin_write": (I, [P, C.c_char_p, P, U64, U32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _lib = lib
    return lib


def exported_symbols():
    """Names of the C entry points this binding requires."""
    return ["hj_ctx_create", "hj_ctx_destroy", "hj_strerror", "hj_last_error", "hj_algo_count",
            "hj_algo_name", "hj_kernel_name", "hj_join", "hj_join_noalloc", "hj_dev_alloc", "hj_dev_free", "hj_copy_h2d",
            "hj_copy_d2h", "hj_sync", "hj_generate", "hj_device_info", "hj_copy_d2d", "hj_batch_create",
            "hj_batch_run", "hj_batch_destroy", "hj_autotune", "hj_tuned_config", "hj_tuned_pass1_bits",
            "hj_autotune_zipf", "hj_tuned_skew_config", "hj_tuned_batch_bits",
            "hj_serve_open",
            "hj_serve_register_build", "hj_serve_submit", "hj_serve_poll", "hj_serve_drain", "hj_serve_stats",
            "hj_serve_close", "hj_ref_spec_for", "hj_ref_generate", "hj_ref_rand", "hj_data_cfg_read",
            "hj_bin_read", "hj_bin_write"]


@dataclass
class DeviceRelation:
    """Tuples resident in HBM: the device analogue of relation_t (store.h:147-150)."""
    ptr: int
    tuple_num: int
    tuple_size: int
    owner: "Context | None" = None

    def as_c(self) -> Relation:
        return Relation(self.ptr, self.tuple_num, self.tuple_size, HJ_LOC_DEVICE)

    @property
    def nbytes(self) -> int:
        return self.tuple_num * self.tuple_size


def kernel_name(kid: int) -> str:
    s = load_library().hj_kernel_name(kid)
    return s.decode() if s else f"k{kid}"


class Context:
    """One GPU context (stream, events, scratch arena)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.hj_ctx_create(device, C.byref(h))
        if rc != HJ_OK:
            raise HJError(rc, "hj_ctx_create: " + self.lib.hj_strerror(rc).decode())
        self.h = h
        self._owned = []

    # -- errors
    def _check(self, rc: int, what: str):
        if rc != HJ_OK:
            detail = self.lib.hj_last_error(self.h).decode()
            raise HJError(rc, f"{what}: {self.lib.hj_strerror(rc).decode()}: {detail}")

    def close(self):
        if getattr(self, "h", None):
            for p in self._owned:
                self.lib.hj_dev_free(self.h, C.c_void_p(p))
            self._owned = []
            self.lib.hj_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.hj_dev_alloc(self.h, int(nbytes), C.byref(p)), "hj_dev_alloc")
        self._owned.append(p.value)
        return p.value

    def free(self, ptr: int):
        if ptr in self._owned:
            self._owned.remove(ptr)
            self._check(self.lib.hj_dev_free(self.h, C.c_void_p(ptr)), "hj_dev_free")

    def free_relation(self, rel: DeviceRelation):
        if rel is not None and rel.ptr:
            self.free(rel.ptr)
            rel.ptr = 0

    def upload(self, arr: np.ndarray, tuple_size: int) -> DeviceRelation:
        """Copy a host tuple array (bytes, row-major) into HBM."""
        a = np.ascontiguousarray(arr)
        nbytes = a.nbytes
        if nbytes % tuple_size:
            raise ValueError("array size is not a multiple of tuple_size")
        ptr = self.alloc(max(nbytes, 1))
        if nbytes:
            self._check(self.lib.hj_copy_h2d(self.h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), nbytes),
                        "hj_copy_h2d")
        return DeviceRelation(ptr, nbytes // tuple_size, tuple_size, self)

    def download(self, rel: DeviceRelation) -> np.ndarray:
        out = np.empty(rel.nbytes, dtype=np.uint8)
        if rel.nbytes:
            self._check(self.lib.hj_copy_d2h(self.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(rel.ptr),
                                             rel.nbytes), "hj_copy_d2h")
        return out

    def copy_d2d(self, dst: int, src: int, nbytes: int) -> float:
        """Streaming device copy; returns its device time in ms."""
        ms = C.c_double()
        self._check(self.lib.hj_copy_d2d(self.h, C.c_void_p(dst), C.c_void_p(src), int(nbytes), C.byref(ms)),
                    "hj_copy_d2d")
        return ms.value

    def sync(self):
        self._check(self.lib.hj_sync(self.h), "hj_sync")

    # -- generator
    def generate(self, workload="pkfk", r_num=1 << 10, s_num=1 << 14, width=16, seed=12345,
                 theta=1.0, key_range=0):
        spec = make_spec(workload, r_num, s_num, width, seed, theta, key_range)
        r = self.alloc(max(spec.r_num * spec.width, 1))
        s = self.alloc(max(spec.s_num * spec.width, 1))
        self._check(self.lib.hj_generate(self.h, C.byref(spec), C.c_void_p(r), C.c_void_p(s)), "hj_generate")
        return (DeviceRelation(r, spec.r_num, spec.width, self), DeviceRelation(s, spec.s_num, spec.width, self))

    # -- join
    @staticmethod
    def _params(radix_bits: int = 0, passes: int = 0, lp_slots_log2: int = 0, kernel_times: bool = False,
                part_variant: int = 0, radix_bits_pass1: int = 0, flags: int = 0) -> "Params":
        return Params(radix_bits, passes, lp_slots_log2, (HJ_F_KERNEL_TIMES if kernel_times else 0) | flags,
                      part_variant, radix_bits_pass1)

    def join(self, algo: str, R, S, **kw) -> dict:
        """Run one join on device relations (DeviceRelation).  Keyword
        arguments are hj_params_t's fields: radix_bits, passes,
        lp_slots_log2, kernel_times, part_variant, radix_bits_pass1, flags."""
        rr = []
        for rel in (R, S):
            if isinstance(rel, DeviceRelation):
                rr.append(rel.as_c())
            else:
                raise TypeError("pass DeviceRelation (use Context.upload or Context.host_relation)")
        prm = self._params(**kw)
        res = Result()
        self._check(self.lib.hj_join(self.h, algo.encode(), C.byref(rr[0]), C.byref(rr[1]), C.byref(prm),
                                     C.byref(res)), f"hj_join({algo})")
        return result_dict(res)

    def join_noalloc(self, algo: str, R, S, **kw) -> dict:
        """hj_join_noalloc: nphj_sc / nphj_lp on device relations with the
        library's fixed per-GPU scratch (no device allocation in the call)."""
        if not (isinstance(R, DeviceRelation) and isinstance(S, DeviceRelation)):
            raise TypeError("hj_join_noalloc takes DeviceRelation inputs")
        prm = self._params(**kw)
        res = Result()
        self._check(self.lib.hj_join_noalloc(self.h, algo.encode(), C.byref(R.as_c()), C.byref(S.as_c()),
                                             C.byref(prm), C.byref(res)), f"hj_join_noalloc({algo})")
        return result_dict(res)

    def join_host(self, algo: str, r_host: np.ndarray, s_host: np.ndarray, tuple_size: int, **kw) -> dict:
        """Join host-resident relations: the library copies them in (untimed).
        Takes the same keyword arguments as join()."""
        ra, sa = np.ascontiguousarray(r_host), np.ascontiguousarray(s_host)
        R = Relation(ra.ctypes.data_as(C.c_void_p), ra.nbytes // tuple_size, tuple_size, HJ_LOC_HOST)
        S = Relation(sa.ctypes.data_as(C.c_void_p), sa.nbytes // tuple_size, tuple_size, HJ_LOC_HOST)
        prm = self._params(**kw)
        res = Result()
        self._check(self.lib.hj_join(self.h, algo.encode(), C.byref(R), C.byref(S), C.byref(prm), C.byref(res)),
                    f"hj_join({algo})")
        return result_dict(res)

    def bin_read(self, path: str, tuple_num: int, tuple_size: int) -> DeviceRelation:
        """Load a raw tuple file (the reference's .bin layout) into HBM."""
        ptr = self.alloc(max(tuple_num * tuple_size, 1))
        self._check(self.lib.hj_bin_read(self.h, path.encode(), tuple_num, tuple_size, C.c_void_p(ptr)),
                    "hj_bin_read")
        return DeviceRelation(ptr, tuple_num, tuple_size, self)

    def bin_write(self, path: str, rel: DeviceRelation):
        self._check(self.lib.hj_bin_write(self.h, path.encode(), C.c_void_p(rel.ptr), rel.tuple_num,
                                          rel.tuple_size), "hj_bin_write")

    def batch(self, algo: str, Rs, Ss, part_variant: int = 0, flags: int = 0) -> "Batch":
        return Batch(self, algo, Rs, Ss, part_variant, flags)

    def autotune(self, width: int = 16, s_ratio: int = 16, min_log2: int = 10, max_log2: int = 20):
        self._check(self.lib.hj_autotune(self.h, width, s_ratio, min_log2, max_log2), "hj_autotune")

    def tuned_config(self, r_num: int):
        b, p = C.c_uint32(), C.c_uint32()
        self._check(self.lib.hj_tuned_config(self.h, r_num, C.byref(b), C.byref(p)), "hj_tuned_config")
        return b.value, p.value

    def autotune_zipf(self, width: int = 16, s_ratio: int = 8, min_log2: int = 10, max_log2: int = 20,
                      theta: float = 1.0):
        self._check(self.lib.hj_autotune_zipf(self.h, width, s_ratio, min_log2, max_log2, float(theta)),
                    "hj_autotune_zipf")

    def tuned_skew_config(self, r_num: int):
        b, p, b1, ps = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self.lib.hj_tuned_skew_config(self.h, r_num, C.byref(b), C.byref(p), C.byref(b1), C.byref(ps)),
                    "hj_tuned_skew_config")
        return {"radix_bits": b.value, "passes": p.value, "radix_bits_pass1": b1.value, "probe_scale": ps.value}

    def tuned_batch_bits(self, r_num: int) -> int:
        b = C.c_uint32()
        self._check(self.lib.hj_tuned_batch_bits(self.h, r_num, C.byref(b)), "hj_tuned_batch_bits")
        return b.value

    def tuned_pass1_bits(self, r_num: int) -> int:
        b1 = C.c_uint32()
        self._check(self.lib.hj_tuned_pass1_bits(self.h, r_num, C.byref(b1)), "hj_tuned_pass1_bits")
        return b1.value

    def device_info(self) -> dict:
        d = DeviceInfo()
        self._check(self.lib.hj_device_info(self.h, C.byref(d)), "hj_device_info")
        return {"cu_count": d.cu_count, "lds_per_block": d.lds_per_block, "lds_per_cu": d.lds_per_cu,
                "warp_size": d.warp_size, "hbm_bytes": d.hbm_bytes, "clock_khz": d.clock_khz,
                "arch": d.arch.decode()}


def make_spec(workload="pkfk", r_num=1 << 10, s_num=1 << 14, width=16, seed=12345, theta=1.0, key_range=0):
    wl = WORKLOADS[workload] if isinstance(workload, str) else int(workload)
    return GenSpec(wl, width, seed, r_num, s_num, key_range or r_num, theta)


def result_dict(res: Result) -> dict:
    kt = {}
    for i in range(res.n_ktimes):
        name = kernel_name(res.ktime_id[i])
        kt.setdefault(name, []).append(res.ktime_ms[i])
    return {
        "matched_cnt": res.matched_cnt, "check_sum": res.check_sum,
        "buildpart_time": res.buildpart_time, "probejoin_time": res.probejoin_time,
        "total_time": res.total_time, "algo_bytes": res.algo_bytes, "radix_bits": res.radix_bits,
        "passes": res.passes, "fallback": res.fallback, "kernel_ms": kt,
        "narrow_tuples": res.narrow_tuples, "narrow": res.narrow, "positional_tuples": res.positional_tuples,
    }


# ------------------------------------------------- the reference's datasets
def ref_spec(workload_name: str, tuple_size: int = 16, exact_order: bool = False) -> RefSpec:
    """Generator spec of a reference workload ("uniform_dupfk_D",
    "zipfian_1.250_A", "payload_pkfk_64", ...)."""
    sp = RefSpec()
    rc = load_library().hj_ref_spec_for(workload_name.encode(), tuple_size, C.byref(sp))
    if rc != HJ_OK:
        raise HJError(rc, f"hj_ref_spec_for({workload_name})")
    sp.exact_order = int(exact_order)
    return sp


def ref_generate(spec: RefSpec):
    """Host tuples (uint8 arrays R, S) drawn as the reference's generator
    draws them."""
    r = np.empty(spec.r_num * spec.tuple_size, dtype=np.uint8)
    s = np.empty(spec.s_num * spec.tuple_size, dtype=np.uint8)
    rc = load_library().hj_ref_generate(C.byref(spec), r.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p))
    if rc != HJ_OK:
        raise HJError(rc, "hj_ref_generate")
    return r, s


def ref_rand(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    load_library().hj_ref_rand(seed, n, out.ctypes.data_as(C.c_void_p))
    return out


def read_data_cfg(path: str) -> dict:
    """A reference data configuration file as a dict of its settings."""
    c = DataCfg()
    rc = load_library().hj_data_cfg_read(path.encode(), C.byref(c))
    if rc != HJ_OK:
        raise HJError(rc, f"hj_data_cfg_read({path})")
    out = {}
    for name, _ in DataCfg._fields_:
        v = getattr(c, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


def algo_names():
    lib = load_library()
    return [lib.hj_algo_name(i).decode() for i in range(lib.hj_algo_count())]


class Batch:
    """Many joins as one captured HIP graph (hj_batch_*); per-join results."""

    def __init__(self, ctx: Context, algo: str, Rs, Ss, part_variant: int = 0, flags: int = 0):
        if len(Rs) != len(Ss):
            raise ValueError("one build and one probe relation per join")
        self.ctx, self.n = ctx, len(Rs)
        self._R = (Relation * max(1, self.n))(*[r.as_c() for r in Rs])
        self._S = (Relation * max(1, self.n))(*[x.as_c() for x in Ss])
        h = C.c_void_p()
        prm = Params(part_variant=part_variant, flags=flags)
        ctx._check(ctx.lib.hj_batch_create(ctx.h, algo.encode(), self.n, self._R, self._S, C.byref(prm),
                                           C.byref(h)), "hj_batch_create")
        self.h = h

    def run(self):
        res = (JoinResult * max(1, self.n))()
        info = BatchInfo()
        self.ctx._check(self.ctx.lib.hj_batch_run(self.h, res, C.byref(info)), "hj_batch_run")
        out = [(res[i].matched_cnt, res[i].check_sum) for i in range(self.n)]
        d = {f: getattr(info, f) for f, _ in BatchInfo._fields_ if f not in ("ktime_ms", "ktime_id")}
        d["kernel_ms"] = {}
        for i in range(info.n_ktimes):
            name = kernel_name(info.ktime_id[i])
            d["kernel_ms"][name] = d["kernel_ms"].get(name, 0.0) + info.ktime_ms[i]
        return out, d

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.hj_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching join server (hj_serve_*)."""

    def __init__(self, ctx: Context, algo: str = "phj_radix_bc", hbm_cap_bytes: int = 8 << 30, slots: int = 3,
                 max_batch: int = 64, solo_lanes: int = 0, solo_min_log2: int = 0, flags: int = 0):
        self.ctx = ctx
        cfg = ServeConfig(hbm_cap_bytes, slots, max_batch, solo_lanes, solo_min_log2, flags)
        h = C.c_void_p()
        ctx._check(ctx.lib.hj_serve_open(ctx.h, algo.encode(), C.byref(cfg), C.byref(h)), "hj_serve_open")
        self.h = h

    def register_build(self, R: DeviceRelation) -> int:
        bid = C.c_uint32()
        r = R.as_c()
        self.ctx._check(self.ctx.lib.hj_serve_register_build(self.h, C.byref(r), C.byref(bid)),
                        "hj_serve_register_build")
        return bid.value

    def submit(self, S: DeviceRelation, R: DeviceRelation = None, build: int = -1) -> int:
        rid = C.c_uint64()
        s = S.as_c()
        r = R.as_c() if R is not None else Relation()
        self.ctx._check(self.ctx.lib.hj_serve_submit(self.h, build, C.byref(r), C.byref(s), C.byref(rid)),
                        "hj_serve_submit")
        return rid.value

    def poll(self, rid: int):
        """(matched_cnt, check_sum), latency_ms; or None while running."""
        res, lat = JoinResult(), C.c_double()
        rc = self.ctx.lib.hj_serve_poll(self.h, rid, C.byref(res), C.byref(lat))
        if rc == HJ_EBUSY:
            return None
        self.ctx._check(rc, "hj_serve_poll")
        return (res.matched_cnt, res.check_sum), lat.value

    def drain(self):
        self.ctx._check(self.ctx.lib.hj_serve_drain(self.h), "hj_serve_drain")

    def stats(self) -> dict:
        st = ServeStats()
        self.ctx._check(self.ctx.lib.hj_serve_stats(self.h, C.byref(st)), "hj_serve_stats")
        return {f: getattr(st, f) for f, _ in ServeStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.hj_serve_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
