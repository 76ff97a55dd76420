"""CPU tests of the C-ABI boundary: libhjmi.so loads, exports every entry
point include/hjmi.h declares, keeps the reference's algorithm names, and the
C host driver rejects what the reference rejects.  No kernel is launched."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "hashjoin-scm_amd", "libhjmi.so")
CLI = os.path.join(ROOT, "hashjoin-scm_amd", "bin", "hjmain")
REFERENCE_ALGOS = ["nphj_sc", "nphj_lp", "phj_radix_bc", "phj_radix_lp", "phj_radix_sc", "phj_radix_hr",
                   "phj_radix_hro", "phj_radix_bw_reg_bc", "phj_radix_bw_reg_lp", "phj_radix_bw_reg_sc",
                   "phj_radix_bw_reg_hr", "phj_radix_bw_reg_hro", "phj_asymm_radix_bc", "phj_asymm_radix_sc",
                   "phj_asymm_radix_hr", "phj_asymm_radix_hro", "phj_asymm_radix_bw_reg_bc",
                   "phj_asymm_radix_bw_reg_sc", "phj_asymm_radix_bw_reg_hr", "phj_asymm_radix_bw_reg_hro", "ptr_nphj_sc", "ptr_nphj_lp", "ptr_phj_radix_bc",
                   "ptr_phj_radix_sc", "ptr_phj_radix_lp", "ptr_phj_radix_hr", "ptr_phj_radix_bw_reg_bc",
                   "ptr_phj_radix_bw_reg_sc", "ptr_phj_radix_bw_reg_lp", "ptr_phj_radix_bw_reg_hr"]


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(hj_\w+)\s*\(", txt, re.M):
                if "static" not in m.group(0) and "HJ_HD" not in m.group(0):
                    names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "hj_join" in names and "hj_generate" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declared_entry_points():
    import hjmi
    import hjmi.tpch
    assert set(hjmi.exported_symbols()) <= set(declared_functions())
    assert set(hjmi.tpch.exported_symbols()) <= set(declared_functions())
    assert "hj_tpch_query" in declared_functions()


def test_registry_keeps_reference_names():
    import hjmi
    assert hjmi.algo_names() == REFERENCE_ALGOS
    lib = hjmi.load_library()
    assert lib.hj_algo_name(len(REFERENCE_ALGOS)) is None
    assert lib.hj_algo_name(-1) is None


def test_status_strings():
    import hjmi
    lib = hjmi.load_library()
    assert lib.hj_strerror(hjmi.HJ_OK) == b"ok"
    assert b"admission" in lib.hj_strerror(hjmi.HJ_EREJECT)
    assert lib.hj_kernel_name(9999) is None


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    import hjmi
    saved = hjmi._lib
    try:
        hjmi._lib = None
        with pytest.raises(ImportError):
            hjmi.load_library(str(tmp_path / "nope.so"))
    finally:
        hjmi._lib = saved
        importlib.reload  # keep flake quiet


def _has_gpu():
    return os.path.exists("/dev/kfd") and subprocess.run(
        ["bash", "-c", "ls /dev/dri/renderD* >/dev/null 2>&1"]).returncode == 0


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU error path")
def test_context_without_gpu_reports_nodev():
    import hjmi
    with pytest.raises(hjmi.HJError) as e:
        hjmi.Context(0)
    assert e.value.code == hjmi.HJ_ENODEV


def test_cli_lists_registry():
    out = subprocess.run([CLI, "--list"], capture_output=True, text=True, check=True).stdout.split()
    assert out == REFERENCE_ALGOS


def test_cli_rejects_unknown_algo_like_reference():
    p = subprocess.run([CLI, "--algo=phj_radix_xx", "--workload=uniform", "--subtype=A", "--param=pkfk"],
                       capture_output=True, text=True)
    assert p.returncode != 0
    assert "does not exist" in p.stdout
