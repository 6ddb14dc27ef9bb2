"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/qnb.h declares; no compute call is made without a GPU."""
import ctypes as C
import os
import re

from conftest import ROOT
from q6_amd import abi


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "qnb.h")).read()
    return sorted(set(re.findall(r"\b(qnb_[a-z_]+)\s*\(", txt)))


def test_header_and_python_export_lists_agree():
    assert header_symbols() == sorted(abi.EXPORTS)


def test_library_exports_every_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_abi_version_and_device_count_without_gpu():
    lib = abi.load()
    assert lib.qnb_abi_version() == abi.ABI_VERSION
    assert lib.qnb_device_count() >= 0


def test_struct_sizes_match_header():
    # the ctypes mirror must have the same size as the C structs
    src = r'''
#include <stdio.h>
#include "qnb.h"
int main(){printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(qnb_system), sizeof(qnb_params), sizeof(qnb_energies),
              sizeof(qnb_qenergies), sizeof(qnb_md_params), sizeof(qnb_qtangent), sizeof(qnb_qbonded));return 0;}
'''
    import subprocess
    import tempfile
    d = tempfile.mkdtemp()
    open(os.path.join(d, "t.c"), "w").write(src)
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(d, "t"), os.path.join(d, "t.c")])
    out = subprocess.check_output([os.path.join(d, "t")]).decode().split()
    assert [int(v) for v in out] == [C.sizeof(abi.QnbSystem), C.sizeof(abi.QnbParams), C.sizeof(abi.QnbEnergies),
                                     C.sizeof(abi.QnbQEnergies), C.sizeof(abi.QnbMdParams),
                                     C.sizeof(abi.QnbQTangent), C.sizeof(abi.QnbQbonded)]


def test_create_fails_loudly_without_gpu():
    import pytest
    lib = abi.load()
    if lib.qnb_device_count() > 0:
        pytest.skip("GPU present")
    from conftest import golden_path
    from q6_amd.system import from_golden
    system = from_golden(golden_path("lig_eq1"))  # keeps the arrays the structs point to alive
    s, p = system.to_structs()
    h = C.c_void_p()
    rc = lib.qnb_create(C.byref(s), C.byref(p), C.byref(h))
    assert rc == 2, lib.qnb_last_error()
    assert b"no HIP device" in lib.qnb_last_error()
