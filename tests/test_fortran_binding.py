"""The Fortran ISO_C_BINDING module compiles with amdflang, links against the
engine library and its BIND(C) structs have the C ABI's sizes (CPU only: the
program calls qnb_abi_version and qnb_device_count, no compute)."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT
from q6_amd import abi

PROG = r'''
program t
  use qnb_bind
  implicit none
  type(qnb_system) :: s
  type(qnb_params) :: p
  type(qnb_energies) :: e
  type(qnb_qenergies) :: q
  type(qnb_qtangent) :: tg
  type(qnb_qbonded) :: qb
  print '(i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,i0,1x,i0)', qnb_abi_version(), c_sizeof(s), c_sizeof(p), &
        c_sizeof(e), c_sizeof(q), c_sizeof(tg), c_sizeof(qb), merge(1, 0, qnb_device_count() >= 0)
end program
'''


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not on PATH")
def test_fortran_binding_links_and_matches_c_layout():
    d = tempfile.mkdtemp()
    src = os.path.join(ROOT, "q6_amd", "fortran", "qnb_bind.f90")
    open(os.path.join(d, "t.f90"), "w").write(PROG)
    lib = os.path.join(ROOT, "q6_amd")
    subprocess.check_call(["amdflang", "-c", src, "-module-dir", d, "-o", os.path.join(d, "qnb_bind.o")])
    subprocess.check_call(["amdflang", "-I", d, os.path.join(d, "t.f90"), os.path.join(d, "qnb_bind.o"),
                           "-L" + lib, "-lq6nb", "-Wl,-rpath," + lib, "-o", os.path.join(d, "t")])
    out = subprocess.check_output([os.path.join(d, "t")]).decode().split()
    assert [int(v) for v in out] == [abi.ABI_VERSION, C.sizeof(abi.QnbSystem), C.sizeof(abi.QnbParams), C.sizeof(abi.QnbEnergies),
                                     C.sizeof(abi.QnbQEnergies), C.sizeof(abi.QnbQTangent),
                                     C.sizeof(abi.QnbQbonded), 1]
