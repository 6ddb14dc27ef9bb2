"""ctypes wrapper of the C restatement (oracle/qnb_oracle.c).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Build: `build()` below (gcc,
-O2 -ffp-contract=off), output oracle/libqnb_oracle.so.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libqnb_oracle.so")
SRC = os.path.join(HERE, "qnb_oracle.c")


def build(openmp=False):
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-o", LIB, SRC, "-lm"]
    subprocess.check_call(cmd)
    return LIB


def _lib():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        build()
    lib = C.CDLL(LIB)
    vp = C.c_void_p
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from q6_amd import abi
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [C.POINTER(abi.QnbSystem), C.POINTER(abi.QnbParams)]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_make_pair_lists.argtypes = [vp, C.POINTER(C.c_double)]
    lib.orc_eval_terms.argtypes = [vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                   C.c_int, C.POINTER(abi.QnbEnergies), C.POINTER(abi.QnbQEnergies)]
    lib.orc_calculate_exc.argtypes = [vp, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_uint8),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.orc_get_pairs.restype = C.c_int64
    lib.orc_get_pairs.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.c_int64]
    lib.orc_get_list_params.restype = C.c_int64
    lib.orc_get_list_params.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_int64]
    lib.orc_get_lrf.argtypes = [vp, C.POINTER(C.c_double)]
    lib.orc_set_lrf.argtypes = [vp, C.POINTER(C.c_double)]
    lib.orc_get_shell_state.argtypes = [vp, C.POINTER(C.c_double)]
    lib.orc_set_shell_state.argtypes = [vp, C.POINTER(C.c_double)]
    lib.orc_get_theta.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    return lib, abi


_pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731


class Oracle:
    """Reference-order CPU evaluation of one system (the checker)."""

    def __init__(self, system, **par_overrides):
        self.lib, self.abi = _lib()
        self.sys = system
        self._s, self._p = system.to_structs(**par_overrides)
        self.h = self.lib.orc_create(C.byref(self._s), C.byref(self._p))
        self.nst = int(system.n["nstates"])

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def make_pair_lists(self, x):
        x = np.ascontiguousarray(x, np.float64)
        self.lib.orc_make_pair_lists(self.h, _pd(x))

    def eval_terms(self, mask, x, Tfree=0.0, istep=0, d=None):
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        E = self.abi.QnbEnergies()
        EQ = (self.abi.QnbQEnergies * self.nst)()
        self.lib.orc_eval_terms(self.h, mask, _pd(x), _pd(d), Tfree, istep, C.byref(E), EQ)
        return d, E.as_dict(), [q.as_dict() for q in EQ]

    def calculate_exc(self, x, masks, types, eq_in):
        """calculate_exc: masks (narr, nat_solute) 0/1, types (narr,) 1 full,
        2 electro, 3 vdW; eq_in (nstates, 11) qq el, vdW, qp el, vdW, qw el,
        vdW, Q bond, angle, torsion, improper, restraint.  Returns (narr,
        nstates, 7): qq el, vdW, qp el, vdW, qx el, vdW, total."""
        x = np.ascontiguousarray(x, np.float64)
        masks = np.ascontiguousarray(masks, np.uint8)
        types = np.ascontiguousarray(types, np.int32)
        eq_in = np.ascontiguousarray(eq_in, np.float64)
        narr = len(types)
        out = np.zeros((narr, self.nst, 7))
        self.lib.orc_calculate_exc(self.h, _pd(x), narr, masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   types.ctypes.data_as(C.POINTER(C.c_int32)), _pd(eq_in), _pd(out))
        return out

    def pairs(self, which):
        n = self.lib.orc_get_pairs(self.h, which, None, 0)
        ab = np.zeros((max(n, 1), 2), np.int32)
        self.lib.orc_get_pairs(self.h, which, ab.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return ab[:n]

    def list_params(self, which, npar):
        n = self.lib.orc_get_list_params(self.h, which, None, 0)
        par = np.zeros((max(n, 1), npar), np.float64)
        self.lib.orc_get_list_params(self.h, which, _pd(par), n)
        return par[:n]

    def lrf(self):
        out = np.zeros((int(self.sys.n["ncgp"]), 43))
        self.lib.orc_get_lrf(self.h, _pd(out))
        return out

    def set_lrf(self, tab):
        tab = np.ascontiguousarray(tab, np.float64)
        self.lib.orc_set_lrf(self.h, _pd(tab))

    def shell_state(self):
        n = int(self.sys.par["nwpolr_shell"])
        st = np.zeros(3 * max(n, 1))
        self.lib.orc_get_shell_state(self.h, _pd(st))
        return st[:3 * n]

    def set_shell_state(self, st):
        st = np.ascontiguousarray(st, np.float64)
        self.lib.orc_set_shell_state(self.h, _pd(st))

    def theta(self):
        th = np.zeros(max(int(self.sys.n["nwat"]), 1))
        ni = np.zeros(8, np.int32)
        self.lib.orc_get_theta(self.h, _pd(th), ni.ctypes.data_as(C.POINTER(C.c_int32)))
        return th[:int(self.sys.n["nwat"])], ni[:int(self.sys.par["nwpolr_shell"])]
