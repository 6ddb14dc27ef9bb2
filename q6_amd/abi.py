"""ctypes mirror of include/qnb.h and loader of the HIP engine library.

The engine is the in-tree shared library q6_amd/libq6nb.so (built by
__graft_entry__.build()).  Loading fails loudly when it is missing: there is no
CPU fallback in the product path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libq6nb.so")

QNB_MAX_STATES = 8
QNB_MAX_SHELLS = 8
QNB_MAX_NBR_RANGE = 25

QNB_PREC_FP64 = 0
QNB_PREC_MIXED = 1

LIST_PP, LIST_PW, LIST_WW, LIST_QP, LIST_QW, LIST_QQ, LIST_QQP = range(7)

T_WW, T_QW, T_PW, T_PP, T_QP, T_LRF, T_QQ, T_QQP, T_RAD, T_POL, T_SINT = (1 << k for k in range(11))
T_NONBONDED = T_WW | T_QW | T_PW | T_PP | T_QP | T_LRF | T_SINT
T_ALL = (1 << 11) - 1

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int32)
_pu8 = C.POINTER(C.c_uint8)


class QnbSystem(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("nat_solute", C.c_int32), ("nwat", C.c_int32), ("solv_atom", C.c_int32),
        ("ncgp", C.c_int32), ("ncgp_solute", C.c_int32), ("nqat", C.c_int32), ("nstates", C.c_int32),
        ("ntypes", C.c_int32), ("ivdw_rule", C.c_int32), ("solvent_spc", C.c_int32),
        ("nexlong", C.c_int32), ("n14long", C.c_int32), ("any_atom_cutoff", C.c_int32),
        ("crg", _pd), ("iac", _pi), ("iqatom", _pi), ("excl", _pu8), ("heavy", _pu8),
        ("cgp_first", _pi), ("cgp_last", _pi), ("cgp_iswitch", _pi), ("cgpatom", _pi),
        ("avdw", _pd), ("bvdw", _pd), ("ljcod", _pi), ("el14_scale", C.c_double),
        ("listex", _pu8), ("list14", _pu8), ("listexlong", _pi), ("list14long", _pi),
        ("pw_A", _pd), ("pw_B", _pd), ("pw_el", _pd),
        ("ww_A", _pd), ("ww_B", _pd), ("ww_el", _pd),
        ("iqseq", _pi), ("qconn3", _pu8),
        ("qp_A", _pd), ("qp_B", _pd), ("qp_el", _pd), ("qp_sc", _pd), ("qp_set", _pu8),
        ("qq_A", _pd), ("qq_B", _pd), ("qq_el", _pd), ("qq_sc", _pd), ("qq_set", _pu8), ("qq_soft", _pu8),
        ("qw_A", _pd), ("qw_B", _pd), ("qw_el", _pd), ("qw_sc", _pd),
        ("n_solv_int", C.c_int32), ("solv_int_ij", _pi), ("solv_int_par", _pd),
    ]


class QnbParams(C.Structure):
    _fields_ = [
        ("Rcpp", C.c_double), ("Rcpw", C.c_double), ("Rcww", C.c_double), ("Rcq", C.c_double),
        ("RcLRF", C.c_double), ("use_LRF", C.c_int32), ("precision", C.c_int32), ("ref_quirks", C.c_int32),
        ("xpcent", C.c_double * 3), ("xwcent", C.c_double * 3), ("rexcl_o", C.c_double),
        ("lambda_", C.c_double * QNB_MAX_STATES),
        ("rwat", C.c_double), ("fk_wsphere", C.c_double), ("Dwmz", C.c_double), ("awmz", C.c_double),
        ("Boltz", C.c_double),
        ("wpol_restr", C.c_int32), ("nwpolr_shell", C.c_int32), ("fkwpol", C.c_double), ("itdis", C.c_double),
        ("wshell_rout", C.c_double * QNB_MAX_SHELLS), ("wshell_dr", C.c_double * QNB_MAX_SHELLS),
        ("wshell_cstb", C.c_double * QNB_MAX_SHELLS), ("wshell_theta_corr", C.c_double * QNB_MAX_SHELLS),
    ]


class QnbEnergies(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "solvent_radial", "water_pol")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class QnbQEnergies(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("qq_el", "qq_vdw", "qp_el", "qp_vdw", "qw_el", "qw_vdw")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class QnbMdParams(C.Structure):
    _fields_ = [
        ("dt_fs", C.c_double), ("T0", C.c_double), ("tau_T_fs", C.c_double),
        ("nbcycle", C.c_int32), ("ene_every", C.c_int32), ("max_frames", C.c_int32),
        ("nbonds", C.c_int32), ("nangles", C.c_int32),
        ("bond_ij", _pi), ("bond_r0", _pd), ("bond_k", _pd),
        ("angle_ijk", _pi), ("angle_t0", _pd), ("angle_k", _pd),
        ("winv", _pd), ("water_roh", C.c_double), ("water_rhh", C.c_double),
        ("separate_scaling", C.c_int32),
    ]


_pu8 = C.POINTER(C.c_uint8)

QTANGENT_FIELDS = ("qp_dq", "qp_dA", "qp_dB", "qp_is14", "qw_dq", "qw_dA", "qw_dB",
                   "qq_dq_i", "qq_dA_i", "qq_dB_i", "qq_dq_j", "qq_dA_j", "qq_dB_j", "qq_is14")


class QnbQTangent(C.Structure):
    _fields_ = [(n, _pu8 if n.endswith("is14") else _pd) for n in QTANGENT_FIELDS]


QNB_ADJ_NV = 7
ADJ_DQ, ADJ_DA, ADJ_DB, ADJ_DA14, ADJ_DB14, ADJ_EEL, ADJ_EVDW = range(7)

QNB_FRAME_LEN = 104
F_Q0, F_BOND, F_ANGLE, F_EKIN, F_TEMP, F_TFREE, F_STEP, F_OFFD = 9, 57, 58, 59, 60, 61, 62, 63
F_QB0 = 71  # per state s: Q bond, angle, torsion, improper at 71 + 4s
QNB_MAX_OFFD = 4
QNB_QB_MAX_TERMS = 4096
QNB_QB_MAX_BONDS = 1024


class QnbQbonded(C.Structure):
    _fields_ = [
        ("nbond", C.c_int32), ("nangle", C.c_int32), ("ntorsion", C.c_int32), ("nimproper", C.c_int32),
        ("nbond_lib", C.c_int32), ("nangle_lib", C.c_int32), ("ntorsion_lib", C.c_int32),
        ("nimproper_lib", C.c_int32),
        ("bond_ij", _pi), ("bond_cod", _pi), ("bond_lib", _pd),
        ("angle_ijk", _pi), ("angle_cod", _pi), ("angle_lib", _pd),
        ("torsion_ijkl", _pi), ("torsion_cod", _pi), ("torsion_lib", _pd),
        ("improper_ijkl", _pi), ("improper_cod", _pi), ("improper_lib", _pd),
        ("nangle_coupl", C.c_int32), ("ntorsion_coupl", C.c_int32), ("nimproper_coupl", C.c_int32),
        ("angle_coupl", _pi), ("torsion_coupl", _pi), ("improper_coupl", _pi),
        ("urey_bradley", C.c_int32),
    ]


# (struct prefix, term-atom count, library width) of the four Q bonded kinds
QB_KINDS = (("bond", 2, 4), ("angle", 3, 4), ("torsion", 4, 3), ("improper", 4, 2))
QB_IDX = {"bond": "bond_ij", "angle": "angle_ijk", "torsion": "torsion_ijkl", "improper": "improper_ijkl"}


def qbonded_struct(tb, urey=False):
    """qnb_qbonded from a table dict: <kind>_ij/_ijk/_ijkl (0-based atoms),
    <kind>_cod (n, nstates) 1-based library rows (0 = absent), <kind>_lib
    rows, <kind>_coupl (n, 3) term, bond, flag.  Returns (struct, arrays to
    keep alive while the struct is in use)."""
    import numpy as np
    q = QnbQbonded()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(_pi if dt == np.int32 else _pd)

    for kind, nat, nlib in QB_KINDS:
        idx = np.asarray(tb.get(QB_IDX[kind], np.zeros((0, nat))), np.int32).reshape(-1, nat)
        n = len(idx)
        lib = np.asarray(tb.get(kind + "_lib", np.zeros((0, nlib))), float).reshape(-1, max(nlib, 1))[:, :nlib]
        setattr(q, "n" + kind, n)
        setattr(q, "n" + kind + "_lib", len(lib))
        setattr(q, QB_IDX[kind], arr(idx, np.int32))
        setattr(q, kind + "_cod", arr(np.asarray(tb.get(kind + "_cod", np.zeros((n, 1))), np.int32), np.int32))
        setattr(q, kind + "_lib", arr(lib, np.float64))
        if kind != "bond":
            c = np.asarray(tb.get(kind + "_coupl", np.zeros((0, 3))), np.int32).reshape(-1, 3)
            setattr(q, "n" + kind + "_coupl", len(c))
            setattr(q, kind + "_coupl", arr(c, np.int32))
    q.urey_bradley = 1 if urey else 0
    return q, keep

# group contributions: array types and values per (array, state)
GC_FULL, GC_ELECTRO, GC_VDW = 1, 2, 3
GC_NV = 7


class QnbFepParams(C.Structure):
    _fields_ = [("kT", C.c_double), ("nskip", C.c_int32), ("nbins", C.c_int32), ("nmin", C.c_int32),
                ("reserved", C.c_int32), ("alpha", C.c_double * QNB_MAX_STATES),
                ("coeff", C.c_double * QNB_MAX_STATES), ("A", C.c_double), ("mu", C.c_double),
                ("eta", C.c_double), ("r0", C.c_double)]


class QnbFepOut(C.Structure):
    _fields_ = [("dgf", _pd), ("dgr", _pd), ("dG", _pd), ("gap_range", _pd), ("dGg", _pd), ("dGv", _pd),
                ("c1", _pd), ("c2", _pd), ("r", _pd), ("npts", _pi), ("bin_dGg", _pd), ("bin_c1", _pd),
                ("bin_c2", _pd), ("bin_r", _pd), ("ptsum", _pi)]


ABI_VERSION = 6

# every symbol include/qnb.h declares (checked by the CPU test suite)
EXPORTS = [
    "qnb_abi_version", "qnb_last_error", "qnb_device_count", "qnb_create", "qnb_destroy",
    "qnb_make_pair_lists", "qnb_nonbonded", "qnb_qq", "qnb_restraints", "qnb_eval_terms",
    "qnb_get_pairs", "qnb_get_lrf", "qnb_get_shell_state", "qnb_set_shell_state",
    "qnb_batch_create", "qnb_batch_destroy", "qnb_batch_run", "qnb_batch_integrate", "qnb_batch_thermostat", "qnb_batch_sync", "qnb_batch_get_state",
    "qnb_batch_energy_history", "qnb_batch_pair_counts", "qnb_batch_time_kernel", "qnb_batch_minimize", "qnb_batch_autotune",
    "qnb_adjoint", "qnb_adjoint_static", "qnb_batch_store_frames", "qnb_batch_frames", "qnb_batch_adjoint", "qnb_batch_set_offdiag",
    "qnb_set_qbonded", "qnb_eval_qbonded", "qnb_batch_set_qbonded", "qnb_batch_free_energy", "qnb_free_energy",
    "qnb_set_groups", "qnb_calculate_exc", "qnb_batch_set_groups", "qnb_batch_group_energies",
]

_lib = None


def _declare(lib):
    vp = C.c_void_p
    lib.qnb_abi_version.restype = C.c_int32
    lib.qnb_last_error.restype = C.c_char_p
    lib.qnb_device_count.restype = C.c_int32
    lib.qnb_create.argtypes = [C.POINTER(QnbSystem), C.POINTER(QnbParams), C.POINTER(vp)]
    lib.qnb_destroy.argtypes = [vp]
    lib.qnb_make_pair_lists.argtypes = [vp, _pd]
    lib.qnb_nonbonded.argtypes = [vp, _pd, _pd, C.POINTER(QnbEnergies), C.POINTER(QnbQEnergies)]
    lib.qnb_qq.argtypes = [vp, _pd, _pd, C.POINTER(QnbQEnergies)]
    lib.qnb_restraints.argtypes = [vp, _pd, _pd, C.c_double, C.c_int32, C.POINTER(QnbEnergies)]
    lib.qnb_eval_terms.argtypes = [vp, C.c_uint32, _pd, _pd, C.c_double, C.c_int32,
                                   C.POINTER(QnbEnergies), C.POINTER(QnbQEnergies)]
    lib.qnb_get_pairs.argtypes = [vp, C.c_int32, _pi, C.c_int64, C.POINTER(C.c_int64)]
    lib.qnb_get_lrf.argtypes = [vp, _pd]
    lib.qnb_get_shell_state.argtypes = [vp, _pd]
    lib.qnb_set_shell_state.argtypes = [vp, _pd]
    pp = C.POINTER(C.c_void_p)
    lib.qnb_batch_create.argtypes = [C.c_int32, pp, pp, pp, pp, pp, C.POINTER(vp)]
    lib.qnb_batch_destroy.argtypes = [vp]
    lib.qnb_batch_run.argtypes = [vp, C.c_int32, C.c_int32]
    lib.qnb_batch_integrate.argtypes = [vp, pp]
    lib.qnb_batch_thermostat.argtypes = [vp, C.c_int32, _pd]
    lib.qnb_batch_sync.argtypes = [vp]
    lib.qnb_batch_get_state.argtypes = [vp, C.c_int32, _pd, _pd]
    lib.qnb_batch_energy_history.argtypes = [vp, C.c_int32, _pd, C.c_int32, C.POINTER(C.c_int32)]
    lib.qnb_batch_pair_counts.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
    lib.qnb_batch_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    lib.qnb_batch_minimize.argtypes = [vp, C.c_int32, C.c_double]
    lib.qnb_batch_autotune.argtypes = [vp, C.c_int32, _pi]
    lib.qnb_adjoint.argtypes = [vp, C.POINTER(QnbQTangent), C.c_int32, _pd, _pd]
    lib.qnb_adjoint_static.argtypes = [vp, C.POINTER(QnbQTangent), C.c_int32, _pd, _pd]
    lib.qnb_batch_store_frames.argtypes = [vp, C.c_int32, C.c_int32]
    lib.qnb_batch_set_offdiag.argtypes = [vp, C.c_int32, C.c_int32, _pi, _pd, _pd]
    lib.qnb_batch_frames.argtypes = [vp, C.c_int32, _pd, C.c_int32, C.POINTER(C.c_int32)]
    lib.qnb_batch_adjoint.argtypes = [vp, pp, _pd, C.c_int64, C.POINTER(C.c_int64)]
    lib.qnb_batch_free_energy.argtypes = [vp, C.POINTER(QnbFepParams), C.POINTER(QnbFepOut)]
    lib.qnb_free_energy.argtypes = [C.c_int32, C.c_int32, _pi, C.POINTER(_pd), C.POINTER(_pd), _pd,
                                    C.POINTER(QnbFepParams), C.POINTER(QnbFepOut)]
    lib.qnb_set_groups.argtypes = [vp, C.POINTER(QnbSystem), C.c_int32, C.POINTER(C.c_uint8), _pi]
    lib.qnb_calculate_exc.argtypes = [vp, _pd, _pd, _pd]
    lib.qnb_batch_set_groups.argtypes = [vp, C.c_int32, C.POINTER(QnbSystem), C.c_int32, C.POINTER(C.c_uint8), _pi]
    lib.qnb_batch_group_energies.argtypes = [vp, C.c_int32, _pd, C.c_int32, C.POINTER(C.c_int32)]
    lib.qnb_set_qbonded.argtypes = [vp, C.POINTER(QnbQbonded)]
    lib.qnb_eval_qbonded.argtypes = [vp, _pd, _pd, _pd]
    lib.qnb_batch_set_qbonded.argtypes = [vp, C.c_int32, C.POINTER(QnbQbonded)]
    for n in EXPORTS:
        if n not in ("qnb_abi_version", "qnb_last_error", "qnb_device_count"):
            getattr(lib, n).restype = C.c_int32
    return lib


def load(path=None):
    """Load the engine library; raises when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP engine library not built: {p} (run __graft_entry__.build())")
    lib = _declare(C.CDLL(p))
    if lib.qnb_abi_version() != ABI_VERSION:
        raise RuntimeError("qnb ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


class QnbError(RuntimeError):
    pass


def check(lib, rc, what):
    if rc != 0:
        msg = lib.qnb_last_error()
        raise QnbError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
