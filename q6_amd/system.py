"""Host-side description of one spherical system for the nonbonded engine.

`QSystem` holds, as numpy arrays in the C layouts documented in include/qnb.h,
everything Qdyn6's set-up hands to the nonbonded path (topology, FEP tables,
sphere and SCAAS constants).  It is filled either from a golden dump of the
reference (`from_golden`) or by the synthetic builder in `q6_amd.synth`, and
turned into the ABI structs by `to_structs`.
"""
import ctypes as C

import numpy as np

from . import abi

_INT = np.int32
_U8 = np.uint8
_F = np.float64


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class QSystem:
    # names of the arrays, with their dtype (the ABI field of the same name)
    ARRAYS = {
        "crg": _F, "iac": _INT, "iqatom": _INT, "excl": _U8, "heavy": _U8,
        "cgp_first": _INT, "cgp_last": _INT, "cgp_iswitch": _INT, "cgpatom": _INT,
        "avdw": _F, "bvdw": _F, "ljcod": _INT,
        "listex": _U8, "list14": _U8, "listexlong": _INT, "list14long": _INT,
        "pw_A": _F, "pw_B": _F, "pw_el": _F, "ww_A": _F, "ww_B": _F, "ww_el": _F,
        "iqseq": _INT, "qconn3": _U8,
        "qp_A": _F, "qp_B": _F, "qp_el": _F, "qp_sc": _F, "qp_set": _U8,
        "qq_A": _F, "qq_B": _F, "qq_el": _F, "qq_sc": _F, "qq_set": _U8, "qq_soft": _U8,
        "qw_A": _F, "qw_B": _F, "qw_el": _F, "qw_sc": _F,
        "solv_int_ij": _INT, "solv_int_par": _F,
    }
    SIZES = ("natom", "nat_solute", "nwat", "solv_atom", "ncgp", "ncgp_solute", "nqat", "nstates",
             "ntypes", "ivdw_rule", "solvent_spc")

    def __init__(self):
        self.a = {}          # arrays
        self.n = {}          # sizes
        self.el14_scale = 0.5
        self.par = {}        # run parameters (qnb_params fields)
        self.x = None        # (natom, 3) coordinates
        self.v = None        # (natom, 3) velocities
        self.winv = None     # 1/mass
        self.Tfree = 0.0
        self.name = ""
        self.ref = {}        # expected values when the system comes from a golden dump

    # ------------------------------------------------------------ structs
    def finalize(self):
        """Fill empty optional arrays, fix dtypes and contiguity."""
        for k, dt in self.ARRAYS.items():
            v = self.a.get(k)
            if v is None or np.asarray(v).size == 0:
                v = np.zeros(1, dt)
            self.a[k] = np.ascontiguousarray(np.asarray(v).ravel(), dtype=dt)
        self.x = np.ascontiguousarray(self.x, dtype=_F).reshape(-1, 3)
        return self

    def to_structs(self, precision=abi.QNB_PREC_FP64, **overrides):
        self.finalize()
        s = abi.QnbSystem()
        for k in self.SIZES:
            setattr(s, k, int(self.n[k]))
        s.nexlong = int(self.n.get("nexlong", 0))
        s.n14long = int(self.n.get("n14long", 0))
        s.any_atom_cutoff = int(self.n.get("any_atom_cutoff", 0))
        s.n_solv_int = int(self.n.get("n_solv_int", 0))
        s.el14_scale = float(self.el14_scale)
        for k, dt in self.ARRAYS.items():
            ct = {np.float64: C.c_double, np.int32: C.c_int32, np.uint8: C.c_uint8}[np.dtype(dt).type]
            setattr(s, k, _p(self.a[k], ct))
        p = abi.QnbParams()
        par = dict(self.par)
        par.update(overrides)
        for k in ("Rcpp", "Rcpw", "Rcww", "Rcq", "RcLRF", "rexcl_o", "rwat", "fk_wsphere", "Dwmz", "awmz",
                  "Boltz", "fkwpol", "itdis"):
            setattr(p, k, float(par[k]))
        p.use_LRF = int(par["use_LRF"])
        p.precision = int(par.get("precision", precision))
        p.ref_quirks = int(par.get("ref_quirks", 0))
        p.wpol_restr = int(par["wpol_restr"])
        p.nwpolr_shell = int(par["nwpolr_shell"])
        for k in range(3):
            p.xpcent[k] = float(par["xpcent"][k])
            p.xwcent[k] = float(par["xwcent"][k])
        lam = np.atleast_1d(par["lambda"])
        for k in range(len(lam)):
            p.lambda_[k] = float(lam[k])
        for nm in ("wshell_rout", "wshell_dr", "wshell_cstb", "wshell_theta_corr"):
            vals = np.atleast_1d(par.get(nm, np.zeros(0)))
            arr = getattr(p, nm)
            for k in range(len(vals)):
                arr[k] = float(vals[k])
        return s, p

    # ------------------------------------------------------------ helpers
    @property
    def natom(self):
        return int(self.n["natom"])

    def pair_count_per_step(self, nlist):
        """Pair interactions per step: nbpp + nbpw + nbww + nstates*(nbqp +
        nbqw) + sum_s (nbqq_s + nbqqp_s), the lengths of the reference's
        explicit lists (src/globals.f90:414-437); nlist holds atom-pair counts
        with qq and qqp already summed over states."""
        ns = int(self.n["nstates"])
        return (nlist["pp"] + nlist["pw"] + nlist["ww"] + ns * (nlist["qp"] + nlist["qw"])
                + nlist["qq"] + nlist["qqp"])


def from_golden(path):
    """Build a QSystem from a fixture written by tests/golden/make_golden.py."""
    g = np.load(path)
    s = QSystem()
    s.name = path
    sc = lambda k: g[k].item() if g[k].size == 1 else g[k]  # noqa: E731
    for k in ("natom", "nat_solute", "nwat", "solv_atom", "ncgp", "ncgp_solute", "nqat", "nstates", "ivdw_rule"):
        s.n[k] = int(sc(k))
    s.n["ntypes"] = int(sc("natyps"))
    s.n["solvent_spc"] = 1 if (int(sc("solvent_type")) == 0 and s.n["ivdw_rule"] == 1) else 0
    s.n["any_atom_cutoff"] = 1 if int(sc("iuse_switch_atom")) == 0 else 0
    s.el14_scale = float(sc("el14_scale"))
    nat, nsol = s.n["natom"], s.n["nat_solute"]
    a = s.a
    a["crg"] = g["crg"]
    a["iac"] = g["iac"] - 1
    a["iqatom"] = g["iqatom"]
    a["excl"] = g["excl"]
    a["heavy"] = g["heavy"]
    a["cgp_first"] = g["cgp_first"] - 1
    a["cgp_last"] = g["cgp_last"] - 1
    a["cgp_iswitch"] = g["cgp_iswitch"] - 1
    a["cgpatom"] = g["cgpatom"] - 1
    a["avdw"] = g["avdw"].ravel(order="F")
    a["bvdw"] = g["bvdw"].ravel(order="F")
    # ljcod covers types 1..maxval(iac) (src/simprep.f90:4555-4560); types
    # above carry no atoms and keep the default code 1
    lj = np.ones((s.n["ntypes"], s.n["ntypes"]), g["ljcod"].dtype)
    nl = g["ljcod"].shape[0]
    lj[:nl, :nl] = g["ljcod"]
    a["ljcod"] = lj.ravel(order="F")
    if nsol > 0:
        a["listex"] = g["listex"].ravel(order="F")
        a["list14"] = g["list14"].ravel(order="F")
        if "listexlong" in g.files:
            a["listexlong"] = g["listexlong"].ravel(order="F") - 1
            s.n["nexlong"] = g["listexlong"].shape[1]
        if "list14long" in g.files:
            a["list14long"] = g["list14long"].ravel(order="F") - 1
            s.n["n14long"] = g["list14long"].shape[1]
    for k in ("ww_A", "ww_B", "ww_el", "pw_A", "pw_B", "pw_el"):
        if k in g.files:
            a[k] = g[k].ravel(order="F")
    if s.n["nqat"] > 0:
        a["iqseq"] = g["iqseq"] - 1
        if nsol > 0:
            qc = g["qconn"]  # (ns, nsol, nq)
            a["qconn3"] = (qc <= 3).any(axis=(0, 2)).astype(np.uint8)
        for k in ("qp_A", "qp_B", "qp_el", "qp_sc", "qp_set", "qq_A", "qq_B", "qq_el", "qq_sc", "qq_set",
                  "qq_soft", "qw_A", "qw_B", "qw_el", "qw_sc"):
            if k in g.files:
                a[k] = g[k].ravel(order="F")
    if "solv_int_ij" in g.files:  # nonbnd_solv_int of a topology with solvent torsions
        a["solv_int_ij"] = g["solv_int_ij"].ravel(order="F") - 1
        a["solv_int_par"] = g["solv_int_par"].ravel(order="F")
        s.n["n_solv_int"] = g["solv_int_ij"].shape[1]
    s.x = g["x"].T.copy()
    s.v = g["v"].T.copy()
    s.winv = g["winv"].copy()
    s.Tfree = float(sc("Tfree"))
    p = s.par
    for k in ("Rcpp", "Rcpw", "Rcww", "Rcq", "RcLRF", "rexcl_o", "rwat", "fk_wsphere", "Dwmz", "awmz",
              "Boltz", "fkwpol", "itdis"):
        p[k] = float(sc(k))
    p["use_LRF"] = int(sc("use_LRF"))
    p["wpol_restr"] = int(sc("wpol_restr"))
    p["nwpolr_shell"] = int(sc("nwpolr_shell"))
    p["xpcent"] = g["xpcent"]
    p["xwcent"] = g["xwcent"]
    p["lambda"] = np.atleast_1d(g["lambda"])
    for k in ("wshell_rout", "wshell_dr", "wshell_cstb", "wshell_theta_corr"):
        if k in g.files:
            p[k] = np.atleast_1d(g[k])
    s.ref = {k: g[k] for k in g.files}
    return s.finalize()
