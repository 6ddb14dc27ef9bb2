"""Host-side mirror of the reference's nonbonded interface over the HIP engine.

`NonbondedEngine` exposes the entry points Qdyn6's MD loop calls in the
reference's NONBONDENE/POTENE modules (make_pair_lists, pot_energy_nonbonds,
nonbond_qq/qqp, restrain_solvent/watpol), implemented by the C ABI of
q6_amd/libq6nb.so.  There is no CPU path: construction fails when the library
or a GPU is missing.
"""
import ctypes as C

import numpy as np

from . import abi

_pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731


class NonbondedEngine:
    def __init__(self, system, precision=abi.QNB_PREC_FP64, **par_overrides):
        self.lib = abi.load()
        self.sys = system
        self._s, self._p = system.to_structs(precision=precision, **par_overrides)
        self._p.precision = precision
        h = C.c_void_p()
        abi.check(self.lib, self.lib.qnb_create(C.byref(self._s), C.byref(self._p), C.byref(h)), "qnb_create")
        self.h = h
        self.nst = int(system.n["nstates"])

    def close(self):
        if getattr(self, "h", None):
            self.lib.qnb_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # group contributions (calculate_exc, src/nonbondene.f90:6825-6920)
    def set_groups(self, masks, types):
        """Energy arrays of the group contribution: masks (narr, nat_solute)
        0/1 over the solute atoms, types (narr,) abi.GC_FULL / GC_ELECTRO /
        GC_VDW; an empty list removes them."""
        masks = np.ascontiguousarray(np.asarray(masks, np.uint8).reshape(len(types), self._s.nat_solute))
        types = np.ascontiguousarray(types, np.int32)
        self._gc = (masks, types)
        abi.check(self.lib, self.lib.qnb_set_groups(self.h, C.byref(self._s), len(types),
                                                     masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                     types.ctypes.data_as(C.POINTER(C.c_int32))), "qnb_set_groups")

    def calculate_exc(self, x, eq_in):
        """eq_in (nstates, 11): qq el, vdW, qp el, vdW, qw el, vdW, Q bond,
        angle, torsion, improper, restraint per state.  Returns (narr,
        nstates, 7): qq el, vdW, qp el, vdW, qx el, vdW, total."""
        x = np.ascontiguousarray(x, np.float64)
        eq_in = np.ascontiguousarray(eq_in, np.float64)
        narr = len(self._gc[1]) if getattr(self, "_gc", None) else 0
        out = np.zeros((narr, self.nst, abi.GC_NV))
        abi.check(self.lib, self.lib.qnb_calculate_exc(self.h, _pd(x), _pd(eq_in), _pd(out)), "qnb_calculate_exc")
        return out

    # Q-atom bonded terms (qatom_load_fep's tables, src/qatom.f90:952-1800)
    def set_qbonded(self, tables, urey=False):
        """Loads the Q bonded terms (abi.qbonded_struct layout; None removes
        them) and zeroes their Morse coupling state."""
        if tables is None:
            abi.check(self.lib, self.lib.qnb_set_qbonded(self.h, None), "qnb_set_qbonded")
            return
        q, keep = abi.qbonded_struct(tables, urey)
        abi.check(self.lib, self.lib.qnb_set_qbonded(self.h, C.byref(q)), "qnb_set_qbonded")

    # the per-state Q bonded loop of pot_energy (src/potene.f90:183-190)
    def qbonded(self, x, d=None):
        """Returns (E, d): E (nstates, 4) = EQ(s)%q%bond, %angle, %torsion,
        %improper; the lambda-weighted gradient is added into d."""
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        e = np.zeros(4 * self.nst)
        abi.check(self.lib, self.lib.qnb_eval_qbonded(self.h, _pd(x), _pd(d), _pd(e)), "qnb_eval_qbonded")
        return e.reshape(self.nst, 4), d

    # make_pair_lists (src/nonbondene.f90:749-837)
    def make_pair_lists(self, x):
        x = np.ascontiguousarray(x, np.float64)
        abi.check(self.lib, self.lib.qnb_make_pair_lists(self.h, _pd(x)), "qnb_make_pair_lists")

    def eval_terms(self, mask, x, Tfree=0.0, istep=0, d=None):
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        E = abi.QnbEnergies()
        EQ = (abi.QnbQEnergies * self.nst)()
        abi.check(self.lib, self.lib.qnb_eval_terms(self.h, mask, _pd(x), _pd(d), Tfree, istep, C.byref(E), EQ),
                  "qnb_eval_terms")
        return d, E.as_dict(), [q.as_dict() for q in EQ]

    # pot_energy_nonbonds (src/potene.f90:320-380)
    def pot_energy_nonbonds(self, x, d=None):
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        E = abi.QnbEnergies()
        EQ = (abi.QnbQEnergies * self.nst)()
        abi.check(self.lib, self.lib.qnb_nonbonded(self.h, _pd(x), _pd(d), C.byref(E), EQ), "qnb_nonbonded")
        return d, E.as_dict(), [q.as_dict() for q in EQ]

    # nonbond_qq + nonbond_qqp (src/nonbondene.f90:5013-5153)
    def nonbond_qq(self, x, d=None, EQ=None):
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        eq = (abi.QnbQEnergies * self.nst)()
        if EQ is not None:
            for k, q in enumerate(EQ):
                for n, v in q.items():
                    setattr(eq[k], n, v)
        abi.check(self.lib, self.lib.qnb_qq(self.h, _pd(x), _pd(d), eq), "qnb_qq")
        return d, [q.as_dict() for q in eq]

    # restrain_solvent + watpol (src/nonbondene.f90:6466-6746)
    def restraints(self, x, Tfree, istep, d=None):
        x = np.ascontiguousarray(x, np.float64)
        if d is None:
            d = np.zeros_like(x)
        E = abi.QnbEnergies()
        abi.check(self.lib, self.lib.qnb_restraints(self.h, _pd(x), _pd(d), Tfree, istep, C.byref(E)),
                  "qnb_restraints")
        return d, E.as_dict()

    def pairs(self, which):
        n = C.c_int64()
        abi.check(self.lib, self.lib.qnb_get_pairs(self.h, which, None, 0, C.byref(n)), "qnb_get_pairs")
        ab = np.zeros((max(n.value, 1), 2), np.int32)
        abi.check(self.lib, self.lib.qnb_get_pairs(self.h, which, ab.ctypes.data_as(C.POINTER(C.c_int32)), n.value,
                                                   C.byref(n)), "qnb_get_pairs")
        return ab[:n.value]

    def lrf(self):
        out = np.zeros((int(self.sys.n["ncgp"]), 43))
        abi.check(self.lib, self.lib.qnb_get_lrf(self.h, _pd(out)), "qnb_get_lrf")
        return out

    def shell_state(self):
        n = int(self.sys.par["nwpolr_shell"])
        st = np.zeros(3 * max(n, 1))
        abi.check(self.lib, self.lib.qnb_get_shell_state(self.h, _pd(st)), "qnb_get_shell_state")
        return st[:3 * n]

    def set_shell_state(self, st):
        st = np.ascontiguousarray(st, np.float64)
        abi.check(self.lib, self.lib.qnb_set_shell_state(self.h, _pd(st)), "qnb_set_shell_state")

    def adjoint(self, tangent, frames):
        """Adjoint of the per-state Q-atom energies over frames (nframes,
        natom, 3): (nframes, nstates, nqat, QNB_ADJ_NV) records (qnb_adjoint)."""
        fr = np.ascontiguousarray(frames, np.float64)
        nf = fr.shape[0]
        nq = int(self.sys.n["nqat"])
        out = np.zeros((nf, self.nst, nq, abi.QNB_ADJ_NV))
        abi.check(self.lib, self.lib.qnb_adjoint(self.h, C.byref(tangent), nf, _pd(fr), _pd(out)), "qnb_adjoint")
        return out

    def adjoint_static(self, tangent, frames):
        """As adjoint(), through the entry point that allocates no device
        memory (qnb_adjoint_static: a fixed per-GPU workspace)."""
        fr = np.ascontiguousarray(frames, np.float64)
        nf = fr.shape[0]
        nq = int(self.sys.n["nqat"])
        out = np.zeros((nf, self.nst, nq, abi.QNB_ADJ_NV))
        abi.check(self.lib, self.lib.qnb_adjoint_static(self.h, C.byref(tangent), nf, _pd(fr), _pd(out)),
                  "qnb_adjoint_static")
        return out


def pot_energy_summary(E, EQ, lambdas, bonded=None, EQ_bonded=None, EQ_restraint=None):
    """The energy summary of pot_energy (src/potene.f90:224-251) in its
    order of operations.

    E: class energies of the nonbonded engine (pp/pw/ww el and vdw, lrf,
    solvent_radial, water_pol); EQ[s]: per-state qq/qp/qw el and vdw (qp
    includes the Q-neighbour pairs); lambdas[s].  The terms the engine does
    not compute come from the caller: bonded = dict of p_bond, w_bond,
    p_angle, w_angle, p_torsion, p_improper, restraint_fix,
    restraint_shell and restraint_protein (the position/distance restraint
    sum p_restrain leaves before the states' restraints are added); EQ_bonded[s] = (bond, angle, torsion, improper) of the
    Q-atom bonded terms; EQ_restraint[s] = the state's restraint energy.
    Returns (Esum, EQsum): Esum holds q_bond..q_improper, qx_el, qx_vdw,
    restraint_protein, restraint_total and potential; EQsum[s] holds qx_el,
    qx_vdw and total."""
    bonded = bonded or {}
    ns = len(EQ)
    qb = [(0.0, 0.0, 0.0, 0.0)] * ns if EQ_bonded is None else [tuple(map(float, b)) for b in EQ_bonded]
    qr = [0.0] * ns if EQ_restraint is None else [float(r) for r in EQ_restraint]
    out = dict(q_bond=0.0, q_angle=0.0, q_torsion=0.0, q_improper=0.0, qx_el=0.0, qx_vdw=0.0,
               restraint_protein=float(bonded.get("restraint_protein", 0.0)))
    eqs = []
    for s in range(ns):
        q, lam = EQ[s], float(lambdas[s])
        qx_el = q["qq_el"] + q["qp_el"] + q["qw_el"]
        qx_vdw = q["qq_vdw"] + q["qp_vdw"] + q["qw_vdw"]
        total = qb[s][0] + qb[s][1] + qb[s][2] + qb[s][3] + qx_el + qx_vdw + qr[s]
        eqs.append(dict(qx_el=qx_el, qx_vdw=qx_vdw, total=total))
        out["q_bond"] = out["q_bond"] + qb[s][0] * lam
        out["q_angle"] = out["q_angle"] + qb[s][1] * lam
        out["q_torsion"] = out["q_torsion"] + qb[s][2] * lam
        out["q_improper"] = out["q_improper"] + qb[s][3] * lam
        out["qx_el"] = out["qx_el"] + qx_el * lam
        out["qx_vdw"] = out["qx_vdw"] + qx_vdw * lam
        out["restraint_protein"] = out["restraint_protein"] + qr[s] * lam
    g = lambda k: float(bonded.get(k, 0.0))  # noqa: E731
    out["restraint_total"] = (g("restraint_fix") + g("restraint_shell") + out["restraint_protein"]
                              + E["solvent_radial"] + E["water_pol"])
    out["potential"] = (g("p_bond") + g("w_bond") + g("p_angle") + g("w_angle") + g("p_torsion")
                        + g("p_improper") + E["pp_el"] + E["pp_vdw"] + E["pw_el"] + E["pw_vdw"] + E["ww_el"]
                        + E["ww_vdw"] + out["q_bond"] + out["q_angle"] + out["q_torsion"] + out["q_improper"]
                        + out["qx_el"] + out["qx_vdw"] + out["restraint_total"] + E["lrf"])
    return out, eqs
