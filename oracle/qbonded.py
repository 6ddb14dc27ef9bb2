"""CPU restatement of the Q-atom bonded terms per state (test infrastructure).

Imported only by tests/, as the checker of the device kernel k_qbonded.
Follows, term by term and in the reference's loop order:

  * pot_energy's per-state loop, src/potene.f90:183-190: for each state,
    qbond, qangle, (qurey_bradley for CHARMM force fields), qtorsion,
    qimproper;
  * qbond src/bondene.f90:506-551 (Morse + harmonic; the Morse coupling state
    EMorseD / dMorse_i / dMorse_j of each bond is rewritten when the bond is
    present in the state, and kept from the last evaluation otherwise);
  * qangle :387-446 (coupling: the FIRST matching angle coupling, `exit`),
    qurey_bradley :450-502 (the LAST matching coupling), qtorsion :624-685 and
    qimproper :555-620 (the LAST matching coupling);
  * geometry: bond_calc, angle_calc, torsion_calc, improper_calc,
    src/bonded.f90:53-62, 78-118, 151-238, 240-320;
  * EMorseD = dMorse = 0 at set-up, src/qatom.f90:1000-1004.

Pinned against tests/golden/lig_qb2.npz (the reference's own routines on the
reference ligand) and, independently, against central finite differences of
its own energy (tests/test_oracle_qbonded.py).

Data (0-based atom and term indices; codes 1-based, 0 = absent in the state):
  bond_ij (nb, 2), bond_cod (nb, ns), bond_lib (:, 4) = Dmz, amz, r0, fk
  angle_ijk (na, 3), angle_cod, angle_lib (:, 4) = fk, ang0, ureyfk, ureyr0
  torsion_ijkl (nt, 4), torsion_cod, torsion_lib (:, 3) = fk, rmult, deltor
  improper_ijkl (ni, 4), improper_cod, improper_lib (:, 2) = fk, imp0
  angle_coupl / torsion_coupl / improper_coupl (n, 3) = term, bond, flag
  (flag 1: the term scales with 1 - EMorseD instead of EMorseD).
"""
import math

import numpy as np

QREAL_EPS = 1.0e-10  # src/sizes.f90:50 (QDOUBLE)


def _bond_calc(a, b):
    v = b - a
    return v, math.sqrt(float(v @ v))


def _angle_calc(a, b, c):
    ab = a - b
    bc = c - b
    ab2 = 1.0 / float(ab @ ab)
    bc2 = 1.0 / float(bc @ bc)
    abd = math.sqrt(ab2)
    bcd = math.sqrt(bc2)
    sc = float(ab @ bc) * abd * bcd
    sc = min(1.0, max(-1.0, sc))
    ang = math.acos(sc)
    inv = math.sin(ang)
    if abs(inv) < QREAL_EPS:
        inv = QREAL_EPS
    inv = -1.0 / inv
    af = inv * ((bc * abd * bcd) - (ab2 * sc * ab))
    cf = inv * ((ab * abd * bcd) - (bc2 * sc * bc))
    return ang, af, af + cf, cf


def _torsion_calc(a, b, c, d):
    ab = a - b
    bc = c - b
    cd = d - c
    cabc = np.cross(ab, bc)
    cbcd = np.cross(-bc, cd)
    a2 = float(cabc @ cabc)
    b2 = float(cbcd @ cbcd)
    aa = math.sqrt(a2)
    bb = math.sqrt(b2)
    sc = float(cabc @ cbcd) / (aa * bb)
    sc = min(1.0, max(-1.0, sc))
    ang = math.acos(sc)
    if float(bc @ np.cross(cabc, cbcd)) < 0.0:
        ang = -ang
    inv = math.sin(ang)
    if abs(inv) < QREAL_EPS:
        inv = QREAL_EPS
    inv = -1.0 / inv
    v1 = ((cbcd / (aa * bb)) - ((cabc / a2) * sc)) * inv
    v2 = ((cabc / (aa * bb)) - ((cbcd / b2) * sc)) * inv
    ca = ab - bc
    db = -bc - cd
    return (ang, np.cross(bc, v1), np.cross(ca, v1) + np.cross(cd, v2), np.cross(db, v2) - np.cross(ab, v1),
            np.cross(bc, v2))


def _coupling(table, term, first):
    """(bond, flag) of a term's Morse coupling, or None."""
    hit = None
    for t, b, f in np.asarray(table, int).reshape(-1, 3):
        if t == term:
            hit = (int(b), int(f))
            if first:
                break
    return hit


def new_state(nbond):
    return np.zeros(nbond), np.zeros((nbond, 3))


def qbonded(x, tb, lam, state, urey=False):
    """One evaluation of the per-state loop at x (natom, 3).  Returns
    E (ns, 4) = bond, angle (+ Urey-Bradley), torsion, improper per state,
    the gradient d (natom, 3) (d = dE/dx, each state's forces weighted by its
    lambda) and the Morse coupling state after the call.  `state` =
    (EMorseD (nb,), dMorse_i (nb, 3)) from the previous call."""
    x = np.asarray(x, float).reshape(-1, 3)
    lam = np.asarray(lam, float)
    ns = len(lam)
    emd = np.array(state[0], float, copy=True)
    dmi = np.array(state[1], float, copy=True).reshape(-1, 3)
    d = np.zeros_like(x)
    E = np.zeros((ns, 4))
    bij = np.asarray(tb.get("bond_ij", np.zeros((0, 2))), int).reshape(-1, 2)
    aijk = np.asarray(tb.get("angle_ijk", np.zeros((0, 3))), int).reshape(-1, 3)
    tijkl = np.asarray(tb.get("torsion_ijkl", np.zeros((0, 4))), int).reshape(-1, 4)
    iijkl = np.asarray(tb.get("improper_ijkl", np.zeros((0, 4))), int).reshape(-1, 4)
    acp = tb.get("angle_coupl", np.zeros((0, 3)))
    tcp = tb.get("torsion_coupl", np.zeros((0, 3)))
    icp = tb.get("improper_coupl", np.zeros((0, 3)))

    def couple(cp, pe):
        b, _ = cp
        i, j = bij[b]
        d[i] += dmi[b] * pe
        d[j] += -dmi[b] * pe

    def gamma_of(cp):
        if cp is None:
            return 1.0
        g = emd[cp[0]]
        return 1.0 - g if cp[1] == 1 else g

    for s in range(ns):
        L = lam[s]
        # qbond
        for ib, (i, j) in enumerate(bij):
            ic = int(tb["bond_cod"][ib][s])
            if ic <= 0:
                continue
            Dmz, amz, r0, fk = tb["bond_lib"][ic - 1]
            av, dist = _bond_calc(x[i], x[j])
            db = dist - r0
            fexp = math.exp(-amz * db)
            E[s, 0] += Dmz * (fexp * fexp - 2.0 * fexp) + 0.5 * fk * db ** 2
            dv = (2.0 * Dmz * amz * (fexp - fexp * fexp) + fk * db) * L / dist
            d[i] -= av * dv
            d[j] += av * dv
            if db > 0.0:
                emd[ib] = -(fexp * fexp - 2.0 * fexp)
                dmi[ib] = 2.0 * amz * (fexp - fexp * fexp) * L / dist * av
            else:
                emd[ib] = 1.0
                dmi[ib] = 0.0
        # qangle
        for ia, (i, j, k) in enumerate(aijk):
            ic = int(tb["angle_cod"][ia][s])
            if ic <= 0:
                continue
            fk, ang0 = tb["angle_lib"][ic - 1][:2]
            cp = _coupling(acp, ia, first=True)
            g = gamma_of(cp)
            ang, av, bv, cv = _angle_calc(x[i], x[j], x[k])
            da = ang - ang0
            ae = 0.5 * fk * da ** 2
            E[s, 1] += ae * g
            dv = g * fk * da * L
            d[i] += av * dv
            d[j] -= bv * dv
            d[k] += cv * dv
            if cp is not None:
                couple(cp, ae)
        # qurey_bradley
        if urey:
            for ia, (i, j, k) in enumerate(aijk):
                ic = int(tb["angle_cod"][ia][s])
                if ic <= 0 or tb["angle_lib"][ic - 1][2] == 0.0:
                    continue
                ufk, ur0 = tb["angle_lib"][ic - 1][2:4]
                cp = _coupling(acp, ia, first=False)
                g = gamma_of(cp)
                av, dist = _bond_calc(x[i], x[k])
                ru = dist - ur0
                eu = ufk * ru ** 2
                E[s, 1] += eu * g
                du = g * 2.0 * (ufk * ru / dist) * L
                d[i] -= av * du
                d[k] += av * du
                if cp is not None:
                    couple(cp, eu)
        # qtorsion
        for it, (i, j, k, l) in enumerate(tijkl):
            ic = int(tb["torsion_cod"][it][s])
            if ic <= 0:
                continue
            fk, rm, dt = tb["torsion_lib"][ic - 1][:3]
            cp = _coupling(tcp, it, first=False)
            g = gamma_of(cp)
            ang, av, bv, cv, dvv = _torsion_calc(x[i], x[j], x[k], x[l])
            arg = rm * ang - dt
            pe = fk * (1.0 + math.cos(arg))
            E[s, 2] += pe * g
            dv = -rm * fk * math.sin(arg) * g * L
            d[i] += av * dv
            d[j] += bv * dv
            d[k] += cv * dv
            d[l] += dvv * dv
            if cp is not None:
                couple(cp, pe)
        # qimproper
        for ip, (i, j, k, l) in enumerate(iijkl):
            ic = int(tb["improper_cod"][ip][s])
            if ic <= 0:
                continue
            fk, imp0 = tb["improper_lib"][ic - 1][:2]
            cp = _coupling(icp, ip, first=False)
            g = gamma_of(cp)
            ang, av, bv, cv, dvv = _torsion_calc(x[i], x[j], x[k], x[l])
            arg = ang - imp0
            arg = arg - 2.0 * math.pi * _nint(arg / (2.0 * math.pi))
            dv = fk * arg
            pe = 0.5 * dv * arg
            E[s, 3] += pe * g
            dv = dv * g * L
            d[i] += av * dv
            d[j] += bv * dv
            d[k] += cv * dv
            d[l] += dvv * dv
            if cp is not None:
                couple(cp, pe)
    return E, d, (emd, dmi)


def _nint(v):
    """Fortran NINT: nearest integer, halves away from zero."""
    return math.floor(v + 0.5) if v >= 0.0 else -math.floor(-v + 0.5)


def from_dump(z):
    """Tables in the layout above from a qnb_dump record set (Fortran 1-based
    atom and term numbers, arrays with the per-term record along the last
    axis)."""
    ns = int(z["nstates"])
    tb = {}
    if "qbnd" in z:
        q = np.asarray(z["qbnd"]).T
        tb.update(bond_ij=q[:, :2] - 1, bond_cod=q[:, 2:2 + ns], bond_lib=np.asarray(z["qbondlib"]).T)
    if "qang" in z:
        q = np.asarray(z["qang"]).T
        tb.update(angle_ijk=q[:, :3] - 1, angle_cod=q[:, 3:3 + ns], angle_lib=np.asarray(z["qanglib"]).T)
    if "qtor" in z:
        q = np.asarray(z["qtor"]).T
        tb.update(torsion_ijkl=q[:, :4] - 1, torsion_cod=q[:, 4:4 + ns], torsion_lib=np.asarray(z["qtorlib"]).T)
    if "qimp" in z:
        q = np.asarray(z["qimp"]).T
        tb.update(improper_ijkl=q[:, :4] - 1, improper_cod=q[:, 4:4 + ns], improper_lib=np.asarray(z["qimplib"]).T)
    for key, name in (("iang_coupl", "angle_coupl"), ("itor_coupl", "torsion_coupl"),
                      ("iimp_coupl", "improper_coupl")):
        if key in z:
            c = np.asarray(z[key]).T.copy()
            c[:, :2] -= 1
            tb[name] = c
    return tb
