"""Reader of Q's input files: topology (.top), FEP file (.fep) and MD input
(.inp), and the set-up Qdyn6 performs on them before its first step, so that
any Qprep6 system with a spherical boundary reaches the engine without the
reference's own modules.

  * `read_top`: topo_read (src/topo.f90:522-1116).  Records are read the way
    Fortran list-directed input reads them: a count line, then that many
    values over as many lines as they take; the 1-4 and exclusion windows
    are 80 one-digit flags per line ('(80i1)'), the excluded atoms 80 T/F
    flags per line ('(80l1)').
  * `read_fep`: qatom_load_atoms / qatom_load_fep (src/qatom.f90:322-628,
    639-1991), the sections that feed the nonbonded path and the Q-atom
    connectivity: [FEP], [atoms], [change_charges], [atom_types],
    [change_atoms], [soft_pairs], [el_scale], [excluded_pairs],
    [change_bonds], [softcore] and [off_diagonals].  Sections that change
    bonded terms only ([angle_types], [change_angles], ...) are read as raw
    lines; the selection forms of [atoms] (offsets, residue names) are refused.
  * `read_inp`: the keys of src/md.f90:113-1369 this path uses ([MD],
    [cut-offs], [sphere], [solvent], [intervals], [files], [lambdas]), with
    Qdyn6's defaults (src/globals.f90:204-229).
  * `build_system`: the set-up of src/simprep.f90 in Qdyn6's order: get_fep
    (:1000-1048: Q flags, topology charges as default Q charges, sqrt of the
    Q-atom epsilons under the arithmetic rule), topology (:4521-4612: ljcod,
    sqrt of the epsilons under the arithmetic rule), prep_sim (:3595-3725:
    solvent tables from the first solvent molecule, wat_sphere :4613-4703,
    wat_shells :4707-4891, the sqrt(coulomb_constant) charge scaling), the
    solvent internal pairs (:4004-4138), make_qconn (src/nonbondene.f90:
    3087-3170) and precompute_interactions (:2860-2870: pw/qp/qq/qw/ww
    tables, :3321-3589), giving a `QSystem` in the layouts of include/qnb.h.
"""
import math
import os
import re

import numpy as np

from .system import QSystem

MAX_NBR_RANGE = 25          # src/topo.f90:37
NLJTYP = 3                  # src/topo.f90:35
BOLTZ = 0.001986            # src/globals.f90:46
ITDIS_UPDATE = 100          # src/globals.f90:204
WPOLR_LAYER = 3.0001        # src/globals.f90:209
DROUT = 0.5                 # src/globals.f90:210
RC_DEFAULT = 10.0           # rcpp/rcww/rcpw defaults, src/globals.f90:212-214
RCQ_DEFAULT_SPH = 99.0      # src/globals.f90:217
RCLRF_DEFAULT = 99.0        # src/globals.f90:218
FKWPOL_DEFAULT = 20.0       # src/globals.f90:227
FK_WSPHERE_DEFAULT = 60.0   # src/globals.f90:228


class QFileError(ValueError):
    pass


def _num(tok):
    return float(tok.replace("d", "e").replace("D", "E"))


def _lead(line, convs):
    """The leading values of a line read list-directed into len(convs)
    variables; stops at the first token that does not convert (the text
    after the numbers), so optional trailing values come back as missing."""
    out = []
    for tok, cv in zip(line.replace(",", " ").split(), convs):
        try:
            out.append(cv(tok) if cv is not float else _num(tok))
        except ValueError:
            break
    return out


class _Records:
    """Lines of a file with Fortran's record semantics: every read starts on
    a new line and a list-directed read of n values continues over lines."""

    def __init__(self, text):
        self.lines = text.split("\n")
        self.i = 0

    def line(self):
        if self.i >= len(self.lines):
            raise QFileError("unexpected end of file")
        ln = self.lines[self.i]
        self.i += 1
        return ln

    def skip(self):
        self.line()

    def values(self, n, conv=float):
        out = []
        while len(out) < n:
            for tok in self.line().replace(",", " ").split():
                if len(out) == n:
                    break
                out.append(_num(tok) if conv is float else conv(tok))
        return out

    def flags(self, n, true_chars):
        """n one-character flags, 80 per record ('(80i1)' / '(80l1)')."""
        out = []
        while len(out) < n:
            rec = self.line()[:80].ljust(80)
            out.extend(c in true_chars for c in rec[:min(80, n - len(out))])
        return np.array(out, dtype=np.uint8)


# ---------------------------------------------------------------- topology
def read_top(path):
    """topo_read (src/topo.f90:522-1116) of a spherical-boundary topology."""
    with open(path) as f:
        r = _Records(f.read())
    t = {"version": 2.0}
    title = r.line()
    if title.strip() == "Q topology file":
        while True:
            ln = r.line()
            key = ln.split()[0].upper() if ln.split() else ""
            rest = ln.strip()[len(key):].strip()
            if key == "VERSION":
                t["version"] = float(re.match(r"[0-9]+(\.[0-9]+)?", rest.split()[0]).group(0))
            elif key == "END":
                break
    v = _lead(r.line(), (int, int, int, int))
    if not v:
        raise QFileError("atom count line")
    natom = v[0]
    nsol = v[1] if len(v) > 1 else natom
    sa = v[2] if len(v) > 2 else 3
    t["dielectric"] = v[3] if len(v) > 3 and v[3] > 0 else 80000
    if sa == 0:
        nwat, sa = 0, 1
    else:
        nwat = (natom - nsol) // sa
    t.update(natom=natom, nat_solute=nsol, solv_atom=sa, nwat=nwat)
    t["x"] = np.array(r.values(3 * natom), float).reshape(natom, 3)
    r.skip()
    t["iac"] = np.array(r.values(natom, int), np.int64)

    def counted(ncols, conv=int):
        v = _lead(r.line(), (int, int))
        n = v[0]
        nsolute = v[1] if len(v) > 1 else n
        return np.array(r.values(ncols * n, conv), np.int64).reshape(n, ncols), nsolute

    t["bonds"], t["nbonds_solute"] = counted(3)
    nb = _lead(r.line(), (int,))[0]
    t["bondlib"] = np.array([_lead(r.line(), (int, float, float))[1:3] for _ in range(nb)], float).reshape(nb, 2)
    t["angles"], t["nangles_solute"] = counted(4)
    na = _lead(r.line(), (int,))[0]
    alib = []
    for _ in range(na):
        v = _lead(r.line(), (int, float, float, float, float))[1:]
        alib.append((v + [0.0, 0.0, 0.0, 0.0])[:4])
    t["anglib"] = np.array(alib, float).reshape(na, 4)
    t["torsions"], t["ntors_solute"] = counted(5)
    nt = _lead(r.line(), (int,))[0]
    t["torlib"] = np.array([r.values(5)[1:] for _ in range(nt)], float).reshape(nt, 4)
    t["impropers"], t["nimps_solute"] = counted(5)
    v = _lead(r.line(), (int, int))
    t["imp_type"] = v[1] if len(v) > 1 else 1
    t["implib"] = np.array([r.values(3)[1:] for _ in range(v[0])], float).reshape(v[0], 2)
    n = _lead(r.line(), (int,))[0]
    t["crg"] = np.array(r.values(n), float)
    v = _lead(r.line(), (int, int, int))
    ncgp = v[0]
    t["ncgp_solute"] = v[1] if len(v) > 1 else ncgp
    t["iuse_switch_atom"] = v[2] if len(v) > 2 else 1
    first, last, sw, atoms = [], [], [], []
    for _ in range(ncgp):
        nat, isw = r.values(2, int)
        first.append(len(atoms))
        atoms += r.values(nat, int)
        last.append(len(atoms) - 1)
        sw.append(isw)
    t.update(ncgp=ncgp, cgp_first=np.array(first), cgp_last=np.array(last), cgp_iswitch=np.array(sw),
             cgpatom=np.array(atoms, np.int64))
    ntypes = r.values(1, int)[0]
    t["ivdw_rule"] = r.values(1, int)[0]
    v = _lead(r.line(), (float, float))
    t["el14_scale"] = v[0]
    t["coulomb"] = v[1] if len(v) > 1 and v[1] > 0.0 else 332.0
    r.skip()
    t["mass"] = np.array(r.values(ntypes), float)
    avdw = np.zeros((ntypes, NLJTYP))
    bvdw = np.zeros((ntypes, NLJTYP))
    for k in range(NLJTYP):
        r.skip()
        avdw[:, k] = r.values(ntypes)
        r.skip()
        bvdw[:, k] = r.values(ntypes)
    t.update(ntypes=ntypes, avdw=avdw, bvdw=bvdw)
    nlj2 = r.values(1, int)[0]
    t["lj2"] = np.array([r.values(2, int) for _ in range(nlj2)], np.int64).reshape(nlj2, 2)
    r.skip()
    t["list14"] = r.flags(MAX_NBR_RANGE * nsol, "1") if nsol > 0 else np.zeros(0, np.uint8)
    n = r.values(1, int)[0]
    t["list14long"] = np.array([r.values(2, int) for _ in range(n)], np.int64).reshape(n, 2)
    r.skip()
    t["listex"] = r.flags(MAX_NBR_RANGE * nsol, "1") if nsol > 0 else np.zeros(0, np.uint8)
    n = r.values(1, int)[0]
    t["listexlong"] = np.array([r.values(2, int) for _ in range(n)], np.int64).reshape(n, 2)
    ln = r.line()
    if ln.strip() == "":
        ln = r.line()
    v = _lead(ln, (int, int))
    nres = v[0]
    t["nres"], t["nres_solute"] = nres, (v[1] if len(v) > 1 else nres)
    t["res_start"] = np.array(r.values(nres, int)) if nres > 0 else np.zeros(0, int)
    r.skip()
    t["res_name"] = []
    for i in range((nres + 15) // 16):
        ln = r.line()
        t["res_name"] += [ln[5 * j:5 * j + 4].strip() for j in range(min(16, nres - 16 * i))]
    nmol = r.values(1, int)[0]
    t["nmol"] = nmol
    t["mol_start"] = np.array(r.values(nmol, int)) if nmol > 0 else np.zeros(0, int)
    r.skip()
    t["type_names"] = []
    for i in range((ntypes + 7) // 8):
        ln = r.line()
        t["type_names"] += [ln[9 * j:9 * j + 8].strip() for j in range(min(8, ntypes - 8 * i))]
    r.skip()
    for i in range((ntypes + 12) // 13):
        r.skip()
    if t["version"] < 4:
        raise QFileError("topology version %g: the sphere record needs version 4 or later" % t["version"])
    t["solvent_type"] = r.values(1, int)[0]
    ln = r.line()
    if ln.split()[0].upper() == "PBC":
        raise QFileError("periodic-box topologies are not supported (spherical boundary only)")
    if t["version"] < 5.01:
        v = _lead(ln, (float, float, float))
        t["rexcl_o"], t["rwat"], t["rho_wat"] = v[0], v[2], 0.0335
    else:
        v = _lead(ln, (float, float, float))
        t["rexcl_o"], t["rwat"] = v[0], v[1]
        t["rho_wat"] = v[2] if len(v) > 2 and v[2] > 0.0 else 0.0335
    t["xpcent"] = np.array(r.values(3), float)
    t["xwcent"] = np.array(r.values(3), float)
    r.values(2, int)
    t["excl"] = r.flags(natom, "Tt")
    return t


# ---------------------------------------------------------------- FEP file
def _sections(text):
    """Lines of each [section], comments ('!', '#') and blanks removed."""
    out, cur = {}, None
    for raw in text.split("\n"):
        ln = re.split(r"[!#]", raw, 1)[0].strip()
        if not ln:
            continue
        m = re.match(r"^\[(.*)\]$", ln)
        if m:
            cur = m.group(1).strip().lower()
            out.setdefault(cur, [])
            continue
        if cur is not None:
            out[cur].append(ln)
    return out


def _keys(lines):
    out = {}
    for ln in lines:
        p = ln.split(None, 1)
        out[p[0].lower()] = p[1].strip() if len(p) > 1 else ""
    return out


def _flag(v):
    return v.strip().lower() in ("on", "yes", "true", "1", ".true.", "t")


def read_fep(path, top):
    """The FEP file's Q-atoms, charges, Q-atom types and the pair and bond
    changes the nonbonded set-up reads (src/qatom.f90:322-1991)."""
    with open(path) as f:
        sec = _sections(f.read())
    kv = _keys(sec.get("fep", []))
    for k in ("offset", "offset_name", "offset_residue"):
        if k in kv:
            raise QFileError("[FEP] %s is not supported: give topology atom numbers in [atoms]" % k)
    ns = int(kv.get("states", "1"))
    fe = {"nstates": ns, "qq_use_library_charges": _flag(kv.get("qq_use_library_charges", "off")),
          "softcore_use_max_potential": _flag(kv.get("softcore_use_max_potential", "off"))}
    rows = [ln.split() for ln in sec.get("atoms", [])]
    nq = max((int(p[0]) for p in rows), default=0)
    iqseq = np.zeros(nq, np.int64)
    for p in rows:
        if len(p) < 2 or not p[1].lstrip("-").isdigit():
            raise QFileError("[atoms] takes 'q-atom topology-atom' pairs")
        iqseq[int(p[0]) - 1] = int(p[1])
    if nq and (iqseq.min() < 1 or iqseq.max() > top["nat_solute"]):
        raise QFileError("[atoms]: Q-atoms must be solute atoms")
    fe["iqseq"] = iqseq  # 1-based topology numbers
    qcrg = np.repeat(top["crg"][iqseq - 1][:, None], ns, axis=1) if nq else np.zeros((0, ns))
    for ln in sec.get("change_charges", []):
        v = ln.split()
        qcrg[int(v[0]) - 1, :] = [_num(x) for x in v[1:1 + ns]]
    fe["qcrg"] = qcrg
    types = sec.get("atom_types", [])
    fe["vdw_from_topo"] = not types
    if types:
        names, qa, qb, qm = [], [], [], []
        for ln in types:
            v = ln.split()
            names.append(v[0])
            qa.append([_num(v[1]), _num(v[3]), _num(v[5])])
            qb.append([_num(v[2]), _num(v[4]), _num(v[6])])
            qm.append(_num(v[7]))
        index = {n: i for i, n in enumerate(names)}
    else:
        iac = top["iac"][iqseq - 1] - 1
        qa, qb = top["avdw"][iac].tolist(), top["bvdw"][iac].tolist()
        index = None
    fe["qavdw"], fe["qbvdw"] = np.array(qa, float).reshape(-1, 3), np.array(qb, float).reshape(-1, 3)
    ch = sec.get("change_atoms", [])
    if not ch:
        fe["qvdw_flag"] = False
        fe["qiac"] = np.repeat(np.arange(nq)[:, None], ns, axis=1)
    else:
        if len(ch) != nq or index is None:
            raise QFileError("[change_atoms] must list every Q-atom, with [atom_types]")
        fe["qvdw_flag"] = True
        qiac = np.zeros((nq, ns), np.int64)
        for ln in ch:
            v = ln.split()
            qiac[int(v[0]) - 1, :] = [index[n] for n in v[1:1 + ns]]
        fe["qiac"] = qiac
    fe["soft_pairs"] = [tuple(int(x) for x in ln.split()[:2]) for ln in sec.get("soft_pairs", [])]
    fe["el_scale"] = [(int(v[0]), int(v[1]), [_num(x) for x in v[2:2 + ns]])
                      for v in (ln.split() for ln in sec.get("el_scale", []))]
    fe["excluded_pairs"] = [(int(v[0]), int(v[1]), [int(x) == 1 for x in v[2:2 + ns]])
                            for v in (ln.split() for ln in sec.get("excluded_pairs", []))]
    fe["change_bonds"] = [(int(v[0]), int(v[1]), [int(x) for x in v[2:2 + ns]])
                          for v in (ln.split() for ln in sec.get("change_bonds", []))]
    sc = sec.get("softcore")
    fe["alpha_max"] = None
    if sc is not None:
        if not fe["qvdw_flag"] or len(sc) != nq:
            raise QFileError("[softcore] needs [change_atoms] and one line per Q-atom")
        am = np.zeros((nq, ns))
        for ln in sc:
            v = ln.split()
            am[int(v[0]) - 1, :] = [_num(x) for x in v[1:1 + ns]]
        fe["alpha_max"] = am
    fe["off_diagonals"] = [ln.split() for ln in sec.get("off_diagonals", [])]
    fe["raw"] = sec
    return fe


# ---------------------------------------------------------------- MD input
def read_inp(path):
    """Sections of an MD input as {section: {key: value}} plus the
    [lambdas] values; file names resolve against the input's folder."""
    with open(path) as f:
        sec = _sections(f.read())
    out = {k: _keys(v) for k, v in sec.items() if k not in ("lambdas", "sequence_restraints",
                                                            "distance_restraints", "group_contribution")}
    out["lambdas"] = [_num(x) for ln in sec.get("lambdas", []) for x in ln.split()]
    out["_dir"] = os.path.dirname(os.path.abspath(path))
    return out


def _get(d, sec, key, default=None, conv=float):
    v = d.get(sec, {}).get(key)
    if v is None:
        return default
    return conv(v.split()[0]) if conv is not str else v


def _softcore_lookup(top, fe, bvdw_topo):
    """sc_lookup(iq, type or natyps + jq, state) (src/qatom.f90:1865-1991)."""
    nq, ns, nt = len(fe["iqseq"]), fe["nstates"], top["ntypes"]
    sc = np.zeros((nq, nt + nq, ns))
    am = fe["alpha_max"]
    if am is None:
        return sc
    qa, qb, qiac = fe["qavdw"], fe["qbvdw"], fe["qiac"]
    umax, rule = fe["softcore_use_max_potential"], top["ivdw_rule"]
    for i in range(nq):
        for s in range(ns):
            for j in range(nt):
                if umax:
                    aq, bq = qa[qiac[i, s], 0], qb[qiac[i, s], 0]
                    aj, bj = top["avdw"][j, 0], bvdw_topo[j, 0]
                    if am[i, s] > 1e-6:
                        if rule == 1:
                            sc[i, j, s] = (-bq * bj + math.sqrt(bq * bq * bj * bj + 4.0 * am[i, s] * aq * aj)) / (
                                2.0 * am[i, s])
                        else:
                            sc[i, j, s] = (-2.0 * math.sqrt(bq) * bj + 2.0 * math.sqrt(
                                bq * bj ** 2 + am[i, s] * math.sqrt(bq) * bj)) * (aq + aj) ** 6 / (2.0 * am[i, s])
                else:
                    sc[i, j, s] = am[i, s]
            for j in range(nq):
                if am[i, s] > 1e-6 or am[j, s] > 1e-6:
                    aq, bq = qa[qiac[i, s], 0], qb[qiac[i, s], 0]
                    aj, bj = qa[qiac[j, s], 0], qb[qiac[j, s], 0]
                    if umax:
                        amt = min(am[i, s], am[j, s])
                        if am[i, s] < 1e-6 or am[j, s] < 1e-6:
                            amt = max(am[i, s], am[j, s])
                        if rule == 1:
                            sc[i, nt + j, s] = (-bq * bj + math.sqrt(bq * bq * bj * bj + 4.0 * amt * aq * aj)) / (
                                2.0 * amt)
                        else:
                            sc[i, nt + j, s] = (-2.0 * math.sqrt(bq * bj) + 2.0 * math.sqrt(
                                bq * bj + amt * math.sqrt(bq * bj))) * (aq + aj) ** 6 / (2.0 * amt)
                    else:
                        sc[i, nt + j, s] = max(am[i, s], am[j, s])
    return sc


def _pair(rule, a1, a2, b1, b2):
    """vdWA, vdWB of one pair from the per-atom factors
    (precompute_set_values_*, src/simprep.f90:3321-3589)."""
    if rule == 1:
        return a1 * a2, b1 * b2
    t = a1 + a2
    t = t ** 2
    t = t * t * t
    tb = b1 * b2
    return (t ** 2) * tb, 2.0 * t * tb


def _qconn(top, fe, bonds_active):
    """make_qconn / find_bonded (src/nonbondene.f90:3087-3170): bonds
    between each solute atom and each Q-atom per state, 1 (itself) to 4,
    9 beyond, 0 for the FEP file's excluded pairs."""
    nsol, iqseq, ns = top["nat_solute"], fe["iqseq"], fe["nstates"]
    nq = len(iqseq)
    qconn = np.full((ns, nsol + 1, nq), 9, np.int64)  # 1-based atoms
    iqatom = np.zeros(top["natom"] + 1, np.int64)
    iqatom[iqseq] = np.arange(1, nq + 1)
    for iq in range(nq):
        qconn[:, iqseq[iq], iq] = 1
    qb = fe["change_bonds"]
    for iq in range(nq):
        for s in range(ns):
            # find_bonded's depth-first walk; a shorter path found later
            # lowers an atom's level and walks on from it
            def walk(cur, level):
                nl = level + 1
                for (i, j) in bonds_active:
                    for a, b in ((i, j), (j, i)):
                        if a == cur:
                            if qconn[s, b, iq] > nl:
                                qconn[s, b, iq] = nl
                                if nl < 4:
                                    walk(b, nl)
                            break
                for (i, j, cods) in qb:
                    if cods[s] > 0:
                        for a, b in ((i, j), (j, i)):
                            if a == cur:
                                if qconn[s, b, iq] > nl:
                                    qconn[s, b, iq] = nl
                                    if nl < 4:
                                        walk(b, nl)
                                break
            walk(iqseq[iq], 1)
    for (i, j, flags) in fe["excluded_pairs"]:
        for a, b in ((i, j), (j, i)):
            iq = iqatom[a]
            if iq > 0:
                for s in range(ns):
                    if flags[s]:
                        qconn[s, b, iq - 1] = 0
    return qconn[:, 1:, :]


def _wat_shells(top, chg1, bnd_last_cod, ang_last_cod, crgQtot):
    """wat_shells (src/simprep.f90:4707-4891) without a restart file: the
    solvent dipole from the last bond's and the last angle's codes (the
    angle in radians, deg2rad = pi/180, src/math.f90:621-624)."""
    ang0 = (math.pi / 180.0) * top["anglib"][ang_last_cod - 1, 1]
    mu_w = -chg1 * top["bondlib"][bnd_last_cod - 1, 1] * math.cos(ang0 / 2)
    drs = WPOLR_LAYER / DROUT
    nsh = int(-0.5 + math.sqrt(2.0 * drs + 0.25))
    rout, rho = top["rwat"], top["rho_wat"]
    outs, drs_, cstb = [], [], []
    for i in range(1, nsh + 1):
        dr = DROUT * i
        ri = rout - dr
        rshell = (0.5 * (rout ** 3 + ri ** 3)) ** (1.0 / 3.0)
        eps = float(top["dielectric"]) / 1000.0
        eps = 1.0 - (1.0 / eps)
        outs.append(rout)
        drs_.append(dr)
        cstb.append(crgQtot * eps / (rho * mu_w * 4.0 * math.pi * rshell ** 2))
        rout = rout - dr
    return nsh, np.array(outs), np.array(drs_), np.array(cstb)


def build_system(inp_path, x=None):
    """A QSystem from an MD input and the topology and FEP file it names,
    set up as Qdyn6 sets it up (positions from the topology unless x is
    given; velocities zero: use q6_amd.qdyn_init for the seeded start)."""
    inp = read_inp(inp_path)
    d = inp["_dir"]
    top_name = _get(inp, "files", "topology", conv=str)
    if top_name is None:
        raise QFileError("[files] names no topology")
    top = read_top(os.path.join(d, top_name.split()[0]))
    fep_name = _get(inp, "files", "fep", conv=str)
    fe = read_fep(os.path.join(d, fep_name.split()[0]), top) if fep_name else None
    natom, nsol, sa, nwat = top["natom"], top["nat_solute"], top["solv_atom"], top["nwat"]
    rule, nt = top["ivdw_rule"], top["ntypes"]
    iac = top["iac"]                       # 1-based types
    # Q-atoms and states (get_fep, src/simprep.f90:1000-1048)
    nq = len(fe["iqseq"]) if fe else 0
    ns = fe["nstates"] if fe else 0
    lam = inp["lambdas"][:ns] if inp["lambdas"] else ([1.0] + [0.0] * (ns - 1) if ns else [])
    if ns and len(lam) != ns:
        raise QFileError("[lambdas] needs one value per state")
    iqatom = np.zeros(natom, np.int64)
    if nq:
        iqatom[fe["iqseq"] - 1] = np.arange(1, nq + 1)
    qavdw = fe["qavdw"].copy() if fe else np.zeros((0, 3))
    qbvdw = fe["qbvdw"].copy() if fe else np.zeros((0, 3))
    qvdw_flag = bool(fe and fe["qvdw_flag"])
    if qvdw_flag and rule == 2:
        qbvdw[:, 0] = np.sqrt(qbvdw[:, 0])
        qbvdw[:, 2] = np.sqrt(qbvdw[:, 2])
    bvdw_fep = top["bvdw"]  # the softcore lookup reads the topology's epsilons as loaded
    sc = _softcore_lookup(top, dict(fe, qavdw=qavdw, qbvdw=qbvdw), bvdw_fep) if fe else np.zeros((0, nt, 0))
    # topology (src/simprep.f90:4540-4565)
    num_atyp = int(iac.max())
    ljcod = np.ones((nt, nt), np.int64)
    for i, j in top["lj2"]:
        ljcod[i - 1, j - 1] = ljcod[j - 1, i - 1] = 2
    if num_atyp < nt:
        ljcod[num_atyp:, :] = 1
        ljcod[:, num_atyp:] = 1
    avdw = top["avdw"]
    bvdw = np.sqrt(np.abs(top["bvdw"])) if rule == 2 else top["bvdw"]
    # run parameters ([MD], [cut-offs], [sphere], [solvent]; src/md.f90)
    use_lrf = _flag(_get(inp, "md", "lrf", "on", str))
    stepsize = _get(inp, "md", "stepsize")
    cut = inp.get("cut-offs", {})
    Rcpp = _get(inp, "cut-offs", "solute_solute", RC_DEFAULT)
    Rcww = _get(inp, "cut-offs", "solvent_solvent", RC_DEFAULT)
    Rcpw = _get(inp, "cut-offs", "solute_solvent", RC_DEFAULT)
    Rcq = _get(inp, "cut-offs", "q_atom", RCQ_DEFAULT_SPH)
    # with a [cut-offs] section RcLRF is read only when LRF is on (and left
    # at zero otherwise, src/md.f90:608-650)
    RcLRF = (_get(inp, "cut-offs", "lrf", RCLRF_DEFAULT) if use_lrf else 0.0) if cut else RCLRF_DEFAULT
    if use_lrf and (RcLRF < Rcpp or RcLRF < Rcpw or RcLRF < Rcww):
        raise QFileError("LRF cut-off must not be smaller than solute or solvent cut-offs")
    solv = inp.get("solvent", inp.get("water"))
    fk_wsphere = _get(inp, "solvent", "radial_force", -1.0) if solv is not None else -1.0
    wpol_restr = _flag(solv["polarisation"]) if solv and "polarisation" in solv else True
    wpol_born = _flag(solv["charge_correction"]) if solv and "charge_correction" in solv else wpol_restr
    fkwpol = _get(inp, "solvent", "polarisation_force", -1.0) if solv is not None else -1.0
    Dwmz = _get(inp, "solvent", "morse_depth", -1.0) if solv is not None else -1.0
    awmz = _get(inp, "solvent", "morse_width", -1.0) if solv is not None else -1.0
    itdis = ITDIS_UPDATE / stepsize if stepsize else float(ITDIS_UPDATE)
    if solv and wpol_restr and "pol_update" in solv:
        itdis = float(int(solv["pol_update"].split()[0]))
    if solv and "radius" in solv and _num(solv["radius"].split()[0]) > 0.0:
        top["rwat"] = _num(solv["radius"].split()[0])
    # charges and solvent tables (prep_sim, wat_sphere: unscaled charges)
    crg = top["crg"].copy()
    qcrg = fe["qcrg"].copy() if fe else np.zeros((0, 0))
    crgQtot = 0.0
    for i in range(nq):
        for s in range(ns):
            crgQtot = crgQtot + qcrg[i, s] * lam[s]
    if not wpol_born:
        crgQtot = 0.0
    if nwat > 0:
        if top["solvent_type"] not in (0, 1):
            raise QFileError("mixed solvents (solvent type 2) are not supported by Qdyn6 either")
        chg_solv = crg[nsol:nsol + sa].copy()
        aLJ = avdw[iac[nsol:nsol + sa] - 1].copy()
        bLJ = bvdw[iac[nsol:nsol + sa] - 1].copy()
        if fk_wsphere == -1:
            fk_wsphere = FK_WSPHERE_DEFAULT
        if fkwpol == -1:
            fkwpol = FKWPOL_DEFAULT
        rwat = top["rwat"]
        if Dwmz == -1:
            Dwmz = 0.26 * math.exp(-0.19 * (rwat - 15.0)) + 0.74
        if awmz == -1:
            awmz = 0.2 / (1.0 + math.exp(0.4 * (rwat - 25.0))) + 0.3
        nsh = 0
        sh = (np.zeros(0), np.zeros(0), np.zeros(0))
        if wpol_restr:
            bcod = int(top["bonds"][-1, 2])
            acod = int(top["angles"][-1, 3])
            nsh, *sh = _wat_shells(top, chg_solv[0], bcod, acod, crgQtot)
    sq = math.sqrt(top["coulomb"])
    crg = crg * sq
    qcrg = qcrg * sq
    s = QSystem()
    s.name = inp_path
    n = s.n
    n.update(natom=natom, nat_solute=nsol, nwat=nwat, solv_atom=sa, ncgp=top["ncgp"],
             ncgp_solute=top["ncgp_solute"], nqat=nq, nstates=ns, ntypes=nt, ivdw_rule=rule,
             solvent_spc=1 if (top["solvent_type"] == 0 and rule == 1) else 0,
             any_atom_cutoff=1 if top["iuse_switch_atom"] == 0 else 0)
    s.el14_scale = top["el14_scale"]
    a = s.a
    a["crg"] = crg
    a["iac"] = (iac - 1).astype(np.int32)
    a["iqatom"] = iqatom
    a["excl"] = top["excl"]
    a["heavy"] = (top["mass"][iac - 1] >= 4.0).astype(np.uint8)
    a["cgp_first"], a["cgp_last"] = top["cgp_first"], top["cgp_last"]
    a["cgp_iswitch"] = top["cgp_iswitch"] - 1
    a["cgpatom"] = top["cgpatom"] - 1
    a["avdw"] = avdw.ravel()
    a["bvdw"] = bvdw.ravel()
    a["ljcod"] = ljcod.ravel(order="F")
    if nsol > 0:
        a["listex"], a["list14"] = top["listex"], top["list14"]
        a["listexlong"] = (top["listexlong"] - 1).ravel()
        a["list14long"] = (top["list14long"] - 1).ravel()
        n["nexlong"], n["n14long"] = len(top["listexlong"]), len(top["list14long"])
    # solute atoms of the solute charge groups (pw_int_comp's loop)
    in_group = np.zeros(natom, bool)
    for g in range(top["ncgp_solute"]):
        in_group[top["cgpatom"][top["cgp_first"][g]:top["cgp_last"][g] + 1] - 1] = True
    if nwat > 0:
        chg = chg_solv * sq
        wA, wB, wE = np.zeros((sa, sa)), np.zeros((sa, sa)), np.zeros((sa, sa))
        for i in range(sa):
            for j in range(sa):
                v = ljcod[iac[nsol + i] - 1, iac[nsol + j] - 1] - 1
                wA[i, j], wB[i, j] = _pair(rule, aLJ[i, v], aLJ[j, v], bLJ[i, v], bLJ[j, v])
                wE[i, j] = chg[i] * chg[j]
        a["ww_A"], a["ww_B"], a["ww_el"] = wA.ravel(order="F"), wB.ravel(order="F"), wE.ravel(order="F")
        if nsol > 0:
            pA, pB, pE = np.zeros((nsol, sa)), np.zeros((nsol, sa)), np.zeros((nsol, sa))
            for i in range(nsol):
                if iqatom[i] != 0 or not in_group[i]:
                    continue
                for k in range(sa):
                    v = ljcod[iac[i] - 1, iac[nsol + k] - 1] - 1
                    pA[i, k], pB[i, k] = _pair(rule, avdw[iac[i] - 1, v], aLJ[k, v], bvdw[iac[i] - 1, v], bLJ[k, v])
                    pE[i, k] = crg[i] * chg[k]
            a["pw_A"], a["pw_B"], a["pw_el"] = pA.ravel(order="F"), pB.ravel(order="F"), pE.ravel(order="F")
        if top["ntors_solute"] < len(top["torsions"]):
            a["solv_int_ij"], a["solv_int_par"] = _solvent_internal(top, ljcod, aLJ, bLJ, chg, rule)
            n["n_solv_int"] = len(a["solv_int_ij"]) // 2
    if nq:
        _q_tables(s, top, fe, iqatom, ljcod, avdw, bvdw, qavdw, qbvdw, qvdw_flag, crg, qcrg, sc,
                  aLJ if nwat else None, bLJ if nwat else None, chg_solv * sq if nwat else None)
    s.x = np.array(top["x"] if x is None else x, float).reshape(-1, 3)
    s.v = np.zeros_like(s.x)
    s.winv = 1.0 / top["mass"][iac - 1]
    p = s.par
    p.update(Rcpp=Rcpp, Rcpw=Rcpw, Rcww=Rcww, Rcq=Rcq, RcLRF=RcLRF, rexcl_o=top["rexcl_o"], rwat=top["rwat"],
             fk_wsphere=fk_wsphere, Dwmz=Dwmz, awmz=awmz, Boltz=BOLTZ, fkwpol=fkwpol, itdis=itdis,
             use_LRF=int(use_lrf), wpol_restr=int(wpol_restr and nwat > 0), nwpolr_shell=0,
             xpcent=top["xpcent"], xwcent=top["xwcent"], lambda_=None)
    p.pop("lambda_")
    p["lambda"] = np.array(lam, float) if ns else np.zeros(1)
    if nwat > 0 and wpol_restr:
        p["nwpolr_shell"] = nsh
        p["wshell_rout"], p["wshell_dr"], p["wshell_cstb"] = sh
        p["wshell_theta_corr"] = np.zeros(nsh)
    s.top, s.fep, s.inp = top, fe, inp
    return s.finalize()


def _solvent_internal(top, ljcod, aLJ, bLJ, chg, rule):
    """prep_sim_precompute_solvent (src/simprep.f90:4004-4138): pairs of
    one solvent molecule outside its bonds and angles, 1-4 pairs (from the
    solvent torsions) with the 1-4 parameters and scaled charges."""
    nsol, sa, iac = top["nat_solute"], top["solv_atom"], top["iac"]
    inter = np.zeros((sa, sa), np.int64)
    for i in range(sa):
        for j in range(sa):
            inter[i, j] = ljcod[iac[nsol + i] - 1, iac[nsol + j] - 1]
    for i in range(sa):
        inter[i, i] = 0
    last = nsol + sa
    for i, j, _ in top["bonds"][top["nbonds_solute"]:]:
        if i > last or j > last:
            continue
        inter[i - nsol - 1, j - nsol - 1] = inter[j - nsol - 1, i - nsol - 1] = 0
    for i, _, k, _ in top["angles"][top["nangles_solute"]:]:
        if i > last or k > last:
            continue
        inter[i - nsol - 1, k - nsol - 1] = inter[k - nsol - 1, i - nsol - 1] = 0
    for i, _, _, l, _ in top["torsions"][top["ntors_solute"]:]:
        if i > last or l > last:
            continue
        a, b = i - nsol - 1, l - nsol - 1
        if inter[a, b] != 0 or inter[b, a] != 0:
            inter[a, b] = 3
            inter[b, a] = 0
    ij, par = [], []
    for i in range(sa):
        for j in range(sa):
            if inter[i, j] != 0:
                v = inter[i, j] - 1
                A, B = _pair(rule, aLJ[i, v], aLJ[j, v], bLJ[i, v], bLJ[j, v])
                el = chg[i] * chg[j]
                if inter[i, j] == 3:
                    el = el * top["el14_scale"]
                ij += [i, j]
                par += [A, B, el]
    return np.array(ij, np.int32), np.array(par, float)


def _q_tables(s, top, fe, iqatom, ljcod, avdw, bvdw, qavdw, qbvdw, qvdw_flag, crg, qcrg, sc, aLJ, bLJ, chg):
    """qconn and the qp/qq/qw tables (qp_int_comp, qq_int_comp,
    qw_int_comp, src/simprep.f90:3175-3300, 3388-3559)."""
    nsol, sa, rule, nt = top["nat_solute"], top["solv_atom"], top["ivdw_rule"], top["ntypes"]
    iac, iqseq, ns = top["iac"], fe["iqseq"], fe["nstates"]
    nq = len(iqseq)
    qiac = fe["qiac"]
    lib = fe["qq_use_library_charges"]
    # topology bonds with the FEP file's redefined ones removed (their code
    # set to 0, src/simprep.f90:1049-1064); only solute bonds are walked
    redef = {(min(i, j), max(i, j)) for i, j, _ in fe["change_bonds"]}
    bonds = [(int(i), int(j)) for i, j, c in top["bonds"][:top["nbonds_solute"]]
             if c != 0 and (min(i, j), max(i, j)) not in redef]
    qconn = _qconn(top, fe, bonds) if nsol > 0 else np.zeros((ns, 0, nq), np.int64)
    s.qconn = qconn
    a = s.a
    qA, qB, qE = np.zeros((nsol, nq, ns)), np.zeros((nsol, nq, ns)), np.zeros((nsol, nq, ns))
    qS, qset = np.zeros((nsol, nq, ns)), np.zeros((nsol, nq, ns), np.uint8)

    def set_qp(jq, j, st, vdw):
        i = iqseq[jq] - 1
        qv = 1 if vdw == 2 else vdw
        if qvdw_flag:
            A, B = _pair(rule, qavdw[qiac[jq, st], qv - 1], avdw[iac[j] - 1, vdw - 1],
                         qbvdw[qiac[jq, st], qv - 1], bvdw[iac[j] - 1, vdw - 1])
        else:
            A, B = _pair(rule, avdw[iac[i] - 1, vdw - 1], avdw[iac[j] - 1, vdw - 1],
                         bvdw[iac[i] - 1, vdw - 1], bvdw[iac[j] - 1, vdw - 1])
        el = qcrg[jq, st] * crg[j] if not lib else crg[i] * crg[j]
        qA[j, jq, st], qB[j, jq, st], qE[j, jq, st] = A, B, el
        if abs(A) > _EPS or abs(B) > _EPS or abs(el) > _EPS:
            qset[j, jq, st] = 1
        qS[j, jq, st] = sc[jq, iac[j] - 1, st]
        if vdw == 3:
            qE[j, jq, st] = qE[j, jq, st] * top["el14_scale"]

    for ig in range(nsol):
        if (qconn[:, ig, :] <= 3).any():
            continue
        for jq in range(nq):
            vdw = ljcod[iac[ig] - 1, iac[iqseq[jq] - 1] - 1]
            for st in range(ns):
                if qconn[st, ig, jq] == 4:
                    vdw = 3
                set_qp(jq, ig, st, vdw)
    # Q-Q pairs
    QA, QB, QE = np.zeros((nq, nq, ns)), np.zeros((nq, nq, ns)), np.zeros((nq, nq, ns))
    QS, Qset, Qsoft = np.zeros((nq, nq, ns)), np.zeros((nq, nq, ns), np.uint8), np.zeros((nq, nq, ns), np.uint8)
    soft = fe["soft_pairs"]
    for iq in range(nq - 1):
        ia = iqseq[iq] - 1
        for jq in range(iq + 1, nq):
            ja = iqseq[jq] - 1
            for st in range(ns):
                if qconn[st, ja, iq] < 4:
                    continue
                els = 1.0
                for (k, l, sc_) in fe["el_scale"]:
                    if (iq + 1, jq + 1) in ((k, l), (l, k)):
                        els = sc_[st]
                        break
                if qconn[st, ja, iq] == 4:
                    vdw = 3
                elif not qvdw_flag:
                    vdw = ljcod[iac[ia] - 1, iac[ja] - 1]
                else:
                    vdw = 1
                    for (k, l) in soft:
                        if (iq + 1, jq + 1) in ((k, l), (l, k)):
                            vdw = 2
                            break
                ti, tj = qiac[iq, st], qiac[jq, st]
                if rule == 1:
                    if qvdw_flag:
                        A, B = qavdw[ti, vdw - 1] * qavdw[tj, vdw - 1], qbvdw[ti, vdw - 1] * qbvdw[tj, vdw - 1]
                    else:
                        A = avdw[iac[ia] - 1, vdw - 1] * avdw[iac[ja] - 1, vdw - 1]
                        B = bvdw[iac[ia] - 1, vdw - 1] * bvdw[iac[ja] - 1, vdw - 1]
                else:
                    if qvdw_flag:
                        tA = (qavdw[ti, vdw - 1] * qavdw[tj, vdw - 1] if vdw == 2
                              else qavdw[ti, vdw - 1] + qavdw[tj, vdw - 1])
                        tB = qbvdw[ti, vdw - 1] * qbvdw[tj, vdw - 1]
                    else:
                        tA = avdw[iac[ia] - 1, vdw - 1] + avdw[iac[ja] - 1, vdw - 1]
                        tB = bvdw[iac[ia] - 1, vdw - 1] * bvdw[iac[ja] - 1, vdw - 1]
                    if vdw == 2 and qvdw_flag:
                        A, B = tA, tB
                    else:
                        tA = tA ** 2
                        tA = tA * tA * tA
                        A, B = (tA ** 2) * tB, 2.0 * tA * tB
                el = qcrg[iq, st] * qcrg[jq, st] if not lib else crg[ia] * crg[ja]
                el = el * els
                QA[iq, jq, st], QB[iq, jq, st], QE[iq, jq, st] = A, B, el
                QS[iq, jq, st] = sc[iq, nt + jq, st]
                if abs(A) > _EPS or abs(B) > _EPS or abs(el) > _EPS:
                    Qset[iq, jq, st] = 1
                if vdw == 3:
                    QE[iq, jq, st] = QE[iq, jq, st] * top["el14_scale"]
                if vdw == 2 and qvdw_flag:
                    Qsoft[iq, jq, st] = 1
    # Q-atom - solute pairs bonded or angled to some Q-atom (qq_int_comp's second loop)
    for ja in range(nsol):
        if iqatom[ja] != 0 or not (qconn[:, ja, :] <= 3).any():
            continue
        for iq in range(nq):
            ia = iqseq[iq] - 1
            for st in range(ns):
                if qconn[st, ja, iq] >= 4:
                    if qconn[st, ja, iq] == 4:
                        vdw = 3
                    elif qvdw_flag:
                        vdw = 1
                    else:
                        vdw = ljcod[iac[ia] - 1, iac[ja] - 1]
                    set_qp(iq, ja, st, vdw)
    F = lambda v: v.ravel(order="F")  # noqa: E731
    a["iqseq"] = iqseq - 1
    if nsol > 0:
        a["qconn3"] = (qconn <= 3).any(axis=(0, 2)).astype(np.uint8)
        a["qp_A"], a["qp_B"], a["qp_el"], a["qp_sc"], a["qp_set"] = F(qA), F(qB), F(qE), F(qS), F(qset)
    a["qq_A"], a["qq_B"], a["qq_el"], a["qq_sc"] = F(QA), F(QB), F(QE), F(QS)
    a["qq_set"], a["qq_soft"] = F(Qset), F(Qsoft)
    if aLJ is not None:
        wA, wB, wE, wS = (np.zeros((nq, sa, ns)) for _ in range(4))
        for iq in range(nq):
            i = iqseq[iq] - 1
            for k in range(sa):
                iacj = iac[top["nat_solute"] + k]
                vdw = ljcod[iacj - 1, iac[i] - 1]
                qv = 1 if vdw == 2 else vdw
                for st in range(ns):
                    if qvdw_flag:
                        A, B = _pair(rule, qavdw[qiac[iq, st], qv - 1], aLJ[k, vdw - 1], qbvdw[qiac[iq, st], qv - 1],
                                     bLJ[k, vdw - 1])
                    else:
                        A, B = _pair(rule, avdw[iac[i] - 1, vdw - 1], aLJ[k, vdw - 1], bvdw[iac[i] - 1, vdw - 1],
                                     bLJ[k, vdw - 1])
                    wA[iq, k, st], wB[iq, k, st] = A, B
                    wE[iq, k, st] = qcrg[iq, st] * chg[k] if not lib else crg[i] * chg[k]
                    wS[iq, k, st] = sc[iq, iacj - 1, st]
        a["qw_A"], a["qw_B"], a["qw_el"], a["qw_sc"] = F(wA), F(wB), F(wE), F(wS)


_EPS = 1.0e-10  # QREAL_EPS of the double-precision build (src/sizes.f90:50)


def seeded_start(s):
    """Positions and velocities of Qdyn6's seeded start for a system read by
    build_system (q6_amd.qdyn_init.seeded_start: maxwell from [MD]
    random_seed at initial_temperature, the rigid-water constraints with
    the topology's O-H and H-H bond lengths squared, centre-of-mass
    removal).  Only the default constraint set is covered: shake_solvent on,
    shake_solute and shake_hydrogens off (src/md.f90:338-360)."""
    from . import qdyn_init
    inp, top = s.inp, s.top
    md = inp.get("md", {})
    if "random_seed" not in md or "initial_temperature" not in md:
        raise QFileError("[MD] gives no random_seed and initial_temperature")
    if not _flag(md.get("shake_solvent", "on")) or _flag(md.get("shake_solute", "off")) or _flag(
            md.get("shake_hydrogens", "off")):
        raise QFileError("only the default constraints (rigid solvent) are covered")
    nsol, nwat = top["nat_solute"], top["nwat"]
    if nwat > 0 and top["solv_atom"] != 3:
        raise QFileError("the seeded start covers 3-site solvents")
    roh2 = rhh2 = 0.0
    if nwat > 0:
        b0 = {}
        for i, j, c in top["bonds"][top["nbonds_solute"]:]:
            b0[(min(i, j), max(i, j))] = top["bondlib"][c - 1, 1]
        o = nsol + 1
        roh2, rhh2 = b0[(o, o + 1)] ** 2, b0[(o + 1, o + 2)] ** 2
    mass = top["mass"][top["iac"] - 1]
    return qdyn_init.seeded_start(s.x, mass, _num(md["initial_temperature"].split()[0]),
                                  int(md["random_seed"].split()[0]), _num(md["stepsize"].split()[0]), nsol, nwat,
                                  roh2, rhh2)
