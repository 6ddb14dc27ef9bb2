"""Writes the golden fixtures of the nonbonded hot path from the reference itself.

Runs in the development container only (it needs /root/reference and the
reference build under oracle/_ref/, made by `make -C oracle/ref`).  For each case
it prepares a Qdyn6 input from the reference's own SPH test system
(tests/basic_tests/prep_SPH/lig_w.{top,fep}, tests/basic_tests/eq1.inp with the
substitutions tests/basic_tests/run_test.sh:171-185 makes for SPH), runs
oracle/_ref/qnb_dump (this repository's driver linked against the reference's
modules) and converts the dump into a compressed .npz holding:

  * the system as set up by Qdyn6 (coordinates, scaled charges, types, charge
    groups, exclusions, Q-atom precompute tables, SCAAS constants, cutoffs);
  * the pair lists, as (i, j) arrays; water-water as molecule pairs because each
    one expands to solv_atom**2 atom pairs in a fixed order
    (src/nonbondene.f90:4366-4375, checked when the fixture is written);
  * the LRF expansion table after the list update;
  * energy and gradient of every nonbonded term, evaluated into a zeroed d.

Cases:
  lig_eq1  the reference test exactly (1 state, all 46 solute atoms are Q-atoms);
           its Q-surr. energies are the reference's own known answer
           (tests/basic_tests/eval_test.sh:97-101).
  lig_fep2 the same topology with a 2-state FEP file written here: 8 Q-atoms
           with changed charges and types, soft-core alphas, one exponential
           soft pair and an off-diagonal, lambda (0.7, 0.3), solute cutoffs 16 A (the
           ligand has a 13.4 A charge group), water-water 9 A,
           q_atom 12 A (so the Q lists are radius-limited) -> exercises pp, pw,
           qp, qqp, qq, qw, ww, LRF and SCAAS together.
  lig_fep2_nolrf  as lig_fep2 with LRF off (cut-off-only list builders).
  lig_fep2_arith  lig_fep2 under the arithmetic combination rule: the
           topology's rule flag set to 2 and every type's (sqrt A, sqrt B) pair
           rewritten as the (R*/2, epsilon) pair with the same A_ii, B_ii
           (topology and FEP types alike), so the geometric and arithmetic
           systems agree on like-type pairs and differ on mixed ones.  The
           rule also selects the general solvent loops (nonbond_ww/qw instead
           of the _spc forms, src/potene.f90:347-356).
  lig_fep2_anyatom, lig_fep2_anyatom_nolrf  lig_fep2 (and its LRF-off form)
           with the topology's switch-atom flag 0 (iuse_switch_atom = 0,
           src/topo.f90:815): the nbpplis2/nbpwlis2/nbqplis2 builders and
           their _lrf forms, where a charge-group pair is inside a cut-off
           when any atom pair is.
  cjp_excl the reference's group-contribution test system (2CJP epoxide
           hydrolase, 5,622 atoms, 40 Q-atoms in 2 EVB states with soft
           pairs, OPLS-AA, 1,875 excluded atoms; tests/exclude_tests), with
           its five [group_contribution] lines and three atom groups with
           Q-atoms and their bonded neighbours (twelve energy arrays after
           the first): every list and term, and calculate_exc's energies per
           array and state.  Stored without the per-pair list parameters.
  lig_qb2  lig_fep2 plus changing bonds (Morse, harmonic, one forming bond),
           angles (one with Urey-Bradley terms), torsions and impropers with
           Morse couplings (gamma and 1 - gamma, two couplings on one term);
           the fixture keeps x, lambda, the Q bonded tables as Qdyn6 holds them
           after get_fep, and three evaluations of the per-state loop
           (src/potene.f90:183-190) with the Morse coupling state after each.

Usage: python tests/golden/make_golden.py [case ...]  (writes tests/golden/<case>.npz)
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/tests/basic_tests"
DUMP = os.path.join(REPO, "oracle", "_ref", "qnb_dump")
sys.path.insert(0, HERE)
from qdump import read_dump  # noqa: E402

EQ1_SPH = """[MD]
steps                    100
stepsize                   0.1
temperature                0.1
bath_coupling              1
thermostat	berendsen
integrator	leap-frog
separate_scaling        on
random_seed                1
initial_temperature        0.1
lrf                       {lrf}

[sphere]
shell_radius	16
shell_force	20
[solvent]
pol_update	100

[cut-offs]
q_atom	{rq}
lrf	99
{extra_cut}
[intervals]
non_bond                  25
output                    50

[files]
topology                  lig_w.top
fep                       {fep}
final                     eq1.re
{lambdas}
[sequence_restraints]
1 46 0.5 0 1
"""

# Q-atoms of the 2-state case (topology numbers): a cluster around atom 1 and
# a second one around atom 20, ten bonds away, so a soft pair can join them.
FEP2_ATOMS = [1, 2, 3, 22, 25, 19, 20, 45]
FEP2_DQ = [0.10, -0.20, 0.05, 0.15, -0.10, 0.30, -0.25, 0.12]


def fep2_text(top_charges):
    types = [
        ("QN1", 971.75, 28.31, 180.0, 2.5, 687.13, 20.02, 14.01),
        ("QC1", 1059.13, 23.67, 0.0, 0.0, 748.92, 16.74, 12.01),
        ("QC2", 944.52, 22.03, 250.0, 2.5, 667.88, 15.58, 12.01),
        ("QO1", 616.44, 23.77, 90.0, 1.8, 435.89, 16.81, 16.00),
        ("QH1", 84.57, 5.41, 0.0, 0.0, 59.8, 3.83, 1.008),
    ]
    st1 = ["QN1", "QC1", "QC1", "QC2", "QO1", "QC2", "QO1", "QH1"]
    st2 = ["QC1", "QC2", "QC1", "QC2", "QO1", "QO1", "QC2", "QH1"]
    alphas = [(0.0, 5.0), (2.0, 2.0), (0.0, 0.0), (1.0, 3.0), (0.0, 0.0), (4.0, 0.0), (0.0, 2.5), (0.0, 0.0)]
    out = ["[FEP]", "states 2", "", "[atoms]"]
    for k, a in enumerate(FEP2_ATOMS):
        out.append(f"{k + 1} {a}")
    out += ["", "[change_charges]"]
    for k, a in enumerate(FEP2_ATOMS):
        q = top_charges[a - 1]
        out.append(f"{k + 1} {q:.4f} {q + FEP2_DQ[k]:.4f}")
    out += ["", "[atom_types]"]
    for t in types:
        out.append("%-6s %8.2f %8.2f %7.1f %6.2f %8.2f %8.2f %8.3f" % t)
    out += ["", "[change_atoms]"]
    for k in range(len(FEP2_ATOMS)):
        out.append(f"{k + 1} {st1[k]} {st2[k]}")
    out += ["", "[soft_pairs]", "1 7", "", "[softcore]"]
    for k, (a1, a2) in enumerate(alphas):
        out.append(f"{k + 1} {a1} {a2}")
    out += ["", "[off_diagonals]", "1 2 1 7 80.0 0.0", ""]
    return "\n".join(out)


# Q bonded sections of lig_qb2 (topology atom numbers; src/qatom.f90:952-1800
# The one bond that changes presence (19-20) is not a topology bond: for a
# topology bond that breaks, qatom_load_fep collects the topology torsions
# through it into break_torsions(nqtor+1) without a bound check
# (src/qatom.f90:1362-1385), which overruns the heap on this ligand.
# formats: Morse types "Dmz alpha r0", harmonic "fk r0", angle "fk theta0
# [urey_fk urey_r0]", torsion "fk multiplicity delta", improper "fk imp0").
QB_TEXT = """
[bond_types]
1 95.0 2.1 1.05
2 650.0 1.42
3 80.0 1.8 1.95
4 110.0 2.3 1.60

[change_bonds]
1 2 1 3
1 25 2 2
2 3 3 1
19 20 0 4

[angle_types]
1 70.0 109.5
2 60.0 120.0 25.0 2.3
3 45.0 100.0

[change_angles]
2 1 22 1 2
2 1 25 2 0
1 2 3 0 3
3 2 24 1 2

[torsion_types]
1 1.5 3.0 0.0
2 0.8 2.0 180.0
3 2.0 1.0 35.0

[change_torsions]
22 1 2 3 1 2
25 1 2 24 3 0
2 1 22 4 0 1

[improper_types]
1 10.0 0.0
2 5.0 180.0

[change_impropers]
25 1 2 22 1 2
1 2 3 24 2 0

[angle_couplings]
1 1
3 1 1
4 3
1 4

[torsion_couplings]
1 1
2 3 1
2 4

[improper_couplings]
1 1
2 4 1
"""

QB_KEYS = ("x", "lambda", "nstates", "natom", "qb_counts", "qbnd", "qbondlib", "qang", "qanglib", "qtor", "qtorlib",
           "qimp", "qimplib", "iang_coupl", "itor_coupl", "iimp_coupl")


def _to_arith(sa, sb):
    """(sqrt A, sqrt B) of one type -> (R*/2, epsilon): Qdyn6 takes the square
    root of epsilon on loading (FEP types: src/simprep.f90:1042-1047), so that
    (a_i + a_j)^12 b_i b_j = A and 2 (a_i + a_j)^6 b_i b_j = B for i = j
    (src/simprep.f90:3333-3340)."""
    A, B = sa * sa, sb * sb
    if A <= 0.0 or B <= 0.0:
        return 0.0, 0.0
    rstar = (2.0 * A / B) ** (1.0 / 6.0)
    return 0.5 * rstar, B * B / (4.0 * A)


def arith_top(text):
    """The topology with the arithmetic rule and converted LJ parameters
    (topology section 11, read by src/topo.f90:846-876: natyps, rule, 1-4
    line, 'Masses:' line, masses, then for each LJ code a header and
    natyps avdw values, a header and natyps bvdw values)."""
    lines = text.split("\n")
    k = next(i for i, ln in enumerate(lines) if "vdW combination rule" in ln)
    ntyp = int(lines[k - 1].split()[0])
    assert lines[k].split()[0] == "1"
    lines[k] = lines[k].replace("1 =", "2 =", 1)
    j = k + 3  # after the 1-4 line and the 'Masses:' line

    def take(j):
        vals, start = [], j
        while len(vals) < ntyp:
            vals += [float(t) for t in lines[j].split()]
            j += 1
        assert len(vals) == ntyp
        return vals, start, j

    _, _, j = take(j)
    blocks = []
    for _ in range(3):
        a, a0, j = take(j + 1)
        b, b0, j = take(j + 1)
        blocks.append((a, a0, b, b0))
    out = list(lines)
    for a, a0, b, b0 in reversed(blocks):
        conv = [_to_arith(x, y) for x, y in zip(a, b)]
        # replace the value lines of each array (later blocks first, so the
        # line numbers of the earlier ones stay valid)
        nb_lines = _count_lines(lines, b0, ntyp)
        out[b0:b0 + nb_lines] = fmt([c[1] for c in conv])
        na_lines = _count_lines(lines, a0, ntyp)
        out[a0:a0 + na_lines] = fmt([c[0] for c in conv])
    return "\n".join(out)


def fmt(vs):
    return [" ".join("%.10f" % v for v in vs[i:i + 8]) for i in range(0, len(vs), 8)]


def _count_lines(lines, j0, n):
    got, j = 0, j0
    while got < n:
        got += len(lines[j].split())
        j += 1
    return j - j0


def anyatom_top(text):
    """The topology with the charge groups' switch-atom flag set to 0 (the
    third number of the 'No. of charge groups' line, src/topo.f90:815)."""
    lines = text.split("\n")
    k = next(i for i, ln in enumerate(lines) if "No. of charge groups" in ln)
    t = lines[k].split()
    assert t[2] == "1", lines[k]
    lines[k] = "%9s%8s%8s = " % (t[0], t[1], "0") + lines[k].split("= ", 1)[1]
    return "\n".join(lines)


def arith_fep(text):
    """FEP [atom_types] rows name Ri Ei Ci ai Ri14 Ei14 mass: the (Ri, Ei)
    pairs converted as the topology's; Ci, ai (exponential soft pairs) kept."""
    out, inside = [], False
    for ln in text.split("\n"):
        if ln.startswith("["):
            inside = ln.strip() == "[atom_types]"
            out.append(ln)
            continue
        t = ln.split()
        if inside and len(t) == 8:
            r, e = _to_arith(float(t[1]), float(t[2]))
            r14, e14 = _to_arith(float(t[5]), float(t[6]))
            ln = " ".join([t[0], "%.10f" % r, "%.10f" % e, t[3], t[4], "%.10f" % r14, "%.10f" % e14, t[7]])
        out.append(ln)
    return "\n".join(out)


def top_charges():
    lines = open(os.path.join(REF, "prep_SPH", "lig_w.top")).read().split("\n")
    for k, ln in enumerate(lines):
        if "No. of atomic charges" in ln:
            n = int(ln.split()[0])
            vals = []
            j = k + 1
            while len(vals) < n:
                vals += [float(t) for t in lines[j].split()]
                j += 1
            return vals
    raise RuntimeError("charges not found")


def run_case(name, inp_text, fep_text=None, top_text=None):
    tmp = tempfile.mkdtemp(prefix="qnbgold_")
    try:
        if top_text is None:
            shutil.copy(os.path.join(REF, "prep_SPH", "lig_w.top"), tmp)
        else:
            with open(os.path.join(tmp, "lig_w.top"), "w") as f:
                f.write(top_text)
        shutil.copy(os.path.join(REF, "prep_SPH", "lig_w.fep"), tmp)
        if fep_text is not None:
            with open(os.path.join(tmp, "fep2.fep"), "w") as f:
                f.write(fep_text)
        with open(os.path.join(tmp, "run.inp"), "w") as f:
            f.write(inp_text)
        log = subprocess.run([DUMP, "out.bin", "run.inp"], cwd=tmp, capture_output=True, text=True)
        if log.returncode != 0 or not os.path.exists(os.path.join(tmp, "out.bin")):
            print(log.stdout[-3000:], log.stderr[-3000:])
            raise RuntimeError("qnb_dump failed for " + name)
        return read_dump(os.path.join(tmp, "out.bin")), log.stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def compact(d, fill=True, list_params=True):
    """Pack a dump into fixture arrays (0-based atom indices everywhere).
    fill=False leaves out the fill module's copies (fill_*); list_params=False
    the per-pair list parameters (nb*_par)."""
    out = {}
    for k, v in d.items():
        if k.startswith("nbww"):
            continue
        if k.endswith("_ij") and k.startswith("nb"):
            continue
        if (not fill and k.startswith("fill_")) or (not list_params and k.startswith("nb") and k.endswith("_par")):
            continue
        out[k] = np.asarray(v)
    # 0-based copies of index arrays; Fortran 1-based values are kept verbatim
    # under their own names so the fixture is a faithful record of the dump
    sa = d["solv_atom"]
    ns_ = d["nat_solute"]
    ww = d["nbww_ij"]
    if ww.shape[1]:
        assert ww.shape[1] % (sa * sa) == 0
        blk = ww.reshape(2, sa * sa, -1, order="F").transpose(0, 2, 1)
        iw = (blk[0, :, 0] - ns_ - 1) // sa
        jw = (blk[1, :, 0] - ns_ - 1) // sa
        # the fixed expansion (la, ka) of each molecule pair
        la = np.repeat(np.arange(sa), sa)
        ka = np.tile(np.arange(sa), sa)
        assert np.array_equal(blk[0] - 1, ns_ + sa * iw[:, None] + la[None, :])
        assert np.array_equal(blk[1] - 1, ns_ + sa * jw[:, None] + ka[None, :])
        out["ww_mol_pairs"] = np.stack([iw, jw]).astype(np.int32)
        par = d["nbww_par"].reshape(3, sa * sa, -1, order="F").transpose(0, 2, 1)
        assert np.all(par == par[:, :1, :])
        if list_params:
            out["ww_blk_par"] = par[:, 0, :]
    else:
        out["ww_mol_pairs"] = np.zeros((2, 0), np.int32)
    for lst in ("nbpp", "nbpw", "nbqp", "nbqw"):
        out[lst + "_ij"] = d[lst + "_ij"].astype(np.int32)
    for k in d:
        if k.startswith("nbqq") and k.endswith("_ij"):
            out[k] = d[k].astype(np.int32)
    return out


# The reference's group-contribution test system (tests/exclude_tests): the
# 2CJP epoxide hydrolase in a 20 A sphere (5,622 atoms, OPLS-AA), its 40-atom
# 2-state EVB FEP file, and the [group_contribution] lines of
# tests/exclude_tests/inputs/excl/gen_inps.pl:42-47, plus three atom groups
# with a Q-atom and atoms bonded to Q-atoms (the reference's groups hold
# neither); the cut-offs of that script (:20-24), lambda (0.5, 0.5).  The
# Q-atom is the last one (Q-atom 40, atom 5124): qq_int_comp counts a group's
# Q-Q pairs with either atom in the group (src/simprep.f90:3212-3221), but
# exc_assign lists (iq, jq) only where qq_precomp(iq, jq) is set, which
# precompute_set_values_qq does for iq < jq alone (:3444-3500), so for any
# other Q-atom calculate_exc would walk list entries never filled.  The topology's start is not relaxed for
# dynamics, so the dump leaves out its MD step ("nomd").
EXCL = "/root/reference/tests/exclude_tests/inputs"
CJP_INP = """[MD]
steps                          10
stepsize                       1.0
temperature                    300
bath_coupling                  100
separate_scaling               on
lrf                            on
random_seed                    1
initial_temperature            300

[cut-offs]
solute_solute                  10
solvent_solvent                10
solute_solvent                 10
q_atom                         99

[sphere]

[intervals]
non_bond                       30
output                         100
energy                         10

[group_contribution]
residue all 153
residue electro 153
residue vdw 153
residue full 234
residue all 153 234
atom full 5124 1624 1800
atom electro 4719 4732 2000
atom vdw 5124 1631

[files]
topology                       2cjpFH_ionres_oplsa.top
final                          final.re
energy                         run.en
fep                            lig.fep

[lambdas]
0.5 0.5
"""


def run_cjp():
    tmp = tempfile.mkdtemp(prefix="qnbgold_")
    try:
        for f in ("2cjpFH_ionres_oplsa.top", "lig.fep"):
            shutil.copy(os.path.join(EXCL, f), tmp)
        with open(os.path.join(tmp, "run.inp"), "w") as f:
            f.write(CJP_INP)
        log = subprocess.run([DUMP, "out.bin", "nomd", "run.inp"], cwd=tmp, capture_output=True, text=True)
        if log.returncode != 0 or not os.path.exists(os.path.join(tmp, "out.bin")):
            print(log.stdout[-3000:], log.stderr[-3000:])
            raise RuntimeError("qnb_dump failed for cjp_excl")
        return read_dump(os.path.join(tmp, "out.bin"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_synth(name, system):
    """A synthetic system written as Q files (q6_amd.qfiles) and set up by
    the dump driver (no MD step)."""
    from q6_amd import qfiles
    tmp = tempfile.mkdtemp(prefix="qnbgold_")
    try:
        inp = qfiles.write_case(system, tmp, 10, name=name, seed=0)
        log = subprocess.run([DUMP, "out.bin", "nomd", os.path.basename(inp)], cwd=tmp, capture_output=True,
                             text=True)
        if log.returncode != 0 or not os.path.exists(os.path.join(tmp, "out.bin")):
            print(log.stdout[-3000:], log.stderr[-3000:])
            raise RuntimeError("qnb_dump failed for " + name)
        return read_dump(os.path.join(tmp, "out.bin"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    if not os.path.exists(DUMP):
        sys.exit("build the reference first: make -C oracle/ref")
    q = top_charges()
    cases = {
        "lig_eq1": (EQ1_SPH.format(lrf="on", rq=99, extra_cut="", fep="lig_w.fep", lambdas=""), None),
        "lig_fep2": (EQ1_SPH.format(lrf="on", rq=12,
                                    extra_cut="solute_solute 16\nsolvent_solvent 9\nsolute_solvent 16\n",
                                    fep="fep2.fep", lambdas="[lambdas]\n0.7 0.3\n"), fep2_text(q)),
        "lig_fep2_nolrf": (EQ1_SPH.format(lrf="off", rq=12,
                                          extra_cut="solute_solute 16\nsolvent_solvent 9\nsolute_solvent 16\n",
                                          fep="fep2.fep", lambdas="[lambdas]\n0.7 0.3\n"), fep2_text(q)),
    }
    cases["lig_qb2"] = (cases["lig_fep2"][0], fep2_text(q) + QB_TEXT)
    cases["lig_fep2_arith"] = (cases["lig_fep2"][0], arith_fep(fep2_text(q)))
    top_ref = open(os.path.join(REF, "prep_SPH", "lig_w.top")).read()
    top_arith = arith_top(top_ref)
    top_any = anyatom_top(top_ref)
    cases["lig_fep2_anyatom"] = cases["lig_fep2"]
    cases["lig_fep2_anyatom_nolrf"] = cases["lig_fep2_nolrf"]
    tops = {"lig_fep2_arith": top_arith, "lig_fep2_anyatom": top_any, "lig_fep2_anyatom_nolrf": top_any}
    only = sys.argv[1:]
    if not only or "cjp_excl" in only:
        d = run_cjp()
        arrs = compact(d, fill=False, list_params=False)
        path = os.path.join(HERE, "cjp_excl.npz")
        np.savez_compressed(path, **arrs)
        print("cjp_excl ->", path, os.path.getsize(path) // 1024, "KiB;", "pp", d["nbpp_ij"].shape[1], "pw",
              d["nbpw_ij"].shape[1], "qp", d["nbqp_ij"].shape[1], "arrays", len(d["exc_type"]) + 1)
    if not only or "eoh" in only:
        # a nine-site flexible solvent (q6_amd.synth.cfg_ethanol): the general
        # solvent loops and nonbond_solvent_internal
        sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
        from q6_amd import synth
        d = run_synth("eoh", synth.cfg_ethanol())
        arrs = compact(d, fill=False)
        path = os.path.join(HERE, "eoh.npz")
        np.savez_compressed(path, **arrs)
        print("eoh ->", path, os.path.getsize(path) // 1024, "KiB;", "pw", d["nbpw_ij"].shape[1], "ww",
              d["nbww_ij"].shape[1], "qw", d["nbqw_ij"].shape[1], "solvent internal pairs", d["solv_int_ij"].shape[1])
    if not only or "eoh_pol" in only:
        # the same solvent with the polarisation restraint: watpol's
        # charge-weighted dipole about the molecule's centroid
        sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
        from q6_amd import synth
        s = synth.cfg_ethanol()
        s.par["wpol_restr"] = 1
        d = run_synth("eoh_pol", s)
        arrs = compact(d, fill=False)
        path = os.path.join(HERE, "eoh_pol.npz")
        np.savez_compressed(path, **arrs)
        print("eoh_pol ->", path, os.path.getsize(path) // 1024, "KiB;", "shells", d["nwpolr_shell"],
              "in shells", d["wshell_n_insh"])
    for name, (inp, fep) in cases.items():
        if only and name not in only:
            continue
        d, log = run_case(name, inp, fep, tops.get(name))
        if name == "lig_qb2":
            keep = {k: np.asarray(v) for k, v in d.items()
                    if k in QB_KEYS or k[:4] in ("E_qb", "d_qb") or k.startswith(("EMorseD", "dMorse_i"))}
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **keep)
            print(name, "->", path, os.path.getsize(path) // 1024, "KiB;", "counts", d["qb_counts"])
            continue
        if name == "lig_eq1":
            # the reference's own known answer (tests/basic_tests/eval_test.sh:97-101)
            el, vdw = d["E_qw"]
            assert "%.2f %.2f" % (el, vdw) == "3.12 139.43", (el, vdw)
        arrs = compact(d, fill=False)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrs)
        print(name, "->", path, os.path.getsize(path) // 1024, "KiB;",
              "pp", d["nbpp_ij"].shape[1], "pw", d["nbpw_ij"].shape[1], "ww", d["nbww_ij"].shape[1],
              "qp", d["nbqp_ij"].shape[1], "qw", d["nbqw_ij"].shape[1])


if __name__ == "__main__":
    main()
