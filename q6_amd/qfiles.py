"""Q input files for a system built by q6_amd.synth: topology (.top), FEP
file (.fep) and MD input (.inp), in the layouts stock Qdyn6 reads.

  * topology: the record order and header of topo_save / topo_read
    (src/topo.f90:1117-1425 / 522-1114): header block, atom count line,
    coordinates, atom type codes, bonds with their codes and parameters,
    angles, torsions, impropers, charges, charge groups (size, switch atom,
    atoms), atom types (masses and sqrt(A), sqrt(B) for the normal, polar and
    1-4 codes), type-2 pairs, the 1-4 and exclusion windows (25 flags per
    solute atom, 80 digits per line) with their long-range lists, residues,
    molecules, type names, solvent type and the sphere (radii, density,
    centres, excluded atoms);
  * FEP file: [FEP] states, [atoms], [change_charges] and the EVB
    [off_diagonals] (src/qatom.f90:639-1991); Q-atoms keep their topology
    types, as the synthetic systems do;
  * MD input: [MD], [cut-offs], [sphere], [solvent], [intervals], [files],
    [lambdas] and the switch-atom distance restraints of compact groups
    (src/md.f90:113-1369).
Numbers are written with 17 significant digits (Qdyn6 reads them list-
directed), so Qdyn6 sets up exactly the tables the generator built.
"""
import os

import numpy as np


def _g(v):
    return repr(float(v))


def _rows(vals, per, fmt=_g):
    vals = list(vals)
    return [" ".join(fmt(v) for v in vals[i:i + per]) for i in range(0, len(vals), per)] or []


def _flags80(flags):
    s = "".join("1" if f else "0" for f in flags)
    return [s[i:i + 80] for i in range(0, len(s), 80)]


def write_top(s, path, title="synthetic"):
    """Topology of a synthetic system (needs s.md from q6_amd.synth.build)."""
    n, a, md = s.n, s.a, s.md
    nat, nsol, nwat = int(n["natom"]), int(n["nat_solute"]), int(n["nwat"])
    ntypes = int(n["ntypes"])
    L = ["Q topology file", "TITLE      " + title, "DATE       0", "VERSION    5.06", "FORCEFIELD synthetic",
         "FF_TYPE    1", "PRM_FILE   none", "END        of header"]
    mol = md.get("solvent")  # a general solvent molecule (synth.replace_solvent), else TIP3P
    sa = int(n["solv_atom"])
    L.append("%d %d %d %d = Total no. of atoms, no. of solute atoms, atoms per solvent molecule, solvent dielectric"
             % (nat, nsol, sa, 80000))
    L += _rows(np.asarray(s.x).ravel(), 6)
    L.append("%d = No. of integer atom codes. iac's:" % nat)
    L += _rows(np.asarray(a["iac"]) + 1, 16, str)
    # bonds: solute chains (code 1), water O-H (2) and H-H (3)
    bonds = np.asarray(md["bonds"], int).reshape(-1, 2)
    bl = [(i + 1, j + 1, 1) for i, j in bonds]
    if mol is None:
        for w in range(nwat):
            o = nsol + 3 * w + 1
            bl += [(o, o + 1, 2), (o, o + 2, 2), (o + 1, o + 2, 3)]
        bpar = [(450.0, md["water_roh"]), (450.0, md["water_rhh"])]
    else:
        for w in range(nwat):
            o = nsol + sa * w + 1
            bl += [(o + i, o + j, 2 + c) for i, j, c in mol["bonds"]]
        bpar = mol["bond_par"]
    L.append("%d %d = No. of bonds, no. of solute bonds. i - j - icode:" % (len(bl), len(bonds)))
    L += _rows([v for b in bl for v in b], 15, str)
    L.append("%d = No. of bond codes. Parameters:" % (1 + len(bpar)))
    L.append("1 %s %s" % (_g(md["bond_k"]), _g(md["bond_r0"])))
    L += ["%d %s %s" % (2 + c, _g(k), _g(r0)) for c, (k, r0) in enumerate(bpar)]
    angles = np.asarray(md["angles"], int).reshape(-1, 3)
    al = [(i + 1, j + 1, k + 1, 1) for i, j, k in angles]
    if mol is None:
        for w in range(nwat):
            o = nsol + 3 * w + 1
            al.append((o + 1, o, o + 2, 2))
        apar = [(55.0, 104.52)]
    else:
        for w in range(nwat):
            o = nsol + sa * w + 1
            al += [(o + i, o + j, o + k, 2 + c) for i, j, k, c in mol["angles"]]
        apar = mol["angle_par"]
    L.append("%d %d = No. of angles, no. of solute angles. i - j - k - icode:" % (len(al), len(angles)))
    L += _rows([v for t in al for v in t], 12, str)
    L.append("%d = No. of angle codes. Parameters:" % (1 + len(apar)))
    L.append("1 %s %s" % (_g(md["angle_k"]), _g(np.degrees(md["angle_t0"]))))
    L += ["%d %s %s" % (2 + c, _g(k), _g(t0)) for c, (k, t0) in enumerate(apar)]
    tl, tpar = [], []
    if mol is not None:
        for w in range(nwat):
            o = nsol + sa * w + 1
            tl += [(o + i, o + j, o + k, o + l, 1 + c) for i, j, k, l, c in mol["torsions"]]
        tpar = mol["tor_par"]
    L.append("%d 0 = No. of torsions, solute torsions." % len(tl))
    L += _rows([v for t in tl for v in t], 10, str)
    L.append("%d = No. of torsion codes. Parameters:" % len(tpar))
    L += ["%d %s %s %s %d" % (1 + c, _g(fk), _g(rm), _g(dt), int(pth)) for c, (fk, rm, dt, pth) in enumerate(tpar)]
    L.append("0 0 = No. of impropers, solute impr.")
    L.append("0 1 = No. of improper codes, type (1=harmonic,2=periodic). Parameters:")
    L.append("%d = No. of atomic charges" % nat)
    L += _rows(md["crg_unscaled"], 10)
    ncgp, ncs = int(n["ncgp"]), int(n["ncgp_solute"])
    L.append("%d %d 1 = No. of charge groups, no of solute cgps, switch atoms flag. nat_cgp, iswitch / atom list:"
             % (ncgp, ncs))
    first, last, sw, cgpatom = a["cgp_first"], a["cgp_last"], a["cgp_iswitch"], a["cgpatom"]
    for g in range(ncgp):
        L.append("%d %d" % (last[g] - first[g] + 1, sw[g] + 1))
        L += _rows(np.asarray(cgpatom[first[g]:last[g] + 1]) + 1, 13, str)
    L.append("%d = No. of atom types" % ntypes)
    L.append("1 = vdW combination rule (1 = Geom. / 2 = Arit.)")
    L.append("%s %s = Electrostatic 1-4 scaling factor and Coulomb constant" % (_g(s.el14_scale), _g(332.0)))
    L.append("Masses:")
    L += _rows(md["type_mass"], 10)
    avdw = np.asarray(a["avdw"]).reshape(ntypes, 3)
    bvdw = np.asarray(a["bvdw"]).reshape(ntypes, 3)
    for k, name in enumerate(("normal", "polar", "1-4")):
        L.append("sqrt (Aii) %s:" % name)
        L += _rows(avdw[:, k], 8)
        L.append("sqrt (Bii) %s:" % name)
        L += _rows(bvdw[:, k], 8)
    L.append("0 = No. of type-2 vdW interactions. pairs:")
    l14 = np.asarray(a["list14"]).reshape(-1, 25)[:nsol]
    lex = np.asarray(a["listex"]).reshape(-1, 25)[:nsol]
    n14long, nexlong = int(n.get("n14long", 0)), int(n.get("nexlong", 0))
    L.append("%d = No. of 1-4 neighbours. nborlist (range=max_nbr_range):" % (int(l14.sum()) + n14long))
    L += _flags80(l14.ravel())
    L.append("%d = No. of long 1-4 nbrs (>max_nbr_range). pairlist:" % n14long)
    p14 = np.asarray(a["list14long"]).ravel()[:2 * n14long].reshape(-1, 2)
    L += ["%d %d" % (i + 1, j + 1) for i, j in p14]
    L.append("%d = No. of exclusions. exclusion list (range=max_nbr_range):" % (int(lex.sum()) + nexlong))
    L += _flags80(lex.ravel())
    L.append("%d = No. of long exclusions (>max_nbr_range). pairlist:" % nexlong)
    pex = np.asarray(a["listexlong"]).ravel()[:2 * nexlong].reshape(-1, 2)
    L += ["%d %d" % (i + 1, j + 1) for i, j in pex]
    # residues and molecules: one per chain (charge group) and one per water
    starts = [int(cgpatom[first[g]]) + 1 for g in range(ncs)] + [nsol + sa * w + 1 for w in range(nwat)]
    L.append("%d %d = No. of residues, No of solute residues. start atoms:" % (len(starts), ncs))
    L += _rows(starts, 13, str)
    L.append("Sequence:")
    names = ["CHN"] * ncs + [("HOH" if mol is None else mol["name"])] * nwat
    L += [" ".join("%-4s" % v for v in names[i:i + 16]) for i in range(0, len(names), 16)]
    L.append("%d = No. of separate molecules. start atoms:" % len(starts))
    L += _rows(starts, 13, str)
    tn = md["type_names"]
    L.append("%d = No. of atom types:" % ntypes)
    L += [" ".join("%-8s" % v for v in tn[i:i + 8]) for i in range(0, ntypes, 8)]
    L.append("%d = No. of SYBYL atom types:" % ntypes)
    L += [" ".join("%-5s" % v for v in ["O.3", "H"] + ["C.3"] * (ntypes - 2))]
    L.append("%d = solvent type (0=SPC,1=3-atom,2=general)" % (0 if mol is None else 1))
    p = s.par
    L.append("%s %s %s = Exclusion, solvent radii, Solvent density"
             % (_g(p["rexcl_o"]), _g(p["rwat"]), _g(p.get("solvent_density", 0.0334))))
    L.append("%s %s %s = Solute centre" % tuple(_g(v) for v in p["xpcent"]))
    L.append("%s %s %s = Solvent centre" % tuple(_g(v) for v in p["xwcent"]))
    excl = np.asarray(a["excl"])[:nat]
    L.append("%d %d = No. of excluded atoms (incl. water), no. of excluded waters" % (int(excl.sum()), 0))
    flags = "".join("T" if e else "F" for e in excl)
    L += [flags[i:i + 80] for i in range(0, len(flags), 80)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_fep(s, path, offdiag=None):
    """FEP file: Q-atoms, their charges per state (unscaled) and optional EVB
    off-diagonals [(state_i, state_j, q_k, q_l, A, mu)] with 1-based numbers."""
    md = s.md
    iqseq = np.asarray(s.a["iqseq"])[:int(s.n["nqat"])]
    ns = int(s.n["nstates"])
    L = ["[FEP]", "states %d" % ns, "", "[atoms]"]
    L += ["%d %d" % (k + 1, i + 1) for k, i in enumerate(iqseq)]
    L += ["", "[change_charges]"]
    for k in range(len(iqseq)):
        L.append("%d " % (k + 1) + " ".join(_g(v) for v in md["qcrg"][k, :ns]))
    if offdiag:
        L += ["", "[off_diagonals]"]
        L += [" ".join(str(v) if isinstance(v, (int, np.integer)) else _g(v) for v in od) for od in offdiag]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_inp(s, path, top, fep, steps, lambdas=None, dt_fs=1.0, T=300.0, tau_T=100.0, seed=4, nbcycle=10,
              energy_every=0, final="final.re", energy_file=None, group_contribution=None):
    """MD input for the synthetic system's cut-offs, sphere and restraints.
    group_contribution: lines of the [group_contribution] section
    (src/md.f90:1011-1070), e.g. "residue full 3" (residues are the solute
    chains, then the waters); Qdyn6 computes them only with an energy
    interval and file."""
    p, md = s.par, s.md
    lam = np.asarray(lambdas if lambdas is not None else p["lambda"], float)
    L = ["[MD]", "steps %d" % steps, "stepsize %s" % _g(dt_fs), "temperature %s" % _g(T),
         "bath_coupling %s" % _g(tau_T), "separate_scaling on", "random_seed %d" % seed,
         "initial_temperature %s" % _g(T), "lrf %s" % ("on" if p["use_LRF"] else "off"), "shake_solvent on",
         "shake_solute off", "",
         "[cut-offs]", "solute_solute %s" % _g(p["Rcpp"]), "solvent_solvent %s" % _g(p["Rcww"]),
         "solute_solvent %s" % _g(p["Rcpw"]), "q_atom %s" % _g(p["Rcq"]), "lrf %s" % _g(p["RcLRF"]), "",
         "[sphere]", "shell_force 10.0", "",
         "[solvent]", "radial_force %s" % _g(p["fk_wsphere"]),
         "polarisation %s" % ("on" if p.get("wpol_restr", 1) else "off"),
         "polarisation_force %s" % _g(p["fkwpol"]), "pol_update 100", "",
         "[intervals]", "non_bond %d" % nbcycle, "output %d" % max(steps, 1)]
    if energy_every:
        L.append("energy %d" % energy_every)
    L += ["", "[files]", "topology %s" % top, "fep %s" % fep, "final %s" % final]
    if energy_file:
        L.append("energy %s" % energy_file)
    if group_contribution:
        L += ["", "[group_contribution]"] + list(group_contribution)
    L += ["", "[lambdas]", " ".join(_g(v) for v in lam)]
    rst = np.asarray(md.get("restraints", np.zeros((0, 2))), int).reshape(-1, 2)
    if len(rst):
        r0 = np.broadcast_to(md["restraint_r0"], len(rst))
        L += ["", "[distance_restraints]"]
        L += ["%d %d %s %s %s 0" % (i + 1, j + 1, _g(d), _g(d), _g(md["restraint_k"])) for (i, j), d in zip(rst, r0)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_case(s, directory, steps, name="sys", **kw):
    """Writes name.top, name.fep and name.inp into directory; returns the
    input file's path."""
    os.makedirs(directory, exist_ok=True)
    offd = None
    if int(s.n["nstates"]) >= 2 and int(s.n["nqat"]) >= 2:
        offd = [(1, 2, 1, 2, 80.0, 0.0)]
    write_top(s, os.path.join(directory, name + ".top"), title=getattr(s, "name", name)[:60] or name)
    write_fep(s, os.path.join(directory, name + ".fep"), offdiag=offd)
    inp = os.path.join(directory, name + ".inp")
    write_inp(s, inp, name + ".top", name + ".fep", steps, **kw)
    return inp
