"""Seeded synthetic spherical systems for the benchmark workloads.

Builds, directly in the engine's data layout (q6_amd.system.QSystem), the
systems the workloads are defined on: a sphere of rigid TIP3P water on a
jittered cubic lattice at 0.0334 molecules/A^3 with random orientations, around
a solute of self-avoiding bead chains (one neutral charge group per chain),
with Q-atoms in 2-3 FEP/EVB states.  Everything Qdyn6's set-up would derive is
derived the same way here:
  * charges scaled by sqrt(332) (prep_sim, src/simprep.f90:3714-3721);
  * geometric LJ A_i = 2 sqrt(eps) sigma^6, B_i = 2 sqrt(eps) sigma^3, pair
    products as precompute_set_values_* (src/simprep.f90:3321-3589);
  * exclusions 1-2 and 1-3, 1-4 list from the bond graph (as Qprep6 writes
    listex/list14), Q connectivity levels (make_qconn, src/nonbondene.f90:3087);
  * SCAAS constants: Dwmz, awmz (src/simprep.f90:4667-4672), polarisation
    shells and cstb (src/simprep.f90:4707-4891).
Exact atom and pair counts are printed by `describe`.
"""
import math

import numpy as np

from .system import QSystem

SQRT332 = math.sqrt(332.0)
RHO_WAT = 0.0334
BOLTZ = 0.001986  # src/globals.f90:46
# TIP3P (ff/oplsaa/qoplsaa.prm:41,69, qoplsaa.lib:1484-1497 of the reference)
TIP3P_Q = (-0.834, 0.417, 0.417)
TIP3P_ROH = 0.9572
TIP3P_HOH = 104.52
OW_A, OW_B = 762.89, 24.39
M_O, M_H = 15.9994, 1.008


def _lj(sigma, eps):
    return 2.0 * math.sqrt(eps) * sigma ** 6, 2.0 * math.sqrt(eps) * sigma ** 3


def _rand_rot(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


class _Grid:
    """Spatial hash for rejection tests."""

    def __init__(self, cell):
        self.cell = cell
        self.d = {}

    def _key(self, p):
        return tuple(np.floor(p / self.cell).astype(int))

    def add(self, p, idx):
        self.d.setdefault(self._key(p), []).append((idx, p))

    def near(self, p, r, skip=()):
        k = self._key(p)
        r2 = r * r
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for idx, q in self.d.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if idx in skip:
                            continue
                        t = p - q
                        if t @ t < r2:
                            return True
        return False


def bead_chains(rng, nchain, nbead, r_region, bond=1.5, angle=109.5, dmin=2.2, max_region=60.0, compact=None):
    """Self-avoiding bead chains grown inside a sphere.  Non-bonded beads
    (beyond 1-3) stay at least dmin apart.  When a chain cannot be placed the
    region radius grows by 0.5 A (the returned radius says what was used).
    With `compact`, every bead of a chain stays within that distance of the
    chain's first bead (a folded chain: a compact charge group)."""
    grid = _Grid(max(dmin, 2.0))
    pos = np.zeros((nchain * nbead, 3))
    cos_t = math.cos(math.radians(angle))
    sin_t = math.sin(math.radians(angle))
    R = r_region
    for c in range(nchain):
        placed = False
        for attempt in range(400):
            if attempt and attempt % 100 == 0:
                R += 0.5
            chain = []
            p0 = rng.uniform(-R, R, 3)
            while p0 @ p0 > R * R:
                p0 = rng.uniform(-R, R, 3)
            if grid.near(p0, dmin):
                continue
            chain.append(p0)
            ok = True
            for b in range(1, nbead):
                good = False
                for _ in range(60):
                    if b == 1:
                        u = rng.normal(size=3)
                        u /= np.linalg.norm(u)
                    else:
                        prev = chain[-1] - chain[-2]
                        prev /= np.linalg.norm(prev)
                        a = rng.normal(size=3)
                        a -= (a @ prev) * prev
                        a /= np.linalg.norm(a)
                        # the new bond makes the bond angle `angle` with the previous one
                        u = -cos_t * prev + sin_t * a
                    p = chain[-1] + bond * u
                    if p @ p > R * R:
                        continue
                    if compact is not None and (p - chain[0]) @ (p - chain[0]) > compact * compact:
                        continue
                    # self-avoidance against everything except the last two beads of this chain
                    if grid.near(p, dmin):
                        continue
                    if len(chain) >= 3 and any((p - q) @ (p - q) < dmin * dmin for q in chain[:-2]):
                        continue
                    chain.append(p)
                    good = True
                    break
                if not good:
                    ok = False
                    break
            if ok:
                for b, p in enumerate(chain):
                    grid.add(p, c * nbead + b)
                pos[c * nbead:(c + 1) * nbead] = np.array(chain)
                placed = True
                break
        if not placed:
            raise RuntimeError("could not place chain %d" % c)
        if R > max_region:
            raise RuntimeError("solute region exceeded %.1f A" % max_region)
    return pos, R


def tip3p_geometry():
    th = math.radians(TIP3P_HOH) / 2
    return np.array([[0.0, 0.0, 0.0],
                     [TIP3P_ROH * math.sin(th), 0.0, TIP3P_ROH * math.cos(th)],
                     [-TIP3P_ROH * math.sin(th), 0.0, TIP3P_ROH * math.cos(th)]])


def water_sphere(rng_lat, rng_rot, radius, solute_pos, jitter=0.3, dmin_o=2.4, dmin_h=1.8):
    """Water centres on a jittered cubic lattice at RHO_WAT within radius,
    random orientations, dropped where they overlap the solute."""
    a = RHO_WAT ** (-1.0 / 3.0)
    n = int(math.ceil(radius / a)) + 1
    g = np.arange(-n, n + 1) * a
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    cen = cen + rng_lat.uniform(-jitter, jitter, cen.shape)
    cen = cen[np.einsum("ij,ij->i", cen, cen) <= radius * radius]
    rot = _rand_rot(rng_rot, len(cen))
    geo = tip3p_geometry()
    at = cen[:, None, :] + np.einsum("nij,kj->nki", rot, geo)
    if len(solute_pos):
        grid = _Grid(3.0)
        for k, p in enumerate(solute_pos):
            grid.add(p, k)
        keep = np.array([not (grid.near(w[0], dmin_o) or grid.near(w[1], dmin_h) or grid.near(w[2], dmin_h))
                         for w in at])
        at = at[keep]
    return at


def _exclusions(nat, bonds):
    """1-2/1-3 exclusions and 1-4 pairs from the bond graph."""
    adj = [[] for _ in range(nat)]
    for i, j in bonds:
        adj[i].append(j)
        adj[j].append(i)
    ex, p14 = set(), set()
    for s in range(nat):
        dist = {s: 0}
        frontier = [s]
        for lev in range(1, 4):
            nxt = []
            for u in frontier:
                for w in adj[u]:
                    if w not in dist:
                        dist[w] = lev
                        nxt.append(w)
            frontier = nxt
        for t, lev in dist.items():
            if t <= s:
                continue
            if lev in (1, 2):
                ex.add((s, t))
            elif lev == 3:
                p14.add((s, t))
    return ex, p14, adj


def shell_cstb(qcrg, lam, wshell_rout, wshell_dr):
    """Born-correction constant cstb of each polarisation shell as wat_shells
    forms it (src/simprep.f90:4863-4883) for TIP3P solvent:
      * crgQtot from the FEP file's (unscaled) Q charges, Q-atoms outer and
        states inner (wat_sphere, src/simprep.f90:4640-4646);
      * the water dipole mu_w from the oxygen's charge, the length of the
        topology's LAST bond and the cosine of half its last angle
        (:4745-4751); water bonds are listed O-H1, O-H2, H1-H2, so that
        length is the H-H distance;
      * cstb = crgQtot (1 - 1/eps) / (rho mu_w 4 pi rshell^2), rshell the
        radius that halves the shell's volume."""
    qcrg = np.asarray(qcrg, float).reshape(len(qcrg), -1) if np.size(qcrg) else np.zeros((0, 1))
    lam = np.asarray(lam, float)
    crgQtot = 0.0
    for i in range(qcrg.shape[0]):
        for st in range(qcrg.shape[1]):
            crgQtot = crgQtot + float(qcrg[i, st]) * float(lam[st])
    rhh = 2 * TIP3P_ROH * math.sin(math.radians(TIP3P_HOH) / 2)
    mu_w = -TIP3P_Q[0] * rhh * math.cos(math.radians(TIP3P_HOH) / 2)
    eps_diel = 1.0 - 1.0 / 80.0
    out = []
    for rout, dr in zip(wshell_rout, wshell_dr):
        ri = rout - dr
        rshell = (0.5 * (rout ** 3 + ri ** 3)) ** (1.0 / 3.0)
        out.append(crgQtot * eps_diel / (RHO_WAT * mu_w * 4.0 * math.pi * rshell ** 2))
    return np.array(out)


def build(radius, nchain, nbead, n_q_chains, nstates, seed, chain_region=None, q_sigma=0.3,
          cut=10.0, rcq=99.0, rclrf=99.0, use_lrf=True, lambdas=None, q_state0_zero=False,
          name="synthetic", q_per_chain=None, compact=None, solute_dmin=2.2, water_dmin=(2.4, 1.8),
          water_fill=None):
    """A sphere of `radius` A: `nchain` neutral bead chains of `nbead` atoms (one
    charge group each) in the central region, the first `n_q_chains` of them
    (the ones closest to the centre) made of Q-atoms with `nstates` states,
    and TIP3P water filling the rest.  With q_per_chain, only the first
    q_per_chain beads of those chains are Q-atoms (so Q-atoms have bonded
    non-Q neighbours: Q-bonded-neighbour and 1-4 Q-solute pairs).  With
    `compact`, each chain is folded within that radius of its first bead and
    its switch atom is the bead with the smallest largest distance to the
    others, so every atom of a group lies within `compact` of the switch atom
    and no atom pair in contact is missed by the switch-atom cut-off.
    `solute_dmin` is the smallest distance between non-bonded beads as the
    chains are grown, `water_dmin` the smallest O and H distance from a kept
    water to any solute atom, and `water_fill` the radius within which water
    centres are kept (default radius - 0.5)."""
    rng = np.random.default_rng(seed)
    rng_lat = np.random.default_rng(seed + 1)
    rng_rot = np.random.default_rng(seed + 2)
    rng_q = np.random.default_rng(seed + 3)
    # bead types: 6 LJ types, sigma in [3.0, 3.8], eps in [0.05, 0.2]
    sig = np.linspace(3.0, 3.8, 6)
    eps = np.linspace(0.05, 0.2, 6)[::-1]
    bead_masses = [12.011, 14.007, 15.999, 12.011, 32.06, 14.007]
    if nchain:
        spos, r_used = bead_chains(rng, nchain, nbead, chain_region or radius * 0.6, dmin=solute_dmin,
                                   compact=compact)
    else:
        spos, r_used = np.zeros((0, 3)), 0.0
    # Q chains: the chains whose centres are closest to the sphere centre
    if nchain:
        cc = spos.reshape(nchain, nbead, 3).mean(1)
        order = np.argsort(np.einsum("ij,ij->i", cc, cc), kind="stable")
        spos = spos.reshape(nchain, nbead, 3)[order].reshape(-1, 3)
    nsol = nchain * nbead
    wat = water_sphere(rng_lat, rng_rot, radius - 0.5 if water_fill is None else water_fill, spos,
                       dmin_o=water_dmin[0], dmin_h=water_dmin[1])
    nwat = len(wat)
    nat = nsol + 3 * nwat
    x = np.concatenate([spos, wat.reshape(-1, 3)]) if nwat else spos.copy()

    s = QSystem()
    s.name = name
    ntypes = 2 + 6
    avdw = np.zeros((ntypes, 3))
    bvdw = np.zeros((ntypes, 3))
    mass_t = np.zeros(ntypes)
    avdw[0, :] = OW_A
    bvdw[0, :] = OW_B
    mass_t[0], mass_t[1] = M_O, M_H
    for t in range(6):
        A, B = _lj(sig[t], eps[t])
        avdw[2 + t, :] = A
        bvdw[2 + t, :] = B
        mass_t[2 + t] = bead_masses[t]
    iac = np.zeros(nat, np.int32)
    iac[:nsol] = 2 + rng.integers(0, 6, nsol)
    iac[nsol::3] = 0
    iac[nsol + 1::3] = 1
    iac[nsol + 2::3] = 1
    crg = np.zeros(nat)
    for c in range(nchain):
        q = rng.normal(0.0, q_sigma, nbead)
        crg[c * nbead:(c + 1) * nbead] = q - q.mean()
    crg[nsol::3] = TIP3P_Q[0]
    crg[nsol + 1::3] = TIP3P_Q[1]
    crg[nsol + 2::3] = TIP3P_Q[2]
    crg_unscaled = crg.copy()
    crg = crg * SQRT332

    # bonds, angles along each chain; water geometry constrained
    bonds, angles = [], []
    for c in range(nchain):
        o = c * nbead
        bonds += [(o + k, o + k + 1) for k in range(nbead - 1)]
        angles += [(o + k, o + k + 1, o + k + 2) for k in range(nbead - 2)]
    ex, p14, adj = _exclusions(nsol, bonds)
    listex = np.zeros((max(nsol, 1), 25), np.uint8)
    list14 = np.zeros((max(nsol, 1), 25), np.uint8)
    exlong, l14long = [], []
    for (i, j) in ex:
        if j - i <= 25:
            listex[i, j - i - 1] = 1
        else:
            exlong.append((i, j))
    for (i, j) in p14:
        if j - i <= 25:
            list14[i, j - i - 1] = 1
        else:
            l14long.append((i, j))

    # charge groups: one per chain (switch atom in the middle), one per water
    ncgp = nchain + nwat
    first = np.zeros(ncgp, np.int32)
    last = np.zeros(ncgp, np.int32)
    sw = np.zeros(ncgp, np.int32)
    for c in range(nchain):
        first[c], last[c], sw[c] = c * nbead, c * nbead + nbead - 1, c * nbead + nbead // 2
        if compact is not None:
            ch = x[c * nbead:(c + 1) * nbead]
            far = np.sqrt(((ch[:, None, :] - ch[None, :, :]) ** 2).sum(-1)).max(1)
            sw[c] = c * nbead + int(np.argmin(far))
    for w in range(nwat):
        g = nchain + w
        first[g], last[g], sw[g] = nsol + 3 * w, nsol + 3 * w + 2, nsol + 3 * w
    cgpatom = np.arange(nat, dtype=np.int32)

    # Q-atoms
    qpc = nbead if q_per_chain is None else int(q_per_chain)
    nq = n_q_chains * qpc
    iqseq = np.array([c * nbead + k for c in range(n_q_chains) for k in range(qpc)], np.int32)
    iqatom = np.zeros(nat, np.int32)
    iqatom[iqseq] = np.arange(1, nq + 1)
    qcrg = np.zeros((nq, nstates))
    for st in range(nstates):
        for c in range(n_q_chains):
            sl = slice(c * qpc, (c + 1) * qpc)
            if st == 0:
                qcrg[sl, 0] = crg_unscaled[iqseq[sl]]
            elif q_state0_zero:
                qcrg[sl, st] = 0.0
            else:
                q = rng_q.normal(0.0, q_sigma, qpc)
                qcrg[sl, st] = q - q.mean()
    if q_state0_zero:  # cfg1: state 1 charged, state 2 neutral
        pass
    qcrg_s = qcrg * SQRT332
    # qconn levels (make_qconn): 1 self, 2 bonded, 3 angle, 4 torsion, 9 far
    qconn = np.full((nstates, max(nsol, 1), max(nq, 1)), 9, np.int32)
    for iq in range(nq):
        i = iqseq[iq]
        dist = {i: 1}
        frontier = [i]
        for lev in (2, 3, 4):
            nxt = []
            for u in frontier:
                for w in adj[u]:
                    if w not in dist:
                        dist[w] = lev
                        nxt.append(w)
            frontier = nxt
        for t, lev in dist.items():
            qconn[:, t, iq] = lev
    qconn3 = (qconn[:, :nsol, :] <= 3).any(axis=(0, 2)).astype(np.uint8) if nq else np.zeros(nsol, np.uint8)

    sa = 3
    aLJ_solv = np.array([avdw[0], avdw[1], avdw[1]])
    bLJ_solv = np.array([bvdw[0], bvdw[1], bvdw[1]])
    chg_solv = np.array(TIP3P_Q) * SQRT332
    # ww (precompute_set_values_ww, geometric, code ljcod=1)
    ww_A = np.zeros((sa, sa))
    ww_B = np.zeros((sa, sa))
    ww_el = np.zeros((sa, sa))
    for la in range(sa):
        for ka in range(sa):
            ww_A[la, ka] = aLJ_solv[la, 0] * aLJ_solv[ka, 0]
            ww_B[la, ka] = bLJ_solv[la, 0] * bLJ_solv[ka, 0]
            ww_el[la, ka] = chg_solv[la] * chg_solv[ka]
    # pw (precompute_set_values_pw)
    pw_A = np.zeros((max(nsol, 1), sa))
    pw_B = np.zeros((max(nsol, 1), sa))
    pw_el = np.zeros((max(nsol, 1), sa))
    for i in range(nsol):
        if iqatom[i]:
            continue
        for j in range(sa):
            pw_A[i, j] = avdw[iac[i], 0] * aLJ_solv[j, 0]
            pw_B[i, j] = bvdw[iac[i], 0] * bLJ_solv[j, 0]
            pw_el[i, j] = crg[i] * chg_solv[j]
    # Q-atom tables (precompute_set_values_qp/qq/qw) from the Q charges and
    # the Q-side LJ factors; qvdw_flag off: a Q-atom's factors are its type's
    qA = np.zeros((max(nq, 1), nstates, 3))
    qB = np.zeros_like(qA)
    for iq in range(nq):
        qA[iq, :, :] = avdw[iac[iqseq[iq]], :][None, :]
        qB[iq, :, :] = bvdw[iac[iqseq[iq]], :][None, :]
    qctx = dict(qconn=qconn, iqseq=iqseq, iqatom=iqatom, iac=iac, avdw=avdw, bvdw=bvdw, crg=crg,
                aLJ_solv=aLJ_solv, bLJ_solv=bLJ_solv, chg_solv=chg_solv, nsol=nsol, nstates=nstates, el14=0.5)
    qt = q_tables(qctx, qcrg, qA, qB)
    a = s.a
    a["crg"], a["iac"], a["iqatom"] = crg, iac, iqatom
    a["excl"] = np.zeros(nat, np.uint8)
    heavy = np.ones(nat, np.uint8)
    heavy[nsol + 1::3] = 0
    heavy[nsol + 2::3] = 0
    a["heavy"] = heavy
    a["cgp_first"], a["cgp_last"], a["cgp_iswitch"], a["cgpatom"] = first, last, sw, cgpatom
    a["avdw"], a["bvdw"] = avdw.ravel(), bvdw.ravel()
    a["ljcod"] = np.ones(ntypes * ntypes, np.int32)
    a["listex"], a["list14"] = listex.ravel(), list14.ravel()
    a["listexlong"] = np.array(exlong, np.int32).ravel()
    a["list14long"] = np.array(l14long, np.int32).ravel()
    s.n["nexlong"], s.n["n14long"] = len(exlong), len(l14long)
    a["pw_A"], a["pw_B"], a["pw_el"] = pw_A.ravel(order="F"), pw_B.ravel(order="F"), pw_el.ravel(order="F")
    a["ww_A"], a["ww_B"], a["ww_el"] = ww_A.ravel(order="F"), ww_B.ravel(order="F"), ww_el.ravel(order="F")
    a["iqseq"], a["qconn3"] = iqseq, qconn3
    a.update(qt)
    s.n.update(natom=nat, nat_solute=nsol, nwat=nwat, solv_atom=sa, ncgp=ncgp, ncgp_solute=nchain, nqat=nq,
               nstates=nstates, ntypes=ntypes, ivdw_rule=1, solvent_spc=1)
    s.el14_scale = 0.5
    s.x = x
    lam = np.asarray(lambdas if lambdas is not None else [1.0 / nstates] * nstates, float)
    # SCAAS (wat_sphere / wat_shells)
    rwat = radius
    Dwmz = 0.26 * math.exp(-0.19 * (rwat - 15.0)) + 0.74
    awmz = 0.2 / (1.0 + math.exp(0.4 * (rwat - 25.0))) + 0.3
    nsh = int(-0.5 + math.sqrt(2.0 * (3.0001 / 0.5) + 0.25))
    rout = rwat
    sh_rout, sh_dr = [], []
    for k in range(1, nsh + 1):
        dr = 0.5 * k
        sh_rout.append(rout)
        sh_dr.append(dr)
        rout -= dr
    sh_cstb = list(shell_cstb(qcrg, lam, sh_rout, sh_dr))
    s.par = dict(Rcpp=cut, Rcpw=cut, Rcww=cut, Rcq=rcq, RcLRF=rclrf, use_LRF=int(use_lrf), rexcl_o=radius,
                 rwat=rwat, fk_wsphere=60.0, Dwmz=Dwmz, awmz=awmz, Boltz=BOLTZ, fkwpol=20.0, itdis=100.0,
                 wpol_restr=1, nwpolr_shell=nsh, xpcent=np.zeros(3), xwcent=np.zeros(3),
                 wshell_rout=np.array(sh_rout), wshell_dr=np.array(sh_dr), wshell_cstb=np.array(sh_cstb),
                 wshell_theta_corr=np.zeros(nsh))
    s.par["lambda"] = lam
    # MD data: masses, bonded terms, constraints
    mass = mass_t[iac]
    # compact groups keep their shape through distance restraints from every
    # bead to the switch atom at the built distance (Qdyn6 [distance_restraints]
    # with d1 = d2, p_restrain src/bondene.f90:905-950): bead chains have no
    # torsions and would otherwise unfold beyond the switch-atom cut-off
    rst = []
    if compact is not None:
        for c in range(nchain):
            for k in range(first[c], last[c] + 1):
                if k != sw[c]:
                    rst.append((sw[c], k, float(np.linalg.norm(x[k] - x[sw[c]]))))
    s.md = dict(mass=mass, bonds=np.array(bonds, np.int32).reshape(-1, 2), bond_r0=1.5, bond_k=600.0,
                restraints=np.array([(i, j) for i, j, _ in rst], np.int32).reshape(-1, 2),
                restraint_r0=np.array([r for _, _, r in rst]), restraint_k=20.0,
                angles=np.array(angles, np.int32).reshape(-1, 3), angle_t0=math.radians(109.5), angle_k=80.0,
                water_roh=TIP3P_ROH, water_rhh=2 * TIP3P_ROH * math.sin(math.radians(TIP3P_HOH) / 2),
                chain_region=r_used, qcrg=qcrg, qA=qA, qB=qB, qctx=qctx,
                crg_unscaled=crg_unscaled, type_mass=mass_t.copy(), nchain=nchain, nbead=nbead,
                type_names=["OW", "HW"] + ["C%d" % (t + 1) for t in range(6)])
    # Q-atom LJ types (the FEP file's Q atom types): one per bead type that
    # carries a Q-atom, sigma/eps as the topology type's
    q_types = np.unique(iac[iqseq]) if nq else np.zeros(0, np.int32)
    s.md.update(q_types=q_types, q_type_of=np.searchsorted(q_types, iac[iqseq]) if nq else np.zeros(0, np.int64),
                q_sig=sig[q_types - 2] if nq else np.zeros(0), q_eps=eps[q_types - 2] if nq else np.zeros(0))
    s.Tfree = 300.0
    return s.finalize()


def q_tables(ctx, qcrg, qA, qB):
    """Q-atom precompute tables in the engine's layout (precompute_set_values_qp
    /qq/qw, src/simprep.f90:3420-3589) from the Q charges qcrg[nq, ns]
    (unscaled), the Q-side LJ factors qA/qB[nq, ns, code] and the Q
    connectivity levels (make_qconn: 1 self, 2 bonded, 3 angle, 4 1-4, 9 far).
    Pairs at level 4 use LJ code 3 and scaled electrostatics."""
    qconn, iqseq, iqatom, iac = ctx["qconn"], ctx["iqseq"], ctx["iqatom"], ctx["iac"]
    avdw, bvdw, crg = ctx["avdw"], ctx["bvdw"], ctx["crg"]
    nsol, ns, el14 = ctx["nsol"], ctx["nstates"], ctx["el14"]
    nq = len(iqseq)
    qs = np.asarray(qcrg, float) * SQRT332
    nq1, nsol1 = max(nq, 1), max(nsol, 1)
    # qp rows of non-Q solute atoms with qconn >= 4 (qp_int_comp, qq_int_comp)
    qp = {k: np.zeros((nsol1, nq1, ns)) for k in ("A", "B", "el")}
    qp_set = np.zeros((nsol1, nq1, ns), np.uint8)
    if nq and nsol:
        lev = np.transpose(qconn[:, :nsol, :nq], (1, 2, 0))            # (j, iq, s)
        code = np.where(lev == 4, 2, 0)
        nonq = (iqatom[:nsol] == 0)[:, None, None]
        ok = nonq & (lev >= 4)
        tj = iac[:nsol]
        Aj = avdw[tj][:, None, None, :]                                 # (j,1,1,code)
        Bj = bvdw[tj][:, None, None, :]
        Aq = np.take_along_axis(np.broadcast_to(qA[None, :nq], (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        Bq = np.take_along_axis(np.broadcast_to(qB[None, :nq], (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        Ajc = np.take_along_axis(np.broadcast_to(Aj, (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        Bjc = np.take_along_axis(np.broadcast_to(Bj, (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        A = Aq * Ajc
        B = Bq * Bjc
        el = qs[None, :nq, :] * crg[:nsol][:, None, None]
        setf = (np.abs(A) > 1e-10) | (np.abs(B) > 1e-10) | (np.abs(el) > 1e-10)
        el = np.where(lev == 4, el * el14, el)
        qp["A"][:nsol, :nq] = np.where(ok, A, 0.0)
        qp["B"][:nsol, :nq] = np.where(ok, B, 0.0)
        qp["el"][:nsol, :nq] = np.where(ok, el, 0.0)
        qp_set[:nsol, :nq] = (ok & setf).astype(np.uint8)
    # qq for pairs iq < jq with qconn >= 4
    qq = {k: np.zeros((nq1, nq1, ns)) for k in ("A", "B", "el")}
    qq_set = np.zeros((nq1, nq1, ns), np.uint8)
    for iq in range(nq - 1):
        for jq in range(iq + 1, nq):
            for st in range(ns):
                lv = qconn[st, iqseq[jq], iq]
                if lv < 4:
                    continue
                k = 2 if lv == 4 else 0
                A = qA[iq, st, k] * qA[jq, st, k]
                B = qB[iq, st, k] * qB[jq, st, k]
                el = qs[iq, st] * qs[jq, st]
                setf = abs(A) > 1e-10 or abs(B) > 1e-10 or abs(el) > 1e-10
                if lv == 4:
                    el *= el14
                qq["A"][iq, jq, st], qq["B"][iq, jq, st], qq["el"][iq, jq, st] = A, B, el
                qq_set[iq, jq, st] = setf
    # qw (precompute_set_values_qw): water atoms with code 1
    sa = len(ctx["chg_solv"])
    qw = {k: np.zeros((nq1, sa, ns)) for k in ("A", "B", "el")}
    if nq:
        qw["A"][:nq] = qA[:nq, None, :, 0] * ctx["aLJ_solv"][None, :, 0, None]
        qw["B"][:nq] = qB[:nq, None, :, 0] * ctx["bLJ_solv"][None, :, 0, None]
        qw["el"][:nq] = qs[:nq, None, :] * ctx["chg_solv"][None, :, None]
    out = {"qp_" + k: v.ravel(order="F") for k, v in qp.items()}
    out["qp_sc"], out["qp_set"] = np.zeros(qp["A"].size), qp_set.ravel(order="F")
    out.update({"qq_" + k: v.ravel(order="F") for k, v in qq.items()})
    out["qq_sc"], out["qq_set"] = np.zeros(qq["A"].size), qq_set.ravel(order="F")
    out["qq_soft"] = np.zeros(qq["A"].size, np.uint8)
    out.update({"qw_" + k: v.ravel(order="F") for k, v in qw.items()})
    out["qw_sc"] = np.zeros(qw["A"].size)
    return out


def q_tangents(ctx, qcrg, qA, qB):
    """Tangent tables of the Q tables (include/qnb.h qnb_qtangent): the
    derivative of every entry with respect to the Q-side parameter it
    contains -- the scaled Q charge (partner charge x el14 for 1-4 pairs) and
    the Q-atom's geometric LJ factors A, B (partner factor of the pair's code)."""
    qconn, iqseq, iqatom, iac = ctx["qconn"], ctx["iqseq"], ctx["iqatom"], ctx["iac"]
    avdw, bvdw, crg = ctx["avdw"], ctx["bvdw"], ctx["crg"]
    nsol, ns, el14 = ctx["nsol"], ctx["nstates"], ctx["el14"]
    nq = len(iqseq)
    qs = np.asarray(qcrg, float) * SQRT332
    nq1, nsol1 = max(nq, 1), max(nsol, 1)
    qp = {k: np.zeros((nsol1, nq1, ns)) for k in ("dq", "dA", "dB")}
    qp14 = np.zeros((nsol1, nq1, ns), np.uint8)
    if nq and nsol:
        lev = np.transpose(qconn[:, :nsol, :nq], (1, 2, 0))
        code = np.where(lev == 4, 2, 0)
        ok = (iqatom[:nsol] == 0)[:, None, None] & (lev >= 4)
        tj = iac[:nsol]
        Ajc = np.take_along_axis(np.broadcast_to(avdw[tj][:, None, None, :], (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        Bjc = np.take_along_axis(np.broadcast_to(bvdw[tj][:, None, None, :], (nsol, nq, ns, 3)), code[..., None], 3)[..., 0]
        dq = np.broadcast_to(crg[:nsol][:, None, None], (nsol, nq, ns)) * np.where(lev == 4, el14, 1.0)
        qp["dq"][:nsol, :nq] = np.where(ok, dq, 0.0)
        qp["dA"][:nsol, :nq] = np.where(ok, Ajc, 0.0)
        qp["dB"][:nsol, :nq] = np.where(ok, Bjc, 0.0)
        qp14[:nsol, :nq] = (ok & (lev == 4)).astype(np.uint8)
    qq = {k: np.zeros((nq1, nq1, ns)) for k in ("dq_i", "dA_i", "dB_i", "dq_j", "dA_j", "dB_j")}
    qq14 = np.zeros((nq1, nq1, ns), np.uint8)
    for iq in range(nq - 1):
        for jq in range(iq + 1, nq):
            for st in range(ns):
                lv = qconn[st, iqseq[jq], iq]
                if lv < 4:
                    continue
                k = 2 if lv == 4 else 0
                f = el14 if lv == 4 else 1.0
                qq["dq_i"][iq, jq, st], qq["dq_j"][iq, jq, st] = qs[jq, st] * f, qs[iq, st] * f
                qq["dA_i"][iq, jq, st], qq["dA_j"][iq, jq, st] = qA[jq, st, k], qA[iq, st, k]
                qq["dB_i"][iq, jq, st], qq["dB_j"][iq, jq, st] = qB[jq, st, k], qB[iq, st, k]
                qq14[iq, jq, st] = lv == 4
    sa = len(ctx["chg_solv"])
    qw = {k: np.zeros((nq1, sa, ns)) for k in ("dq", "dA", "dB")}
    if nq:
        qw["dq"][:nq] = ctx["chg_solv"][None, :, None]
        qw["dA"][:nq] = ctx["aLJ_solv"][None, :, 0, None]
        qw["dB"][:nq] = ctx["bLJ_solv"][None, :, 0, None]
    out = {"qp_" + k: v.ravel(order="F") for k, v in qp.items()}
    out["qp_is14"] = qp14.ravel(order="F")
    out.update({"qw_" + k: v.ravel(order="F") for k, v in qw.items()})
    out.update({"qq_" + k: v.ravel(order="F") for k, v in qq.items()})
    out["qq_is14"] = qq14.ravel(order="F")
    return out


def q_lj_factors(s, sig=None, eps=None):
    """Q-side geometric LJ factors qA/qB[nq, ns, code] from per-Q-type sigma
    and epsilon (A = 2 sqrt(eps) sigma^6, B = 2 sqrt(eps) sigma^3, every code)."""
    md = s.md
    sig = np.asarray(md["q_sig"] if sig is None else sig, float)
    eps = np.asarray(md["q_eps"] if eps is None else eps, float)
    t = md["q_type_of"]
    ns = s.n["nstates"]
    A = 2.0 * np.sqrt(eps[t]) * sig[t] ** 6
    B = 2.0 * np.sqrt(eps[t]) * sig[t] ** 3
    qA = np.repeat(np.repeat(A[:, None, None], ns, 1), 3, 2)
    qB = np.repeat(np.repeat(B[:, None, None], ns, 1), 3, 2)
    return qA, qB


def set_q_params(s, qcrg=None, qA=None, qB=None):
    """Rebuild the Q tables of a synthetic system from (perturbed) Q charges
    and Q-side LJ factors; returns the system (arrays replaced, not mutated)."""
    md = s.md
    if qcrg is not None:
        md["qcrg"] = np.asarray(qcrg, float)
    if qA is not None:
        md["qA"] = np.asarray(qA, float)
    if qB is not None:
        md["qB"] = np.asarray(qB, float)
    s.a.update(q_tables(md["qctx"], md["qcrg"], md["qA"], md["qB"]))
    return s.finalize()


def _zmat(rows):
    """Cartesian positions from a z-matrix: rows (ref_bond, r, ref_angle,
    theta_deg, ref_dihedral, phi_deg), None for the first atoms' missing
    references."""
    pos = []
    for k, (a, r, b, th, c, ph) in enumerate(rows):
        if k == 0:
            pos.append(np.zeros(3))
            continue
        if k == 1:
            pos.append(pos[a] + np.array([r, 0.0, 0.0]))
            continue
        pa, pb = pos[a], pos[b]
        pc = pos[c] if c is not None else pb + np.array([0.0, 1.0, 0.0])
        bc = pa - pb
        bc /= np.linalg.norm(bc)
        n = np.cross(pb - pc, bc)
        if np.linalg.norm(n) < 1e-8:
            n = np.cross(bc, np.array([0.0, 0.0, 1.0]))
        n /= np.linalg.norm(n)
        m = np.cross(n, bc)
        t, f = math.radians(th), math.radians(ph)
        d = np.array([-r * math.cos(t), r * math.sin(t) * math.cos(f), r * math.sin(t) * math.sin(f)])
        pos.append(pa + d[0] * bc + d[1] * m + d[2] * n)
    return np.array(pos)


def ethanol():
    """A nine-site flexible solvent molecule (OPLS-like ethanol): C2 first
    (the group's switch atom), then C1, O, HO, three H on C1, two on C2.
    Bonds, angles and the twelve torsions make three kinds of internal
    nonbonded pairs for nonbond_solvent_internal: 1-4 pairs through the
    torsions (LJ code 3, scaled Coulomb), and the three H(C1)-HO 1-5 pairs,
    which Qdyn6's list holds in both orders."""
    geo = _zmat([(None, 0, None, 0, None, 0),          # C2
                 (0, 1.529, None, 0, None, 0),          # C1
                 (0, 1.410, 1, 109.5, None, 0),         # O
                 (2, 0.945, 0, 108.5, 1, 180.0),        # HO
                 (1, 1.090, 0, 110.7, 2, 60.0),         # H11
                 (1, 1.090, 0, 110.7, 2, 180.0),        # H12
                 (1, 1.090, 0, 110.7, 2, 300.0),        # H13
                 (0, 1.090, 1, 110.7, 2, 120.0),        # H21
                 (0, 1.090, 1, 110.7, 2, 240.0)])       # H22
    # types: name, sigma, eps, mass (1-4 LJ at half the normal factors)
    types = [("CT2", 3.5, 0.066, 12.011), ("CT1", 3.5, 0.066, 12.011), ("OHE", 3.12, 0.17, 15.999),
             ("HOE", 0.0, 0.0, 1.008), ("HCE", 2.5, 0.03, 1.008)]
    atom_type = [0, 1, 2, 3, 4, 4, 4, 4, 4]
    charge = [0.145, -0.18, -0.683, 0.418, 0.06, 0.06, 0.06, 0.03, 0.03]
    bonds = [(0, 1, 0), (0, 2, 1), (2, 3, 2), (1, 4, 3), (1, 5, 3), (1, 6, 3), (0, 7, 3), (0, 8, 3)]
    bond_par = [(536.0, 1.529), (640.0, 1.410), (1106.0, 0.945), (680.0, 1.090)]
    angles = [(1, 0, 2, 0), (1, 0, 7, 1), (1, 0, 8, 1), (2, 0, 7, 1), (2, 0, 8, 1), (7, 0, 8, 2),
              (0, 1, 4, 1), (0, 1, 5, 1), (0, 1, 6, 1), (4, 1, 5, 2), (4, 1, 6, 2), (5, 1, 6, 2), (0, 2, 3, 3)]
    angle_par = [(100.0, 109.5), (75.0, 110.7), (66.0, 107.8), (110.0, 108.5)]
    tors = [(i, 0, 1, l, 0) for i in (2, 7, 8) for l in (4, 5, 6)] + [(i, 0, 2, 3, 1) for i in (1, 7, 8)]
    tor_par = [(0.3, 3.0, 0.0, 1.0), (0.45, 3.0, 0.0, 1.0)]
    return dict(name="EOH", geo=geo, types=types, atom_type=atom_type, charge=charge, bonds=bonds,
                bond_par=bond_par, angles=angles, angle_par=angle_par, torsions=tors, tor_par=tor_par,
                atom_names=["C2", "C1", "O", "HO", "H11", "H12", "H13", "H21", "H22"])


def solvent_internal_pairs(mol, avdw_solv, bvdw_solv, chg_solv, ljcode, el14, rule=1):
    """Nonbonded pairs inside one solvent molecule as Qdyn6 lists them
    (prep_sim_precompute_solvent, src/simprep.f90:4004-4138): every ordered
    pair (i, j), i != j, starts with its LJ code; bonded and angle-end pairs
    are dropped both ways; a torsion's end pair (i, l) becomes a 1-4 pair
    (code 3) in that order and (l, i) is dropped; the rest stay in both
    orders.  Returns (ij [n, 2] 0-based, par [n, 3] A, B, charge product)."""
    sa = len(mol["atom_type"])
    inter = np.array([[ljcode[i][j] for j in range(sa)] for i in range(sa)], int)
    np.fill_diagonal(inter, 0)
    for i, j, _ in mol["bonds"]:
        inter[i, j] = inter[j, i] = 0
    for i, _, k, _ in mol["angles"]:
        inter[i, k] = inter[k, i] = 0
    for i, _, _, l, _ in mol["torsions"]:
        if inter[i, l] != 0 or inter[l, i] != 0:
            inter[i, l] = 3
            inter[l, i] = 0
    ij, par = [], []
    for i in range(sa):
        for j in range(sa):
            c = inter[i, j]
            if c == 0:
                continue
            if rule == 1:
                A = avdw_solv[i, c - 1] * avdw_solv[j, c - 1]
                B = bvdw_solv[i, c - 1] * bvdw_solv[j, c - 1]
            else:
                tA = avdw_solv[i, c - 1] + avdw_solv[j, c - 1]
                tA = tA ** 2
                tA = tA * tA * tA
                tB = bvdw_solv[i, c - 1] * bvdw_solv[j, c - 1]
                A, B = (tA ** 2) * tB, 2.0 * tA * tB
            el = chg_solv[i] * chg_solv[j]
            if c == 3:
                el = el * el14
            ij.append((i, j))
            par.append((A, B, el))
    return np.array(ij, np.int32).reshape(-1, 2), np.array(par, float).reshape(-1, 3)


def replace_solvent(s, mol, density, seed, fill=None, dmin_solute=2.6, dmin_solvent=2.2):
    """The system s (from build) with its water replaced by copies of the
    solvent molecule `mol` (e.g. ethanol()): centres on a jittered cubic
    lattice at `density` molecules per A^3 within `fill` (default the sphere
    radius - 1), random orientations, dropped where a heavy atom comes within
    dmin_solute of a solute atom or dmin_solvent of a kept molecule's heavy
    atom.  Topology types, charges, groups, the pw/ww/qw tables, the
    solvent's internal nonbonded pairs and the MD data follow; the solvent is
    Qdyn6's SOLVENT_ALLATOM type, without the polarisation restraint."""
    rng = np.random.default_rng(seed)
    n, a, md = s.n, s.a, s.md
    nsol = int(n["nat_solute"])
    sa = len(mol["atom_type"])
    radius = float(s.par["rwat"])
    fill = radius - 1.0 if fill is None else fill
    spos = s.x[:nsol]
    geo = mol["geo"] - mol["geo"].mean(0)
    heavy_m = np.array([mol["types"][t][3] >= 4.0 for t in mol["atom_type"]])
    lat = density ** (-1.0 / 3.0)
    m = int(math.ceil(fill / lat)) + 1
    g = np.arange(-m, m + 1) * lat
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    cen = cen + rng.uniform(-0.3, 0.3, cen.shape)
    cen = cen[np.einsum("ij,ij->i", cen, cen) <= fill * fill]
    rot = _rand_rot(rng, len(cen))
    gs, gw = _Grid(3.0), _Grid(3.0)
    for k, p in enumerate(spos):
        gs.add(p, k)
    mols = []
    for c, R in zip(cen, rot):
        at = c + geo @ R.T
        hv = at[heavy_m]
        if any(gs.near(p, dmin_solute) for p in hv) or any(gw.near(p, dmin_solvent) for p in hv):
            continue
        for p in hv:
            gw.add(p, len(mols))
        mols.append(at)
    nw = len(mols)
    nat = nsol + sa * nw
    x = np.concatenate([spos] + mols) if nw else spos.copy()
    # types: the solute's own, then the solvent's (the water types stay, unused)
    nt0 = int(n["ntypes"])
    avdw = np.asarray(a["avdw"]).reshape(nt0, 3)
    bvdw = np.asarray(a["bvdw"]).reshape(nt0, 3)
    ntypes = nt0 + len(mol["types"])
    avdw = np.vstack([avdw, np.zeros((len(mol["types"]), 3))])
    bvdw = np.vstack([bvdw, np.zeros((len(mol["types"]), 3))])
    mass_t = np.concatenate([md["type_mass"], np.zeros(len(mol["types"]))])
    for k, (_, sig, eps, mass) in enumerate(mol["types"]):
        A, B = _lj(sig, eps) if eps > 0 else (0.0, 0.0)
        avdw[nt0 + k] = (A, A, 0.5 * A)
        bvdw[nt0 + k] = (B, B, 0.5 * B)
        mass_t[nt0 + k] = mass
    iac = np.concatenate([np.asarray(a["iac"])[:nsol], np.tile(nt0 + np.array(mol["atom_type"]), nw)]).astype(np.int32)
    crg_u = np.concatenate([md["crg_unscaled"][:nsol], np.tile(mol["charge"], nw)])
    crg = crg_u * SQRT332
    heavy = (mass_t[iac] >= 4.0).astype(np.uint8)
    ncs = int(n["ncgp_solute"])
    first = np.concatenate([np.asarray(a["cgp_first"])[:ncs], nsol + sa * np.arange(nw)]).astype(np.int32)
    last = np.concatenate([np.asarray(a["cgp_last"])[:ncs], nsol + sa * np.arange(nw) + sa - 1]).astype(np.int32)
    sw = np.concatenate([np.asarray(a["cgp_iswitch"])[:ncs], nsol + sa * np.arange(nw)]).astype(np.int32)
    aLJ = avdw[nt0 + np.array(mol["atom_type"])]
    bLJ = bvdw[nt0 + np.array(mol["atom_type"])]
    chg_solv = np.array(mol["charge"]) * SQRT332
    ww_A = np.array([[aLJ[la, 0] * aLJ[ka, 0] for ka in range(sa)] for la in range(sa)])
    ww_B = np.array([[bLJ[la, 0] * bLJ[ka, 0] for ka in range(sa)] for la in range(sa)])
    ww_el = np.array([[chg_solv[la] * chg_solv[ka] for ka in range(sa)] for la in range(sa)])
    nsol1 = max(nsol, 1)
    iqatom = np.concatenate([np.asarray(a["iqatom"])[:nsol], np.zeros(sa * nw, np.int32)]).astype(np.int32)
    pw_A, pw_B, pw_el = (np.zeros((nsol1, sa)) for _ in range(3))
    for i in range(nsol):
        if iqatom[i]:
            continue
        for j in range(sa):
            pw_A[i, j] = avdw[iac[i], 0] * aLJ[j, 0]
            pw_B[i, j] = bvdw[iac[i], 0] * bLJ[j, 0]
            pw_el[i, j] = crg[i] * chg_solv[j]
    ctx = dict(md["qctx"], aLJ_solv=aLJ, bLJ_solv=bLJ, chg_solv=chg_solv, iac=iac, crg=crg, iqatom=iqatom,
               avdw=avdw, bvdw=bvdw)
    ij, par = solvent_internal_pairs(mol, aLJ, bLJ, chg_solv, [[1] * sa for _ in range(sa)], s.el14_scale)
    a.update(crg=crg, iac=iac, iqatom=iqatom, excl=np.zeros(nat, np.uint8), heavy=heavy, cgp_first=first,
             cgp_last=last, cgp_iswitch=sw, cgpatom=np.arange(nat, dtype=np.int32), avdw=avdw.ravel(),
             bvdw=bvdw.ravel(), ljcod=np.ones(ntypes * ntypes, np.int32), pw_A=pw_A.ravel(order="F"),
             pw_B=pw_B.ravel(order="F"), pw_el=pw_el.ravel(order="F"), ww_A=ww_A.ravel(order="F"),
             ww_B=ww_B.ravel(order="F"), ww_el=ww_el.ravel(order="F"), solv_int_ij=ij.ravel(),
             solv_int_par=par.ravel())
    a.update(q_tables(ctx, md["qcrg"], md["qA"], md["qB"]))
    n.update(natom=nat, nwat=nw, solv_atom=sa, ncgp=ncs + nw, ntypes=ntypes, solvent_spc=0, n_solv_int=len(ij))
    s.x = x
    s.par.update(wpol_restr=0, nwpolr_shell=0, wshell_rout=np.zeros(0), wshell_dr=np.zeros(0),
                 wshell_cstb=np.zeros(0), wshell_theta_corr=np.zeros(0), solvent_density=density)
    md.update(mass=mass_t[iac], type_mass=mass_t, crg_unscaled=crg_u, qctx=ctx, solvent=mol,
              type_names=list(md["type_names"]) + [t[0] for t in mol["types"]])
    s.name = s.name + " in " + mol["name"]
    return s.finalize()


def cfg_ethanol(seed=41):
    """18 A sphere of a flexible nine-site solvent (ethanol()) around six
    10-atom bead chains, the innermost with 10 Q-atoms in 2 states: the
    general-solvent loops (nonbond_ww, nonbond_qw, nonbond_pw over nine
    sites) and the solvent's internal pairs."""
    s = build(18.0, 6, 10, 1, 2, seed, chain_region=6.0, q_sigma=0.4, lambdas=[0.6, 0.4], compact=3.0,
              water_fill=17.0, name="18 A sphere, 6x10-atom chains, 10 Q-atoms", **PACK)
    return replace_solvent(s, ethanol(), 0.0085, seed + 7)


def maxwell_velocities(s, T, seed):
    """Seeded Maxwell-Boltzmann velocities in Q units (A per internal time unit,
    dt_internal = 0.020462 * fs, src/md.f90:207), net momentum removed."""
    rng = np.random.default_rng(seed)
    m = s.md["mass"]
    sd = np.sqrt(BOLTZ * T / m)
    v = rng.normal(size=(len(m), 3)) * sd[:, None]
    p = (v * m[:, None]).sum(0) / m.sum()
    return v - p


# workload definitions ----------------------------------------------------

def cfg1(seed=1):
    """20 A sphere of TIP3P (~1,100 waters) around a 12-atom neutral chain, all
    12 atoms Q-atoms in 2 states (state 1 charges N(0,0.4), state 2 zero)."""
    return build(20.0, 1, 12, 1, 2, seed, chain_region=2.0, q_sigma=0.4, lambdas=[0.5, 0.5],
                 q_state0_zero=True, water_fill=20.0, name="cfg1: 20 A TIP3P sphere + 12 Q-atoms")


# Dense packing of the workload systems: 4,000 solute atoms at about water's
# atom density (0.1 per A^3) need non-bonded beads as close as 1.7 A while the
# chains grow (the start is relaxed by steepest descent before MD); water fills
# the whole sphere at 0.0334 per A^3 wherever its O stays 2.2 A and its H 1.6 A
# from every solute atom.
PACK = dict(solute_dmin=1.7, water_dmin=(2.2, 1.6))


def cfg2(seed=11, lambdas=(0.5, 0.5)):
    """30 A sphere of about 11,000 atoms: 400 neutral 10-atom bead-chain
    groups around the centre (they reach about 22.5 A), 30 Q-atoms in 2 EVB
    states, TIP3P filling the rest of the sphere (about 2,290 waters)."""
    return build(30.0, 400, 10, 3, 2, seed, chain_region=18.0, q_sigma=0.3, lambdas=list(lambdas), compact=3.0,
                 water_fill=30.0, name="cfg2: 30 A sphere, 400x10-atom solute groups, 30 Q-atoms, TIP3P", **PACK)


def describe(s):
    n = s.n
    return (f"{s.name}: natom={n['natom']} solute={n['nat_solute']} waters={n['nwat']} groups={n['ncgp']} "
            f"Q-atoms={n['nqat']} states={n['nstates']} chain_region={s.md.get('chain_region', 0):.1f} A")


def cfg5(seed=500, radii=None):
    """Uneven batch: 16 spheres of radius 15, 17, ..., 45 A built as cfg2
    (ten-atom bead chains packed as cfg2's, TIP3P filling the sphere), about
    0.1 atoms per A^3 of which 0.3 are solute (about 1,400 to 38,000 atoms),
    each with 10-60 Q-atoms (1-6 chains) in 2-3 states (seeded)."""
    return [cfg5_member(k, seed, radii) for k in range(len(radii) if radii is not None else 16)]


def cfg5_member(k, seed=500, radii=None):
    """Member k of the uneven batch (the same system cfg5 builds as its k-th)."""
    rng = np.random.default_rng(seed)
    radii = list(radii) if radii is not None else list(range(15, 47, 2))
    draws = [(int(rng.integers(1, 7)), int(rng.integers(2, 4))) for _ in radii]
    R = radii[k]
    nq, nst = draws[k]
    vol = 4.0 / 3.0 * math.pi * R ** 3
    nchain = max(2, int(round(0.3 * 0.1 * vol / 10)))
    lam = np.full(nst, 1.0 / nst)
    return build(float(R), nchain, 10, min(nq, nchain), nst, seed + 10 * k, chain_region=0.55 * R,
                 q_sigma=0.3, lambdas=lam, compact=3.0, water_fill=float(R), name=f"cfg5[{k}]: {R} A sphere", **PACK)
