"""Mathematical pins of the C restatement, independent of the golden vectors.

On a small synthetic sphere (bead-chain solute groups with 1-2/1-3
exclusions and 1-4 pairs, one chain of Q-atoms in 2 states, TIP3P):
  * every pairwise term's gradient equals central finite differences of its
    energy (Q terms: of the lambda-weighted state energies, as the reference
    applies lambda inside the force, src/nonbonded.f90:220);
  * Newton's third law for the pair terms;
  * translation invariance (sphere centres moved with the atoms);
  * the LRF Taylor expansion reproduces the direct Coulomb potential and
    field of the far charges to third order;
  * the SCAAS radial restraint and the polarisation restraint obey FD too.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from q6_amd import abi, synth


@pytest.fixture(scope="module")
def small():
    s = synth.build(11.0, 4, 6, 1, 2, seed=5, chain_region=4.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0,
                    name="small")
    o = Oracle(s)
    o.make_pair_lists(s.x)
    return s, o


def energy(o, mask, x, lam, Tfree=300.0):
    d, E, EQ = o.eval_terms(mask, x, Tfree=Tfree)
    e = 0.0
    for k in ("pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "solvent_radial", "water_pol"):
        e += E[k]
    for st, q in enumerate(EQ):
        e += lam[st] * sum(q.values())
    return e, d


@pytest.mark.parametrize("mask", [abi.T_WW, abi.T_PW, abi.T_PP, abi.T_QP, abi.T_QW, abi.T_QQ | abi.T_QQP,
                                  abi.T_RAD, abi.T_POL])
def test_gradient_matches_finite_differences(small, mask):
    s, o = small
    lam = s.par["lambda"]
    x0 = s.x.copy()
    _, d = energy(o, mask, x0, lam)
    rng = np.random.default_rng(3)
    atoms = list(rng.choice(s.natom, 12, replace=False)) + [0, 1, 2]
    h = 1e-5
    for a in atoms:
        for c in range(3):
            xp = x0.copy()
            xm = x0.copy()
            xp[a, c] += h
            xm[a, c] -= h
            fd = (energy(o, mask, xp, lam)[0] - energy(o, mask, xm, lam)[0]) / (2 * h)
            assert abs(fd - d[a, c]) <= 2e-5 * max(1.0, abs(d[a, c])), (mask, a, c, fd, d[a, c])


@pytest.mark.parametrize("mask", [abi.T_WW, abi.T_PW, abi.T_PP, abi.T_QP, abi.T_QW, abi.T_QQ | abi.T_QQP])
def test_newton_third_law(small, mask):
    s, o = small
    d, _, _ = o.eval_terms(mask, s.x)
    assert np.max(np.abs(d.sum(0))) <= 1e-9 * max(1.0, np.max(np.abs(d)))


def test_translation_invariance():
    s = synth.build(11.0, 4, 6, 1, 2, seed=5, chain_region=4.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0)
    o1 = Oracle(s)
    o1.make_pair_lists(s.x)
    d1, E1, EQ1 = o1.eval_terms(abi.T_ALL, s.x, Tfree=300.0)
    shift = np.array([3.25, -1.5, 7.0])
    s2 = synth.build(11.0, 4, 6, 1, 2, seed=5, chain_region=4.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0)
    s2.x = s.x + shift
    s2.par["xpcent"] = s.par["xpcent"] + shift
    s2.par["xwcent"] = s.par["xwcent"] + shift
    o2 = Oracle(s2)
    o2.make_pair_lists(s2.x)
    d2, E2, EQ2 = o2.eval_terms(abi.T_ALL, s2.x, Tfree=300.0)
    for k in E1:
        assert abs(E1[k] - E2[k]) <= 1e-9 * max(1.0, abs(E1[k])), k
    assert np.max(np.abs(d1 - d2)) <= 1e-8 * np.max(np.abs(d1))


def test_lrf_taylor_matches_direct_coulomb(small):
    """For each charge group a, the expansion about its centre reproduces
    sum_b q_b / |x - x_b| over the far (beyond cut-off) non-Q atoms b at
    points near the centre, with a third-order remainder."""
    s, o = small
    L = o.lrf()
    x = s.x
    n = s.n
    ncs = n["ncgp_solute"]
    ref = np.concatenate([s.a["cgp_iswitch"][:ncs], n["nat_solute"] + 3 * np.arange(n["nwat"])])
    sol = np.arange(n["ncgp"]) < ncs
    checked = 0
    for a in range(0, n["ncgp"], 7):
        c = L[a, :3]
        r2 = ((x[ref] - x[ref[a]]) ** 2).sum(1)
        cut = np.where(sol & sol[a], s.par["Rcpp"] ** 2, np.where(~sol & ~sol[a], s.par["Rcww"] ** 2,
                                                                  s.par["Rcpw"] ** 2))
        far = np.where((r2 > cut) & (np.arange(n["ncgp"]) != a))[0]
        src = [s.a["cgpatom"][k] for g in far for k in range(s.a["cgp_first"][g], s.a["cgp_last"][g] + 1)
               if s.a["iqatom"][s.a["cgpatom"][k]] == 0]
        if not src:
            continue
        q = s.a["crg"][src]
        P = x[src]
        p = c + np.array([0.31, -0.22, 0.17])
        dr = c - p
        phi0, phi1 = L[a, 3], L[a, 4:7]
        phi2 = L[a, 7:16].reshape(3, 3)
        V = phi0 + dr @ phi1 + 0.5 * dr @ (phi2 @ dr)
        Vd = (q / np.linalg.norm(P - p, axis=1)).sum()
        assert abs(V - Vd) <= 2e-3 * np.abs(q / np.linalg.norm(P - p, axis=1)).sum()
        checked += 1
    assert checked > 5
