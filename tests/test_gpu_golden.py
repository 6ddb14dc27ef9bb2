"""HIP engine against the reference's golden vectors (GPU).

Tolerances (FP64 plain path): pair lists bit-exact as sets (and in the
reference's order); energies and gradients within 1e-10 relative, the bar of
the FP64 form.  The gradient tolerance is relative to the largest component of
the term's gradient; LRF table entries relative to the largest entry of the
same kind (phi0, phi1, phi2, phi3).
"""
import numpy as np
import pytest

from conftest import golden_path
from oracle.oracle import Oracle
from q6_amd import abi
from q6_amd.system import from_golden

pytestmark = pytest.mark.gpu

REL = 1e-10
TERMS = [("ww", abi.T_WW), ("qw", abi.T_QW), ("pw", abi.T_PW), ("pp", abi.T_PP), ("qp", abi.T_QP),
         ("lrf", abi.T_LRF), ("qq", abi.T_QQ), ("qqp", abi.T_QQP), ("solvent_radial", abi.T_RAD),
         ("water_pol", abi.T_POL), ("solv_int", abi.T_SINT)]


def engine_for(s, **kw):
    from q6_amd.engine import NonbondedEngine
    return NonbondedEngine(s, **kw)


def term_energy(t, E, EQ):
    if t in ("ww", "pw", "pp"):
        return [E[t + "_el"], E[t + "_vdw"]]
    if t == "solv_int":  # nonbond_solvent_internal adds into the solvent-solvent energies
        return [E["ww_el"], E["ww_vdw"]]
    if t in ("qw", "qp", "qq"):
        return sum([[q[t + "_el"], q[t + "_vdw"]] for q in EQ], [])
    if t == "qqp":
        return sum([[q["qp_el"], q["qp_vdw"]] for q in EQ], [])
    return [E[t]]


@pytest.fixture(scope="module", params=["lig_eq1", "lig_fep2", "lig_fep2_arith", "lig_fep2_anyatom", "cjp_excl",
                                        "eoh", "eoh_pol"])
def gcase(request):
    print(f"\n[gpu] golden case {request.param}", flush=True)
    s = from_golden(golden_path(request.param))
    eng = engine_for(s)
    eng.make_pair_lists(s.x)
    yield request.param, s, eng
    eng.close()


@pytest.mark.parametrize("which,key", [(abi.LIST_PP, "nbpp_ij"), (abi.LIST_PW, "nbpw_ij"),
                                       (abi.LIST_WW, "ww_mol_pairs"), (abi.LIST_QP, "nbqp_ij")])
def test_lists_match_reference(gcase, which, key):
    name, s, eng = gcase
    ref = s.ref[key].T
    if key.startswith("nb"):
        ref = ref - 1
    got = eng.pairs(which)
    assert got.shape == ref.shape
    # sorted-set equality is the contract; the export also keeps the reference's order
    assert np.array_equal(np.unique(got, axis=0), np.unique(ref, axis=0))
    assert np.array_equal(got, ref)


def test_qw_list_matches_reference(gcase):
    name, s, eng = gcase
    ref = s.ref["nbqw_ij"].T - 1  # (iq, water atom) in blocks of solv_atom
    sa = int(s.n["solv_atom"])
    ref_mol = np.stack([ref[::sa, 0], (ref[::sa, 1] - int(s.n["nat_solute"])) // sa], axis=1)
    got = eng.pairs(abi.LIST_QW)
    assert np.array_equal(got, ref_mol)


def test_lrf_table(gcase):
    name, s, eng = gcase
    got, ref = eng.lrf(), s.ref["lrf"].T
    for lo, hi in ((0, 3), (3, 4), (4, 7), (7, 16), (16, 43)):
        scale = np.max(np.abs(ref[:, lo:hi]))
        assert np.max(np.abs(got[:, lo:hi] - ref[:, lo:hi])) <= REL * scale, (lo, hi)


@pytest.mark.parametrize("term,mask", TERMS)
def test_terms_match_reference(gcase, term, mask):
    name, s, eng = gcase
    if "d_" + term not in s.ref:
        pytest.skip("term not in this system")
    d, E, EQ = eng.eval_terms(mask, s.x, Tfree=s.Tfree, istep=0)
    dref = s.ref["d_" + term].T
    eref = np.atleast_1d(s.ref["E_" + term])
    e = np.array(term_energy(term, E, EQ))
    escale = max(np.max(np.abs(eref)), 1e-12)
    assert np.max(np.abs(e - eref)) <= REL * escale, (e, eref)
    dscale = max(np.max(np.abs(dref)), 1e-12)
    assert np.max(np.abs(d - dref)) <= REL * dscale


def test_full_nonbonded_call(gcase):
    """qnb_nonbonded = the whole of pot_energy_nonbonds in one call."""
    name, s, eng = gcase
    d, E, EQ = eng.pot_energy_nonbonds(s.x)
    terms = [t for t in ("ww", "qw", "pw", "pp", "qp", "lrf", "solv_int") if "d_" + t in s.ref]
    dref = sum(s.ref["d_" + t].T for t in terms)
    assert np.max(np.abs(d - dref)) <= REL * np.max(np.abs(dref))
    eref = {t: np.array(s.ref["E_" + t], float) for t in ("ww", "pw", "pp")}
    if "E_solv_int" in s.ref:
        eref["ww"] = eref["ww"] + s.ref["E_solv_int"]
    for t in ("ww", "pw", "pp"):
        assert abs(E[t + "_el"] - eref[t][0]) <= REL * max(abs(eref[t][0]), 1.0)
        assert abs(E[t + "_vdw"] - eref[t][1]) <= REL * max(abs(eref[t][1]), 1.0)


def test_general_solvent_mixed_precision():
    """The nine-site flexible solvent on the mixed path: the force error of
    every term together against stock Qdyn6's FP64 within the 3e-6 bound of
    the water systems."""
    print("\n[gpu] general solvent, mixed precision", flush=True)
    s = from_golden(golden_path("eoh"))
    eng = engine_for(s, precision=abi.QNB_PREC_MIXED)
    eng.make_pair_lists(s.x)
    d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=s.Tfree)
    eng.close()
    dref = sum(s.ref["d_" + t].T for t, _ in TERMS if "d_" + t in s.ref)
    err = np.max(np.abs(d - dref)) / np.max(np.abs(dref))
    print(f"[gpu] general solvent mixed max force error {err:.2e}", flush=True)
    assert err < 3e-6


@pytest.mark.parametrize("quirks", [1, 0])
def test_nolrf_lists(quirks):
    """LRF off, switch atoms (nbpplist, nbpwlist, nbwwlist, nbqplist).  With
    the reference's own mode (ref_quirks bit 0: nbpwlist's distance test
    moves to atom cgp(ig)%last + 1 after a group's first water inside,
    nonbondene.f90:2686-2691) every list equals stock Qdyn6's, in order, and
    every term is within 1e-10 of its dump.  Without it the solute-water list
    follows the switch-atom test for every water; that mode is checked
    against the oracle in the same mode."""
    print(f"\n[gpu] lig_fep2_nolrf ref_quirks={quirks}", flush=True)
    s = from_golden(golden_path("lig_fep2_nolrf"))
    eng = engine_for(s, ref_quirks=quirks)
    eng.make_pair_lists(s.x)
    o = Oracle(s, ref_quirks=quirks)
    o.make_pair_lists(s.x)
    for which in (abi.LIST_PP, abi.LIST_PW, abi.LIST_WW, abi.LIST_QP, abi.LIST_QW):
        assert np.array_equal(eng.pairs(which), o.pairs(which)), which
    if quirks:
        for which, key in ((abi.LIST_PP, "nbpp_ij"), (abi.LIST_PW, "nbpw_ij"), (abi.LIST_QP, "nbqp_ij")):
            assert np.array_equal(eng.pairs(which), s.ref[key].T - 1), key
        assert np.array_equal(eng.pairs(abi.LIST_WW), s.ref["ww_mol_pairs"].T)
        assert len(s.ref["nbpw_ij"].T) == 51072
        for term, mask in [tm for tm in TERMS if "d_" + tm[0] in s.ref]:
            if term == "lrf":
                continue
            d, E, EQ = eng.eval_terms(mask, s.x, Tfree=s.Tfree, istep=0)
            dref, eref = s.ref["d_" + term].T, np.atleast_1d(s.ref["E_" + term])
            e = np.array(term_energy(term, E, EQ))
            assert np.max(np.abs(e - eref)) <= REL * max(np.max(np.abs(eref)), 1e-12), (term, e, eref)
            assert np.max(np.abs(d - dref)) <= REL * max(np.max(np.abs(dref)), 1e-12), term
    else:
        assert len(eng.pairs(abi.LIST_PW)) == 60201
        d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=s.Tfree)
        do, Eo, EQo = o.eval_terms(abi.T_ALL, s.x, Tfree=s.Tfree)
        assert np.max(np.abs(d - do)) <= REL * np.max(np.abs(do))
    eng.close()


def test_anyatom_nolrf_matches_reference():
    """iuse_switch_atom = 0 with LRF off (nbpplis2, nbpwlis2, nbqplis2): these
    builders have no switch-atom quirk, so every list equals the reference's,
    in order, and every term is within 1e-10."""
    print("\n[gpu] lig_fep2_anyatom_nolrf", flush=True)
    s = from_golden(golden_path("lig_fep2_anyatom_nolrf"))
    assert s.n["any_atom_cutoff"] == 1
    eng = engine_for(s)
    eng.make_pair_lists(s.x)
    for which, key in ((abi.LIST_PP, "nbpp_ij"), (abi.LIST_PW, "nbpw_ij"), (abi.LIST_QP, "nbqp_ij")):
        assert np.array_equal(eng.pairs(which), s.ref[key].T - 1), key
    assert np.array_equal(eng.pairs(abi.LIST_WW), s.ref["ww_mol_pairs"].T)
    for term, mask in [tm for tm in TERMS if "d_" + tm[0] in s.ref]:
        if term == "lrf":
            continue
        d, E, EQ = eng.eval_terms(mask, s.x, Tfree=s.Tfree, istep=0)
        dref, eref = s.ref["d_" + term].T, np.atleast_1d(s.ref["E_" + term])
        e = np.array(term_energy(term, E, EQ))
        assert np.max(np.abs(e - eref)) <= REL * max(np.max(np.abs(eref)), 1e-12), (term, e, eref)
        assert np.max(np.abs(d - dref)) <= REL * max(np.max(np.abs(dref)), 1e-12), term
    eng.close()


def pair_energy_scales(s):
    """sqrt(sum of squared pair energies) of each class's electrostatic and
    van der Waals terms, from the fixture's reference lists (1-based pairs
    with their A, B, qq; water-water as molecule pairs over the 9 atom pairs)."""
    x, out = s.x, {}
    for t in ("pp", "pw"):
        ij, (A, B, el) = s.ref[f"nb{t}_ij"] - 1, s.ref[f"nb{t}_par"]
        r = np.linalg.norm(x[ij[1]] - x[ij[0]], axis=1)
        out[t] = (np.sqrt(((el / r) ** 2).sum()), np.sqrt(((A / r ** 12) ** 2 + (B / r ** 6) ** 2).sum()))
    nsol, mp = s.n["nat_solute"], s.ref["ww_mol_pairs"]
    se = sv = 0.0
    for la in range(3):
        for ka in range(3):
            r = np.linalg.norm(x[nsol + 3 * mp[1] + ka] - x[nsol + 3 * mp[0] + la], axis=1)
            se += ((s.ref["ww_el"][la, ka] / r) ** 2).sum()
            sv += ((s.ref["ww_A"][la, ka] / r ** 12) ** 2 + (s.ref["ww_B"][la, ka] / r ** 6) ** 2).sum()
    out["ww"] = (np.sqrt(se), np.sqrt(sv))
    return out


@pytest.mark.parametrize("case", ["lig_fep2", "lig_fep2_arith"])
def test_mixed_precision_error_budget(case):
    """FP32 pair arithmetic with FP64 accumulation: the error against the FP64
    reference is reported and bounded.  Forces: 3e-6 of the largest gradient
    component (about ten times the 2.7e-7 observed on lig_fep2).  Energy of a class: each pair energy carries an FP32 relative
    error of a few ulp (2^-23), independent from pair to pair, so the class sum
    errs by a few ulp times S = sqrt(sum of squared pair energies); bound 16
    ulp x S.  (A relative bound on the sum itself is meaningless for classes
    like water-water Coulomb: -10.6 kcal/mol from 3.4e5 terms of about +-50.)"""
    print(f"\n[gpu] mixed precision {case}", flush=True)
    s = from_golden(golden_path(case))
    eng = engine_for(s, precision=abi.QNB_PREC_MIXED)
    eng.make_pair_lists(s.x)
    d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=s.Tfree)
    dref = sum(s.ref["d_" + t].T for t, _ in TERMS if "d_" + t in s.ref)
    err = np.max(np.abs(d - dref)) / np.max(np.abs(dref))
    print(f"\n[gpu] mixed precision max force error {err:.2e}", flush=True)
    assert err < 3e-6
    ulp = 2.0 ** -23
    scales = pair_energy_scales(s)
    for t in ("ww", "pw", "pp"):
        for k, suf in enumerate(("_el", "_vdw")):
            ref = s.ref["E_" + t][k]
            S = scales[t][k]
            e = abs(E[t + suf] - ref)
            print(f"[gpu] mixed {t}{suf}: {E[t + suf]:.6f} vs {ref:.6f}  |err| = {e:.2e} = {e / (ulp * S):.2f} ulp x S",
                  flush=True)
            assert e <= 16 * ulp * S, (t, suf, E[t + suf], ref, S)
    eng.close()


@pytest.mark.parametrize("case", ["lig_eq1", "lig_fep2", "eoh_pol"])
def test_watpol_sequence(case):
    """The polarisation restraint over istep = 1..7 with itdis = 3 from a zero
    shell state (the theta_corr feedback of src/nonbondene.f90:6653-6666 runs
    at istep 3 and 6): energy, gradient and shell state of every call within
    1e-10 of stock Qdyn6's."""
    from conftest import watpol_frames
    print(f"\n[gpu] watpol sequence {case}", flush=True)
    s = from_golden(golden_path(case))
    eng = engine_for(s, itdis=float(s.ref["wp_itdis"]))
    eng.make_pair_lists(s.x)
    nsh = int(s.par["nwpolr_shell"])
    eng.set_shell_state(np.zeros(3 * nsh))
    for k in range(1, 8):
        x = watpol_frames(s, k)
        d, E, _ = eng.eval_terms(abi.T_POL, x, Tfree=s.Tfree, istep=k)
        eref, dref = float(s.ref[f"E_wp{k}"]), s.ref[f"d_wp{k}"].T
        assert abs(E["water_pol"] - eref) <= REL * abs(eref), (k, E["water_pol"], eref)
        assert np.max(np.abs(d - dref)) <= REL * np.max(np.abs(dref)), k
        st, sref = eng.shell_state(), s.ref[f"wp{k}_shell"].reshape(-1)
        assert np.max(np.abs(st - sref)) <= REL * np.max(np.abs(sref)), (k, st, sref)
    eng.close()


@pytest.mark.parametrize("case", ["lig_fep2", "lig_fep2_arith"])
def test_pot_energy_summary(case):
    """The engine's class energies of one full evaluation, summed by
    pot_energy_summary with the caller's bonded terms, give Qdyn6's E%qx,
    EQ(s)%total and E%potential within 1e-10 (src/potene.f90:224-251)."""
    from conftest import golden_summary_inputs
    from q6_amd.engine import pot_energy_summary
    print(f"\n[gpu] pot_energy summary {case}", flush=True)
    s = from_golden(golden_path(case))
    eng = engine_for(s)
    eng.make_pair_lists(s.x)
    d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=s.Tfree, istep=0)
    eng.close()
    bonded, qb, qr, lam = golden_summary_inputs(s)
    out, eqo = pot_energy_summary(E, EQ, lam, bonded, qb, qr)
    es, eqs = s.ref["E_summary"], s.ref["EQ_summary"]
    scale = max(abs(es[0]), np.abs(es[1:]).max())
    assert abs(out["qx_el"] - es[8]) <= REL * scale and abs(out["qx_vdw"] - es[9]) <= REL * scale
    for st, q in enumerate(eqs.T):
        assert abs(eqo[st]["total"] - q[1]) <= REL * scale, st
    assert abs(out["potential"] - es[0]) <= REL * scale, (out["potential"], es[0])


def test_calculate_exc_matches_reference():
    """qnb_calculate_exc on the reference's group-contribution system (2CJP,
    twelve energy arrays with residue groups, the last Q-atom and atoms bonded
    to Q-atoms) against stock Qdyn6's calculate_exc: every value within 1e-12
    relative (FP64 in the reference's pair order; the soft pairs' exp is the
    device's)."""
    from conftest import exc_inputs
    print("\n[gpu] calculate_exc cjp_excl", flush=True)
    s = from_golden(golden_path("cjp_excl"))
    masks, types, eq_in, ref = exc_inputs(s)
    eng = engine_for(s)
    eng.set_groups(masks, types)
    got = eng.calculate_exc(s.x, eq_in)
    eng.set_groups(np.zeros((0, s.n["nat_solute"])), [])
    assert eng.calculate_exc(s.x, eq_in).shape[0] == 0
    eng.close()
    assert got.shape == ref.shape
    assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1.0)), np.max(np.abs(got - ref))
