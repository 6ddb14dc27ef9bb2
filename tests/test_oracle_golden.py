"""The C restatement (oracle/) against golden vectors of the reference itself.

The fixtures were written by stock Qdyn6 modules built from the reference
sources (tests/golden/make_golden.py).  Every list, the LRF table, and the
energy and gradient of every term must agree BIT FOR BIT: the restatement
follows the reference's loop and summation order exactly.
"""
import numpy as np
import pytest

from conftest import CASES, golden_path
from oracle.oracle import Oracle
from q6_amd import abi
from q6_amd.system import from_golden

TERMS = [("ww", abi.T_WW), ("qw", abi.T_QW), ("pw", abi.T_PW), ("pp", abi.T_PP), ("qp", abi.T_QP),
         ("lrf", abi.T_LRF), ("qq", abi.T_QQ), ("qqp", abi.T_QQP), ("solvent_radial", abi.T_RAD),
         ("water_pol", abi.T_POL), ("solv_int", abi.T_SINT)]


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


@pytest.fixture(scope="module", params=CASES)
def case(request):
    s = from_golden(golden_path(request.param))
    o = Oracle(s, ref_quirks=1)
    o.make_pair_lists(s.x)
    return request.param, s, o


def test_known_answer_qsurr():
    # tests/basic_tests/eval_test.sh:97-101 of the reference: "Q-surr. 1 1.0000 3.12 139.43"
    s = from_golden(golden_path("lig_eq1"))
    o = Oracle(s)
    o.make_pair_lists(s.x)
    _, _, EQ = o.eval_terms(abi.T_QW | abi.T_QP, s.x)
    assert "%.2f %.2f" % (EQ[0]["qw_el"] + EQ[0]["qp_el"], EQ[0]["qw_vdw"] + EQ[0]["qp_vdw"]) == "3.12 139.43"


@pytest.mark.parametrize("which,key", [(abi.LIST_PP, "nbpp_ij"), (abi.LIST_PW, "nbpw_ij"),
                                       (abi.LIST_WW, "ww_mol_pairs"), (abi.LIST_QP, "nbqp_ij")])
def test_lists_exact_order(case, which, key):
    name, s, o = case
    ref = s.ref[key].T
    if key.startswith("nb"):
        ref = ref - 1
    got = o.pairs(which)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)  # same pairs, same orientation, same append order


def test_list_parameters(case):
    name, s, o = case
    if "nbpp_par" not in s.ref:
        pytest.skip("fixture stored without list parameters")
    par = o.list_params(abi.LIST_PP, 3)
    assert np.array_equal(par.T, s.ref["nbpp_par"].reshape(3, -1))
    par = o.list_params(abi.LIST_PW, 3)
    assert np.array_equal(par.T, s.ref["nbpw_par"].reshape(3, -1))


def test_lrf_table_exact(case):
    name, s, o = case
    if not s.par["use_LRF"]:
        pytest.skip("LRF off")
    assert np.array_equal(o.lrf(), s.ref["lrf"].T)


@pytest.mark.parametrize("term,mask", TERMS)
def test_terms_exact(case, term, mask):
    name, s, o = case
    if "d_" + term not in s.ref:
        pytest.skip("term not dumped for this case")
    d, E, EQ = o.eval_terms(mask, s.x, Tfree=s.Tfree, istep=0)
    assert np.array_equal(d, s.ref["d_" + term].T)
    assert np.array_equal(np.array(term_energy(term, E, EQ)), np.atleast_1d(s.ref["E_" + term]))


def test_watpol_sequence_exact(case):
    """watpol over istep = 1..7 with itdis = 3 from a zero shell state, the
    waters moving between calls: energies, gradients and the shell state
    (theta_corr feedback at istep 3 and 6, running averages) equal stock
    Qdyn6's bit for bit (src/nonbondene.f90:6548-6746)."""
    from conftest import watpol_frames
    name, s, _ = case
    if "E_wp1" not in s.ref:
        pytest.skip("no polarisation restraint in this case")
    o = Oracle(s, ref_quirks=1, itdis=float(s.ref["wp_itdis"]))
    nsh = int(s.par["nwpolr_shell"])
    o.set_shell_state(np.zeros(3 * nsh))
    corr_moved = False
    for k in range(1, 8):
        x = watpol_frames(s, k)
        d, E, _ = o.eval_terms(abi.T_POL, x, Tfree=s.Tfree, istep=k)
        assert E["water_pol"] == float(s.ref[f"E_wp{k}"]), k
        assert np.array_equal(d, s.ref[f"d_wp{k}"].T), k
        st = o.shell_state()
        assert np.array_equal(st, s.ref[f"wp{k}_shell"].reshape(-1)), k
        corr_moved = corr_moved or np.any(st[:nsh] != 0.0)
    assert corr_moved  # the feedback branch ran


def test_pot_energy_summary_exact(case):
    """pot_energy's summary (src/potene.f90:224-251) as the engine's host code
    forms it, from the dumped class energies and the caller's bonded terms,
    equals Qdyn6's E%qx, EQ(s)%qx, EQ(s)%total and E%potential bit for bit."""
    from conftest import golden_summary_inputs
    from q6_amd.engine import pot_energy_summary
    name, s, _ = case
    es, eqs = s.ref["E_summary"], s.ref["EQ_summary"]
    bonded, qb, qr, lam = golden_summary_inputs(s)
    E = dict(pp_el=es[2], pp_vdw=es[3], pw_el=es[4], pw_vdw=es[5], ww_el=es[6], ww_vdw=es[7], lrf=es[1],
             solvent_radial=es[10], water_pol=es[11])
    EQ = [dict(qq_el=q[8], qq_vdw=q[9], qp_el=q[10], qp_vdw=q[11], qw_el=q[12], qw_vdw=q[13]) for q in eqs.T]
    out, eqo = pot_energy_summary(E, EQ, lam, bonded, qb, qr)
    assert out["qx_el"] == es[8] and out["qx_vdw"] == es[9]
    for st, q in enumerate(eqs.T):
        assert eqo[st]["qx_el"] == q[6] and eqo[st]["qx_vdw"] == q[7] and eqo[st]["total"] == q[1], st
    assert out["restraint_total"] == s.ref["E_bonded"][13]
    assert out["potential"] == es[0]


def test_calculate_exc_exact(case):
    """calculate_exc (src/nonbondene.f90:6825-6920) on the reference's
    group-contribution system: every energy array (the reference test's
    residue groups, and groups with the last Q-atom and atoms bonded to
    Q-atoms) and state equals stock Qdyn6's bit for bit."""
    from conftest import exc_inputs
    name, s, o = case
    if "exc_EQ" not in s.ref:
        pytest.skip("no group contributions in this case")
    masks, types, eq_in, ref = exc_inputs(s)
    iq, q3 = s.a["iqatom"][:s.n["nat_solute"]], s.a["qconn3"]
    m = masks.astype(bool)
    assert (m & (iq != 0)).any() and (m & (iq == 0) & (q3 != 0)).any() and (m & (iq == 0) & (q3 == 0)).any()
    assert np.array_equal(o.calculate_exc(s.x, masks, types, eq_in), ref)
