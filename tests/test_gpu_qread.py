"""Q input files to the HIP engine with no reference code in between (GPU):
the reference's SPH test system and a general solvent are read by
q6_amd/qread.py, started as Qdyn6 starts them (q6_amd.qread.seeded_start)
and evaluated on the device; every nonbonded term equals stock Qdyn6's
golden energies and gradients.  Tolerance 1e-9 relative: the host's seeded
start reproduces Qdyn6's positions to 1e-12 relative (the golden values are
at Qdyn6's own positions), the engine's FP64 bar is 1e-10."""
import numpy as np
import pytest

from conftest import golden_path
from q6_amd import abi, qread
from q6_amd.system import from_golden
from qfiles_cases import write_lig_case

pytestmark = pytest.mark.gpu

REL = 1e-9
TERMS = [("ww", abi.T_WW), ("qw", abi.T_QW), ("pw", abi.T_PW), ("pp", abi.T_PP), ("qp", abi.T_QP),
         ("lrf", abi.T_LRF), ("qq", abi.T_QQ), ("qqp", abi.T_QQP), ("solvent_radial", abi.T_RAD),
         ("water_pol", abi.T_POL), ("solv_int", abi.T_SINT)]


def term_energy(t, E, EQ):
    if t in ("ww", "pw", "pp"):
        return [E[t + "_el"], E[t + "_vdw"]]
    if t == "solv_int":
        return [E["ww_el"], E["ww_vdw"]]
    if t in ("qw", "qp", "qq"):
        return sum([[q[t + "_el"], q[t + "_vdw"]] for q in EQ], [])
    if t == "qqp":
        return sum([[q["qp_el"], q["qp_vdw"]] for q in EQ], [])
    return [E[t]]


@pytest.mark.parametrize("case", ["lig_eq1", "lig_fep2", "lig_fep2_arith", "eoh"])
def test_files_to_device(case, tmp_path):
    print(f"\n[gpu] files -> engine: {case}", flush=True)
    from q6_amd.engine import NonbondedEngine
    if case == "eoh":
        from q6_amd import qfiles, synth
        s = qread.build_system(qfiles.write_case(synth.cfg_ethanol(), str(tmp_path), 10, name="eoh", seed=0))
        x = s.x
    else:
        s = qread.build_system(write_lig_case(case, str(tmp_path)))
        x, _ = qread.seeded_start(s)
    g = from_golden(golden_path(case))
    eng = NonbondedEngine(s)
    try:
        eng.make_pair_lists(x)
        for term, mask in TERMS:
            if "d_" + term not in g.ref:
                continue
            d, E, EQ = eng.eval_terms(mask, x, Tfree=g.Tfree, istep=0)
            eref = np.atleast_1d(g.ref["E_" + term])
            e = np.array(term_energy(term, E, EQ))
            assert np.max(np.abs(e - eref)) <= REL * max(np.max(np.abs(eref)), 1e-12), (term, e, eref)
            dref = g.ref["d_" + term].T
            assert np.max(np.abs(d - dref)) <= REL * max(np.max(np.abs(dref)), 1e-12), term
        if case == "lig_eq1":
            # the reference's own known answer (tests/basic_tests/eval_test.sh:97-101)
            d, E, EQ = eng.eval_terms(abi.T_QW, x, Tfree=g.Tfree, istep=0)
            assert "%.2f %.2f" % (EQ[0]["qw_el"], EQ[0]["qw_vdw"]) == "3.12 139.43"
    finally:
        eng.close()
