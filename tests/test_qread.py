"""The native reader of Q's input files (q6_amd/qread.py) against the tables
stock Qdyn6 set up from the same files (the golden dumps): every array the
engine takes, its sizes and run parameters, bit for bit, on the reference's
own SPH test system (five input variants: one state, two states with Q-atom
types and soft-core, LRF off, the arithmetic rule, any-atom groups), the
reference's 2CJP system (soft pairs, 40 Q-atoms, excluded atoms; read from
the reference's test folder where present) and a nine-site general solvent
written by q6_amd.qfiles; then Qdyn6's seeded start from the files."""
import os

import numpy as np
import pytest

from conftest import golden_path
from q6_amd import qread
from q6_amd.system import QSystem, from_golden
from qfiles_cases import LIG_CASES, write_cjp_case, write_lig_case


def compare(s, g):
    bad = []
    for k in QSystem.ARRAYS:
        a, b = np.asarray(s.a[k]).ravel(), np.asarray(g.a[k]).ravel()
        if a.shape != b.shape or not np.array_equal(a.astype(np.float64), b.astype(np.float64)):
            bad.append(k)
    for k in g.n:
        if int(s.n.get(k, -1)) != int(g.n[k]):
            bad.append("n." + k)
    for k in g.par:
        a = np.atleast_1d(s.par.get(k, np.nan)).astype(np.float64)
        b = np.atleast_1d(g.par[k]).astype(np.float64)
        if a.shape != b.shape or not np.array_equal(a, b):
            bad.append("par." + k)
    if s.el14_scale != g.el14_scale:
        bad.append("el14_scale")
    if not np.array_equal(s.winv, g.winv):
        bad.append("winv")
    return bad


@pytest.mark.parametrize("case", LIG_CASES)
def test_lig_cases_match_qdyn6(case, tmp_path):
    s = qread.build_system(write_lig_case(case, str(tmp_path)))
    g = from_golden(golden_path(case))
    assert compare(s, g) == []
    assert np.array_equal(s.x, g.ref["x_preconstraint"].T)  # topology coordinates


def test_cjp_matches_qdyn6(tmp_path):
    inp = write_cjp_case(str(tmp_path))
    if inp is None:
        pytest.skip("the reference's 2CJP test folder is not present")
    s = qread.build_system(inp)
    g = from_golden(golden_path("cjp_excl"))
    assert compare(s, g) == []
    assert np.array_equal(s.x, g.ref["x_preconstraint"].T)


@pytest.mark.parametrize("case", ["eoh", "eoh_pol"])
def test_general_solvent_matches_qdyn6(case, tmp_path):
    from q6_amd import qfiles, synth
    sy = synth.cfg_ethanol()
    sy.par["wpol_restr"] = int(case == "eoh_pol")
    inp = qfiles.write_case(sy, str(tmp_path), 10, name="eoh", seed=0)
    s = qread.build_system(inp)
    g = from_golden(golden_path(case))
    assert compare(s, g) == []
    assert s.n["n_solv_int"] == 18


@pytest.mark.parametrize("case", ["lig_eq1", "lig_fep2"])
def test_seeded_start_from_files(case, tmp_path):
    s = qread.build_system(write_lig_case(case, str(tmp_path)))
    g = from_golden(golden_path(case))
    x, v = qread.seeded_start(s)
    assert np.max(np.abs(x - g.x)) <= 1e-12 * np.max(np.abs(g.x))
    assert np.max(np.abs(v - g.ref["v"].T)) <= 1e-9 * np.max(np.abs(g.ref["v"]))


def test_list_directed_records():
    r = qread._Records("3 = count\n1 2\n3\n4 5 6 7\n8\n")
    n = qread._lead(r.line(), (int,))[0]
    assert r.values(n, int) == [1, 2, 3]
    assert r.values(2, int) == [4, 5]  # the rest of the record is not read
    assert r.values(1, int) == [8]
    assert qread._lead("5 16 1 = No. of charge groups", (int, int, int)) == [5, 16, 1]
    assert qread._lead("12 = only one", (int, int)) == [12]
    assert qread._num("1.5D-3") == 1.5e-3


def test_flag_records():
    r = qread._Records("1" * 80 + "\n" + "0101\n")
    f = r.flags(84, "1")
    assert f.sum() == 82 and f[81] == 1 and f[80] == 0
    r = qread._Records("TFT\n")
    assert list(qread._Records("TF T\n").flags(4, "Tt")) == [1, 0, 0, 1]
    assert list(r.flags(3, "Tt")) == [1, 0, 1]


def test_refusals(tmp_path):
    d = str(tmp_path)
    inp = write_lig_case("lig_eq1", d)
    with open(os.path.join(d, "off.fep"), "w") as f:
        f.write("[FEP]\nstates 1\noffset 10\n[atoms]\n1 1\n")
    with pytest.raises(qread.QFileError):
        qread.read_fep(os.path.join(d, "off.fep"), qread.read_top(os.path.join(d, "lig_w.top")))
    with open(inp) as f:
        text = f.read()
    with open(inp, "w") as f:
        f.write(text.replace("lrf\t99", "lrf\t5"))
    with pytest.raises(qread.QFileError):
        qread.build_system(inp)


@pytest.mark.parametrize("cfg", ["cfg1", "cfg2"])
def test_workloads_round_trip(cfg, tmp_path):
    """The benchmark workloads written by q6_amd.qfiles and read back give
    the generator's own tables (3,063 and 10,861 atoms); stock Qdyn6 builds
    the same tables from these files (tests/test_qfiles.py)."""
    from q6_amd import qfiles, synth
    s = getattr(synth, cfg)()
    r = qread.build_system(qfiles.write_case(s, str(tmp_path), 10, name=cfg, seed=0))
    s.finalize()
    assert [k for k in compare(r, s) if k != "winv"] == []
    assert np.array_equal(r.winv, 1.0 / np.asarray(s.md["mass"], float))
    assert np.array_equal(r.x, s.x)
