"""Q input files of the golden cases, written from the committed fixtures
(tests/golden/lig_w.top and lig_w.fep are the reference's
tests/basic_tests/prep_SPH system) and the input templates of
tests/golden/make_golden.py, so a test can set each case up from files."""
import os
import shutil
import sys

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import make_golden as mg  # noqa: E402

LIG_CASES = ["lig_eq1", "lig_fep2", "lig_fep2_nolrf", "lig_fep2_arith", "lig_fep2_anyatom"]
CUT = "solute_solute 16\nsolvent_solvent 9\nsolute_solvent 16\n"
CJP_DIR = "/root/reference/tests/exclude_tests/inputs"


def _fixture(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def write_lig_case(name, d):
    """Writes lig_w.top, the FEP file and run.inp of a golden case into d;
    returns the input's path."""
    from q6_amd import qread
    top = _fixture("lig_w.top")
    q = list(qread.read_top(os.path.join(GOLDEN, "lig_w.top"))["crg"])
    fep2 = mg.fep2_text(q)
    lam = "[lambdas]\n0.7 0.3\n"
    inp, fep = {
        "lig_eq1": (mg.EQ1_SPH.format(lrf="on", rq=99, extra_cut="", fep="lig_w.fep", lambdas=""), None),
        "lig_fep2": (mg.EQ1_SPH.format(lrf="on", rq=12, extra_cut=CUT, fep="fep2.fep", lambdas=lam), fep2),
        "lig_fep2_nolrf": (mg.EQ1_SPH.format(lrf="off", rq=12, extra_cut=CUT, fep="fep2.fep", lambdas=lam), fep2),
        "lig_fep2_arith": (mg.EQ1_SPH.format(lrf="on", rq=12, extra_cut=CUT, fep="fep2.fep", lambdas=lam),
                           mg.arith_fep(fep2)),
        "lig_fep2_anyatom": (mg.EQ1_SPH.format(lrf="on", rq=12, extra_cut=CUT, fep="fep2.fep", lambdas=lam), fep2),
    }[name]
    if name == "lig_fep2_arith":
        top = mg.arith_top(top)
    elif name == "lig_fep2_anyatom":
        top = mg.anyatom_top(top)
    with open(os.path.join(d, "lig_w.top"), "w") as f:
        f.write(top)
    shutil.copy(os.path.join(GOLDEN, "lig_w.fep"), d)
    if fep is not None:
        with open(os.path.join(d, "fep2.fep"), "w") as f:
            f.write(fep)
    with open(os.path.join(d, "run.inp"), "w") as f:
        f.write(inp)
    return os.path.join(d, "run.inp")


def write_cjp_case(d):
    """The reference's 2CJP group-contribution system (read from the
    reference's own test folder; None where it is absent)."""
    if not os.path.isdir(CJP_DIR):
        return None
    for f in ("2cjpFH_ionres_oplsa.top", "lig.fep"):
        shutil.copy(os.path.join(CJP_DIR, f), d)
    with open(os.path.join(d, "run.inp"), "w") as f:
        f.write(mg.CJP_INP)
    return os.path.join(d, "run.inp")
