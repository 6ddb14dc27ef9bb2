"""Writes tests/golden/lig_fep2_qdyn.en (and, with `qb`, lig_qb2_qdyn.en):
energy files written by stock Qdyn6.

Runs the reference's serial Qdyn6 (built from its sources by
`make -C oracle/ref` into oracle/_ref/) for 60 steps of the 2-state lig_fep2
case of make_golden.py (the reference's SPH test system
tests/basic_tests/prep_SPH/lig_w with the FEP file written there), with an
energy record every 5 steps.  The file is the reference's own output, kept
as a fixture for the energy-file reader/writer (q6_amd.enefile) and for
running stock Qfep6 on it in the tests.  Development container only.

lig_qb2_qdyn.en runs the same input with the lig_qb2 FEP file of
make_golden.py (Q bonds, angles with Urey-Bradley parameters, torsions,
impropers and Morse couplings), so its records carry the per-state Q bonded
energies (EQ%q%bond .. %improper).

Usage: python tests/golden/make_enefile.py [qb]
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import make_golden as mg  # noqa: E402

QDYN = os.path.join(REPO, "oracle", "_ref", "Qdyn6")
OUT = os.path.join(HERE, "lig_fep2_qdyn.en")


def main():
    qb = "qb" in sys.argv[1:]
    out = os.path.join(HERE, "lig_qb2_qdyn.en") if qb else OUT
    if not os.path.exists(QDYN):
        sys.exit("build the reference first: make -C oracle/ref")
    inp = mg.EQ1_SPH.format(lrf="on", rq=12, extra_cut="solute_solute 16\nsolvent_solvent 9\nsolute_solvent 16\n",
                            fep="fep2.fep", lambdas="[lambdas]\n0.7 0.3\n")
    inp = inp.replace("steps                    100", "steps                    60")
    inp = inp.replace("output                    50", "output                    50\nenergy                    5")
    inp = inp.replace("final                     eq1.re", "final                     eq1.re\nenergy                    run.en")
    tmp = tempfile.mkdtemp(prefix="qnbene_")
    try:
        shutil.copy(os.path.join(mg.REF, "prep_SPH", "lig_w.top"), tmp)
        with open(os.path.join(tmp, "fep2.fep"), "w") as f:
            f.write(mg.fep2_text(mg.top_charges()) + (mg.QB_TEXT if qb else ""))
        with open(os.path.join(tmp, "run.inp"), "w") as f:
            f.write(inp)
        log = subprocess.run([QDYN, "run.inp"], cwd=tmp, capture_output=True, text=True)
        if log.returncode != 0 or not os.path.exists(os.path.join(tmp, "run.en")):
            print(log.stdout[-3000:], log.stderr[-3000:])
            raise RuntimeError("Qdyn6 failed")
        shutil.copy(os.path.join(tmp, "run.en"), out)
        print(out, os.path.getsize(out), "bytes")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
