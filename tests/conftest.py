import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["lig_eq1", "lig_fep2", "lig_fep2_nolrf", "lig_fep2_arith", "lig_fep2_anyatom", "lig_fep2_anyatom_nolrf",
         "cjp_excl", "eoh", "eoh_pol"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")


def golden_path(name):
    return os.path.join(GOLDEN, name + ".npz")


def watpol_frames(s, k):
    """Coordinates of call k (1..7) of the dump's polarisation-restraint
    sequence (oracle/ref/qnb_dump.f90): every solvent atom (1-based index i)
    displaced by ((7i + 3k) mod 11 - 5, (5i + 2k) mod 13 - 6, (3i + 5k) mod 7 - 3)
    times 2^-7 A, which both sides form exactly."""
    import numpy as np
    nsol, nat = int(s.n["nat_solute"]), int(s.n["natom"])
    i = np.arange(nsol + 1, nat + 1)
    x = np.array(s.x, dtype=np.float64)
    x[nsol:, 0] = s.x[nsol:, 0] + ((7 * i + 3 * k) % 11 - 5).astype(np.float64) * 0.0078125
    x[nsol:, 1] = s.x[nsol:, 1] + ((5 * i + 2 * k) % 13 - 6).astype(np.float64) * 0.0078125
    x[nsol:, 2] = s.x[nsol:, 2] + ((3 * i + 5 * k) % 7 - 3).astype(np.float64) * 0.0078125
    return x


def golden_summary_inputs(s):
    """The terms pot_energy sums besides the nonbonded engine's, from a
    golden dump: (bonded dict, per-state Q bonded energies, per-state
    restraint energies, lambdas)."""
    eb, eqs = s.ref["E_bonded"], s.ref["EQ_summary"]
    bonded = dict(p_bond=eb[0], w_bond=eb[1], p_angle=eb[2], w_angle=eb[3], p_torsion=eb[4], p_improper=eb[5],
                  restraint_fix=eb[10], restraint_shell=eb[11], restraint_protein=eb[12] - float(
                      (eqs[14] * eqs[0]).sum()))
    return bonded, eqs[2:6].T, eqs[14], eqs[0]


def exc_inputs(s):
    """calculate_exc's inputs and stock Qdyn6's results from a golden dump
    with group contributions: masks (narr, nat_solute), types (narr,), the
    per-state energies pot_energy left (nstates, 11) and the reference's
    arrays (narr, nstates, 7)."""
    import numpy as np
    r = s.ref
    eqs = np.asarray(r["EQ_summary"])
    eq_in = np.stack([eqs[[8, 9, 10, 11, 12, 13, 2, 3, 4, 5, 14], st] for st in range(eqs.shape[1])])
    masks = np.stack([np.asarray(r["exc_mask%d" % g]) for g in np.atleast_1d(r["exc_group"])]).astype(np.uint8)
    return masks, np.atleast_1d(r["exc_type"]).astype(np.int32), eq_in, np.asarray(r["exc_EQ"]).transpose(2, 1, 0)
