"""The Q bonded oracle (oracle/qbonded.py) pinned on the CPU.

1. Against the reference's own routines: tests/golden/lig_qb2.npz holds three
   passes of pot_energy's per-state loop (src/potene.f90:183-190) run by the
   reference's qbond/qangle/qurey_bradley/qtorsion/qimproper on its ligand,
   with the Morse coupling state after each pass.
2. Independently of the reference, against central finite differences of the
   oracle's own energy sum_s lambda_s E_s: where every coupled bond is present
   in every state and couplings scale with EMorseD (flag 0), the reference's
   gradient is the derivative of its energy.
3. The one place where it is not: a term coupled with flag 1 (scaled by
   1 - EMorseD) still receives +dMorse * E (src/bondene.f90:431-437), the
   derivative with the sign of the flag-0 form, so the gradient exceeds the
   energy's derivative by 2 lambda E dEMorseD/dx on the bond's atoms.  The
   oracle and the device kernel keep the reference's form.
"""
import math

import numpy as np
import pytest

from conftest import golden_path
from oracle import qbonded as qb

REL = 1e-12


def golden():
    z = np.load(golden_path("lig_qb2"))
    return z, qb.from_dump(z)


def test_oracle_matches_reference_passes():
    z, tb = golden()
    x, lam = z["x"].T, z["lambda"]
    st = qb.new_state(len(tb["bond_ij"]))
    for p, urey in ((1, False), (2, False), (3, True)):
        E, d, st = qb.qbonded(x, tb, lam, st, urey=urey)
        Er = z[f"E_qb{p}"].reshape(-1, 4)
        dr = z[f"d_qb{p}"].T
        assert np.all(np.abs(E - Er) <= REL * np.maximum(1.0, np.abs(Er))), (p, E, Er)
        assert np.max(np.abs(d - dr)) <= REL * np.max(np.abs(dr)), p
        assert np.array_equal(st[0], z[f"EMorseD{p}"])
        assert np.allclose(st[1], z[f"dMorse_i{p}"].T, rtol=1e-14, atol=0)
    # passes 1 and 2 differ only through the Morse state a term read from a
    # bond absent in its state (torsion 2 and improper 2 are coupled to bond
    # 4, absent in state 1): the state carries over between calls
    assert z["E_qb1"][2] != z["E_qb2"][2] and z["E_qb1"][3] != z["E_qb2"][3]


def synthetic(flag=0, urey=True):
    """Two states; eight atoms in a puckered chain; Morse bonds present in
    both states (with different parameters) so every coupling reads the
    Morse state of the same state."""
    k = np.arange(8)
    x = np.stack([1.2 * k, 0.75 * (-1.0) ** k + 0.1 * k, 0.4 * np.sin(1.3 * k)], axis=1)
    tb = dict(
        bond_ij=np.array([[0, 1], [1, 2], [2, 3], [5, 6]]),
        bond_cod=np.array([[1, 2], [2, 1], [3, 3], [4, 1]]),
        bond_lib=np.array([[90.0, 1.2, 0.9, 0.0], [70.0, 1.1, 1.0, 0.0], [0.0, 0.0, 1.4, 600.0],
                           [50.0, 0.9, 0.8, 0.0]]),
        angle_ijk=np.array([[0, 1, 2], [1, 2, 3], [4, 5, 6], [5, 6, 7]]),
        angle_cod=np.array([[1, 2], [2, 2], [1, 0], [3, 1]]),
        angle_lib=np.array([[60.0, 1.9, 20.0, 2.4], [45.0, 2.0, 0.0, 0.0], [80.0, 1.7, 15.0, 2.2]]),
        torsion_ijkl=np.array([[0, 1, 2, 3], [3, 4, 5, 6], [4, 5, 6, 7]]),
        torsion_cod=np.array([[1, 2], [2, 1], [3, 3]]),
        torsion_lib=np.array([[1.5, 3.0, 0.0], [0.8, 2.0, math.pi], [2.0, 1.0, 0.6]]),
        improper_ijkl=np.array([[1, 0, 2, 3], [5, 4, 6, 7]]),
        improper_cod=np.array([[1, 2], [2, 0]]),
        improper_lib=np.array([[10.0, 0.3], [5.0, math.pi]]),
        angle_coupl=np.array([[0, 0, 0], [3, 3, flag]]),
        torsion_coupl=np.array([[0, 1, 0], [2, 3, flag]]),
        improper_coupl=np.array([[1, 3, flag]]),
    )
    return x, tb, np.array([0.35, 0.65]), urey


def total(x, tb, lam, urey):
    E, _, _ = qb.qbonded(x, tb, lam, qb.new_state(4), urey=urey)
    return float(lam @ E.sum(1))


def fd_grad(x, tb, lam, urey, h=1e-6):
    g = np.zeros_like(x)
    for i in range(x.shape[0]):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            g[i, c] = (total(xp, tb, lam, urey) - total(xm, tb, lam, urey)) / (2 * h)
    return g


@pytest.mark.parametrize("urey", [False, True])
def test_oracle_gradient_is_energy_derivative(urey):
    x, tb, lam, _ = synthetic(flag=0, urey=urey)
    E, d, st = qb.qbonded(x, tb, lam, qb.new_state(4), urey=urey)
    cpl = [0, 1, 3]  # the coupled (Morse) bonds are stretched: their couplings are active
    assert st[0][2] == 1.0  # harmonic bond: constant EMorseD
    assert np.all(st[0][cpl] < 0.99) and np.all(st[0][cpl] > 0.01), st[0]
    g = fd_grad(x, tb, lam, urey)
    assert np.max(np.abs(d - g)) <= 1e-6 * np.max(np.abs(g)), np.max(np.abs(d - g))


def test_flag1_coupling_keeps_reference_sign():
    """A Morse bond (atoms 0-1, one library row in both states) and an angle
    2-3-4 present in state 1 only, coupled with flag 1: E_angle = ae (1 - g).
    The reference adds +dMorse_i ae on atom 0 (and -dMorse_i ae on atom 1), the
    flag-0 sign, so d exceeds dE/dx there by 2 lambda_1 ae dg/dx and nowhere
    else."""
    x = np.array([[0.0, 0.0, 0.0], [1.45, 0.2, -0.1], [3.0, 1.0, 0.5], [4.1, 0.3, 1.2], [5.0, 1.4, 1.9]])
    lam = np.array([0.3, 0.7])
    fk, th0 = 60.0, 1.8
    tb = dict(bond_ij=np.array([[0, 1]]), bond_cod=np.array([[1, 1]]), bond_lib=np.array([[90.0, 1.9, 1.1, 0.0]]),
              angle_ijk=np.array([[2, 3, 4]]), angle_cod=np.array([[1, 0]]), angle_lib=np.array([[fk, th0, 0.0, 0.0]]),
              angle_coupl=np.array([[0, 0, 1]]))

    def tot(xx):
        E, _, _ = qb.qbonded(xx, tb, lam, qb.new_state(1))
        return float(lam @ E.sum(1))

    E, d, st = qb.qbonded(x, tb, lam, qb.new_state(1))
    g = np.zeros_like(x)
    h = 1e-6
    for i in range(5):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            g[i, c] = (tot(xp) - tot(xm)) / (2 * h)
    u, w = x[2] - x[3], x[4] - x[3]
    ae = 0.5 * fk * (math.acos(u @ w / (np.linalg.norm(u) * np.linalg.norm(w))) - th0) ** 2
    # the state left after the call is state 2's (lambda 0.7); state 1's is
    # the same vector scaled by lambda_1 / lambda_2 (one library row)
    dmi1 = st[1][0] * lam[0] / lam[1]
    assert 0.0 < st[0][0] < 1.0
    assert np.allclose(d[0] - g[0], 2 * ae * dmi1, rtol=1e-6, atol=1e-9)
    assert np.allclose(d[1] - g[1], -2 * ae * dmi1, rtol=1e-6, atol=1e-9)
    assert np.max(np.abs(d[2:] - g[2:])) <= 1e-6 * np.max(np.abs(g))
    assert np.linalg.norm(dmi1) * ae > 1e-2  # the case exercises the coupling
