"""Q-atom bonded terms per state on the GPU (k_qbonded) against the oracle and
the reference's golden vectors.

Tolerances (FP64): energies within 1e-10 relative (of max(|E|, 1)); gradients
within 1e-10 of the largest gradient component.  The golden case runs the
reference's own qbond/qangle/qurey_bradley/qtorsion/qimproper
(tests/golden/lig_qb2.npz); the randomized case has several hundred terms of
each kind, so every loop of the kernel runs more than one round of its block.
"""
import numpy as np
import pytest

from conftest import golden_path
from oracle import qbonded as qb
from q6_amd import abi
from q6_amd.system import from_golden

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(E, Er, d, dr):
    assert np.all(np.abs(E - Er) <= REL * np.maximum(1.0, np.abs(Er))), (E, Er)
    assert np.max(np.abs(d - dr)) <= REL * max(np.max(np.abs(dr)), 1.0), np.max(np.abs(d - dr))


@pytest.fixture(scope="module")
def lig():
    from q6_amd.engine import NonbondedEngine
    z = np.load(golden_path("lig_qb2"))
    s = from_golden(golden_path("lig_fep2"))
    assert s.n["natom"] == int(z["natom"]) and np.array_equal(s.par["lambda"][:2], z["lambda"])
    print("\n[gpu] Q bonded on lig_fep2 (3397 atoms, 2 states)", flush=True)
    eng = NonbondedEngine(s)
    yield z, s, eng
    eng.close()


def test_golden_passes(lig):
    """Three passes as the reference ran them: passes 1 and 2 in a row (the
    second reads the Morse state the first left), pass 3 with Urey-Bradley
    after one warm-up call, so that it starts from the same Morse state."""
    z, s, eng = lig
    tb = qb.from_dump(z)
    x = np.ascontiguousarray(z["x"].T)
    eng.set_qbonded(tb)
    for p in (1, 2):
        E, d = eng.qbonded(x)
        close(E, z[f"E_qb{p}"].reshape(-1, 4), d, z[f"d_qb{p}"].T)
    eng.set_qbonded(tb, urey=True)
    eng.qbonded(x)
    E, d = eng.qbonded(x)
    close(E, z["E_qb3"].reshape(-1, 4), d, z["d_qb3"].T)
    # the gradient is added into d
    d0 = np.ones_like(x)
    _, d1 = eng.qbonded(x, d=d0.copy())
    close(np.zeros(1), np.zeros(1), d1 - d0, d)
    eng.set_qbonded(None)
    E, d = eng.qbonded(x)
    assert not E.any() and not d.any()


def random_set(rng, natom, nb=300, na=400, nt=500, ni=300, ns=2):
    """Terms over a pool of atoms, random per-state codes (0 = absent in
    about 1 state in 5), couplings of every kind with both flags, and several
    couplings on one term (the first or last one counts)."""
    pool = rng.choice(natom, 60, replace=False)

    def atoms(n, k):
        out = np.zeros((n, k), int)
        for t in range(n):
            out[t] = rng.choice(pool, k, replace=False)
        return out

    def cods(n, nlib):
        c = rng.integers(1, nlib + 1, (n, ns))
        c[rng.random((n, ns)) < 0.2] = 0
        return c

    bl = np.column_stack([rng.uniform(20, 100, 6), rng.uniform(0.5, 2.5, 6), rng.uniform(2.0, 5.0, 6),
                          np.zeros(6)])
    bl[4:, :2] = 0.0
    bl[4:, 3] = rng.uniform(100, 600, 2)  # two harmonic rows
    al = np.column_stack([rng.uniform(30, 90, 5), rng.uniform(1.5, 2.3, 5), rng.uniform(0, 30, 5),
                          rng.uniform(2.0, 3.0, 5)])
    al[0, 2] = 0.0  # one row without Urey-Bradley
    tl = np.column_stack([rng.uniform(0.5, 3, 5), rng.integers(1, 4, 5).astype(float), rng.uniform(-np.pi, np.pi, 5)])
    il = np.column_stack([rng.uniform(2, 20, 4), rng.uniform(-np.pi, np.pi, 4)])

    def coupl(nterm, ncp):
        return np.column_stack([rng.integers(0, nterm, ncp), rng.integers(0, 4, ncp), rng.integers(0, 2, ncp)])

    return dict(bond_ij=atoms(nb, 2), bond_cod=cods(nb, 6), bond_lib=bl,
                angle_ijk=atoms(na, 3), angle_cod=cods(na, 5), angle_lib=al,
                torsion_ijkl=atoms(nt, 4), torsion_cod=cods(nt, 5), torsion_lib=tl,
                improper_ijkl=atoms(ni, 4), improper_cod=cods(ni, 4), improper_lib=il,
                angle_coupl=coupl(na, 120), torsion_coupl=coupl(nt, 150), improper_coupl=coupl(ni, 90))


def test_random_terms_against_oracle(lig):
    """Four calls in a row at slightly moved coordinates: energies, gradient
    and the Morse state carried between calls match the oracle's."""
    z, s, eng = lig
    rng = np.random.default_rng(11)
    tb = random_set(rng, s.n["natom"])
    x = np.ascontiguousarray(z["x"].T)
    lam = z["lambda"]
    eng.set_qbonded(tb, urey=True)
    st = qb.new_state(len(tb["bond_ij"]))
    for call in range(4):
        xc = x + rng.normal(0, 0.02, x.shape)
        E, d = eng.qbonded(xc)
        Er, dr, st = qb.qbonded(xc, tb, lam, st, urey=True)
        close(E, Er, d, dr)
    eng.set_qbonded(None)


def test_bad_tables_raise(lig):
    z, s, eng = lig
    tb = qb.from_dump(z)
    bad = dict(tb, bond_ij=np.array(tb["bond_ij"]) * 0 + s.n["natom"])
    with pytest.raises(abi.QnbError):
        eng.set_qbonded(bad)
    bad = dict(tb, angle_coupl=np.array([[0, 99, 0]]))
    with pytest.raises(abi.QnbError):
        eng.set_qbonded(bad)
    bad = dict(tb, torsion_cod=np.array(tb["torsion_cod"]) + 50)
    with pytest.raises(abi.QnbError):
        eng.set_qbonded(bad)
    # a failed set leaves no terms loaded
    E, d = eng.qbonded(np.ascontiguousarray(z["x"].T))
    assert not E.any() and not d.any()


def chain_terms(s):
    """Q bonded terms on cfg1's 12-atom Q chain: Morse bonds present in both
    states (so every coupling reads the Morse state of its own call)."""
    r0 = float(np.asarray(s.md["bond_r0"]).ravel()[0])
    return dict(bond_ij=np.array([[0, 1], [1, 2], [5, 6]]), bond_cod=np.array([[1, 2], [2, 1], [1, 1]]),
                bond_lib=np.array([[60.0, 2.0, r0 - 0.05, 0.0], [40.0, 1.5, r0 - 0.1, 0.0]]),
                angle_ijk=np.array([[0, 1, 2], [1, 2, 3], [4, 5, 6]]), angle_cod=np.array([[1, 2], [2, 0], [1, 1]]),
                angle_lib=np.array([[30.0, 1.9, 0.0, 0.0], [20.0, 2.0, 0.0, 0.0]]),
                torsion_ijkl=np.array([[0, 1, 2, 3], [3, 4, 5, 6]]), torsion_cod=np.array([[1, 1], [2, 1]]),
                torsion_lib=np.array([[1.0, 3.0, 0.0], [0.5, 2.0, np.pi]]),
                improper_ijkl=np.array([[5, 4, 6, 7]]), improper_cod=np.array([[1, 1]]),
                improper_lib=np.array([[5.0, 0.4]]),
                angle_coupl=np.array([[0, 0, 0], [2, 2, 1]]), torsion_coupl=np.array([[0, 1, 0]]),
                improper_coupl=np.array([[0, 2, 1]]))


def test_batch_md_records_and_applies_qbonded():
    """In the batched MD step: the sampled per-state Q bonded energies equal
    the single engine's at the stored coordinates of the same steps, and
    after one step the only displacement the terms add to a batch without
    them is -dt^2 winv s_T d_qb (leap-frog, src/md.f90:1451-1466)."""
    from q6_amd import synth
    from q6_amd.engine import NonbondedEngine
    from q6_amd.md import BatchMD
    s = synth.cfg1()
    v = synth.maxwell_velocities(s, 300.0, 4)
    tb = chain_terms(s)
    lam = np.array([0.4, 0.6])
    print("\n[gpu] batch MD with Q bonded terms", flush=True)
    b = BatchMD([s, s], [lam, lam], [v, v], nbcycle=10, ene_every=10, max_frames=8)
    b.set_qbonded(0, tb)
    b.store_frames(10, 8)
    b.run(40)
    b.sync()
    e0, e1, fr = b.energies(0), b.energies(1), b.frames(0)
    b.close()
    assert len(e0) == len(fr) == 4
    assert not e1[:, abi.F_QB0:abi.F_QB0 + 8].any()
    eng = NonbondedEngine(s, **{"lambda": lam})
    eng.set_qbonded(tb)
    for f in range(len(fr)):
        E, _ = eng.qbonded(fr[f])
        rec = e0[f, abi.F_QB0:abi.F_QB0 + 8].reshape(2, 4)
        assert np.all(np.abs(rec - E) <= REL * np.maximum(1.0, np.abs(E))), (f, rec, E)
    # one step from the same start with and without the terms
    _, dq = eng.qbonded(s.x)
    eng.close()
    xs = []
    for with_qb in (True, False):
        c = BatchMD([s], [lam], [v], nbcycle=10, ene_every=10, max_frames=2)
        if with_qb:
            c.set_qbonded(0, tb)
        c.run(1, graph=False)
        c.sync()
        xs.append(c.state(0)[0])
        c.close()
    dx = xs[0] - xs[1]
    winv = 1.0 / np.asarray(s.md["mass"], float)
    dt = 0.020462 * 1.0
    pred = -(dt * dt) * winv[:, None] * dq
    moved = np.abs(pred) > 1e-12
    sT = np.sum(dx[moved] * pred[moved]) / np.sum(pred[moved] ** 2)
    assert 0.8 < sT < 1.2, sT
    assert np.max(np.abs(dx - sT * pred)) <= 1e-9 * np.max(np.abs(pred)) + 1e-15
