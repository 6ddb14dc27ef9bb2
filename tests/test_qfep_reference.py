"""q6_amd.enefile and q6_amd.fep against the reference's own programs.

  * The energy files stock Qdyn6 wrote for the 2-state lig_fep2 case, with
    and without Q bonded terms (tests/golden/lig_fep2_qdyn.en,
    lig_qb2_qdyn.en, made by tests/golden/make_enefile.py)
    reads back with consistent sums (EQ%total = Q bonded + qx + restraint,
    qx = qq + qp + qw, src/potene.f90:224-232), the run's lambdas and the
    off-diagonal H12 = 80 it was given, and the writer reproduces its
    per-frame records byte for byte.
  * Energy files written by q6_amd.enefile are read by stock Qfep6 (built
    from the reference's sources into oracle/_ref/), and its Part 1 (Zwanzig
    dG), Part 2 (per-window, per-bin diabatic and ground-state free energies,
    points, c1^2) and Part 3 (bin averages) equal q6_amd.fep's on the same
    series to the printed precision.  Skipped where Qfep6 is not built.
"""
import os
import tempfile

import numpy as np
import pytest

import qfep_ref
from conftest import ROOT
from q6_amd import enefile, fep

GOLD = os.path.join(ROOT, "tests", "golden", "lig_fep2_qdyn.en")
# the same run with Q bonded terms in the FEP file: nonzero EQ%q fields
GOLD_QB = os.path.join(ROOT, "tests", "golden", "lig_qb2_qdyn.en")


@pytest.mark.parametrize("path", [GOLD, GOLD_QB])
def test_reader_on_qdyn6_energy_file(path):
    d = enefile.read_ene(path, 2)
    assert d["canary"] == 1337
    assert np.array_equal(d["lambdas"], [0.7, 0.3])
    st = d["states"]
    assert st.shape == (11, 2, 14)
    total = st[..., 1:5].sum(-1) + st[..., 5:7].sum(-1) + st[..., 13]
    assert np.allclose(total, st[..., 0], rtol=0, atol=1e-12 * np.abs(st[..., 0]).max())
    assert np.allclose(st[..., 7:9] + st[..., 9:11] + st[..., 11:13], st[..., 5:7], atol=1e-12 * np.abs(st).max())
    assert all(len(o) == 1 and o[0][:3] == (1, 2, 80.0) and o[0][3] > 0 for o in d["offd"])


@pytest.mark.parametrize("path", [GOLD, GOLD_QB])
def test_writer_reproduces_qdyn6_records(path):
    d = enefile.read_ene(path, 2)
    if path == GOLD_QB:
        assert np.all(d["states"][:, :, 1:5] != 0.0)  # Q bond, angle, torsion, improper in both states
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "rt.en")
        enefile.write_ene(p, d["lambdas"], d["states"], offd_frames=d["offd"], version=d["version"])
        assert open(p, "rb").read() == open(path, "rb").read()


def test_frame_states_places_q_bonded_fields():
    """Batch frames -> energy-file records: the four Q bonded energies of state
    s (QNB_F_QB0 + 4s) go to q(1:4) and into EQ(s)%total with qx
    (src/potene.f90:224-231); md.q_state_energies gives the same totals."""
    from q6_amd import abi
    from q6_amd.md import q_state_energies
    rng = np.random.default_rng(3)
    fr = rng.normal(0.0, 10.0, (5, abi.QNB_FRAME_LEN))
    st = enefile.frame_states(fr, 2)
    for s in range(2):
        q = fr[:, abi.F_QB0 + 4 * s:abi.F_QB0 + 4 * s + 4]
        qx = fr[:, abi.F_Q0 + 6 * s:abi.F_Q0 + 6 * s + 6].reshape(5, 3, 2).sum(1)
        assert np.array_equal(st[:, s, 1:5], q)
        assert np.allclose(st[:, s, 5:7], qx, rtol=0, atol=1e-12)
        assert np.allclose(st[:, s, 0], q.sum(1) + qx.sum(1), rtol=0, atol=1e-12)
        assert np.allclose(q_state_energies(fr, 2)[:, s], st[:, s, 0], rtol=0, atol=1e-12)


@pytest.mark.skipif(not qfep_ref.available(), reason="reference Qfep6 not built (make -C oracle/ref)")
@pytest.mark.parametrize("seed", [1, 2])
def test_fep_and_evb_match_stock_qfep6(seed):
    rng = np.random.default_rng(seed)
    nw, npts, nskip, nbins, nmin, kT, H12 = 7, 40, 5, 12, 3, 0.596, 20.0
    alpha = np.array([0.0, 1.5])
    coeff = np.array([1.0, -1.0])
    lam1 = np.linspace(1, 0, nw)
    lams = np.stack([lam1, 1 - lam1], 1)
    Vraw = [np.stack([rng.normal(-20 + 7 * i, 2.5, npts), rng.normal(10 - 5 * i, 2.5, npts)], 1) for i in range(nw)]
    with tempfile.TemporaryDirectory() as tmp:
        files = []
        for i, V in enumerate(Vraw):
            st = np.zeros((npts, 2, 14))
            st[:, :, 0] = V
            st[:, :, 5] = V
            files.append(f"w{i}.en")
            enefile.write_ene(os.path.join(tmp, files[-1]), lams[i], st, offd=[(1, 2, H12, 3.0)])
        ref = qfep_ref.run(tmp, files, 2, kT, nskip, nbins, nmin, alpha, coeff)
    Vs = [fep.state_energies(V, alpha) for V in Vraw]
    _, _, dG = fep.zwanzig(Vs, lams, kT, nskip)
    got1 = np.array([r[5] for r in ref["part1"]])
    assert np.abs(got1 - dG).max() <= 6e-4
    res = fep.evb_pmf(Vs, lams, H12, kT, coeff, nbins, nmin, nskip)
    rows = ref["part2"]
    assert len(rows) == int(np.sum(res["npts"] >= nmin))
    for r in rows:
        i = int(np.argmin(np.abs(lam1 - r["lam"])))
        k = r["bin"]
        assert r["pts"] == res["npts"][i, k]
        assert abs(r["gap"] - res["centres"][k]) <= 6e-3
        assert abs(r["dGg"] - res["dGg"][i, k]) <= 6e-3
        assert abs(r["dGa"] - res["dGv"][i, k, 0]) <= 6e-3
        assert abs(r["dGb"] - res["dGv"][i, k, 1]) <= 6e-3
        assert abs(r["c1"] - res["c1"][i, k]) <= 6e-4
    for r in ref["part3"]:
        k = r["bin"]
        assert r["pts"] == res["ptsum"][k]
        assert abs(r["dGg"] - res["binned"][k]) <= 6e-3
        assert abs(r["c1"] - res["binned_c1"][k]) <= 6e-4


@pytest.mark.skipif(not qfep_ref.available(), reason="reference Qfep6 not built (make -C oracle/ref)")
def test_distance_dependent_coupling_matches_stock_qfep6():
    """Off-diagonal records carry a per-frame r_kl (as the batch samples it);
    Qfep6 given A, mu, eta, r0 computes H12 from it (qfep.f90:476-480)."""
    rng = np.random.default_rng(3)
    nw, npts, nskip, nbins, nmin, kT = 6, 50, 4, 10, 3, 0.596
    A, mu, eta, r0 = 35.0, 0.6, 0.05, 2.0
    alpha = np.array([0.0, -2.0])
    coeff = np.array([1.0, -1.0])
    lam1 = np.linspace(1, 0, nw)
    lams = np.stack([lam1, 1 - lam1], 1)
    Vraw = [np.stack([rng.normal(-15 + 6 * i, 2.0, npts), rng.normal(12 - 5 * i, 2.0, npts)], 1) for i in range(nw)]
    rkl = [rng.uniform(2.0, 4.5, npts) for _ in range(nw)]
    with tempfile.TemporaryDirectory() as tmp:
        files = []
        for i, V in enumerate(Vraw):
            st = np.zeros((npts, 2, 14))
            st[:, :, 0] = V
            st[:, :, 5] = V
            files.append(f"w{i}.en")
            enefile.write_ene(os.path.join(tmp, files[-1]), lams[i], st,
                              offd_frames=[[(1, 2, 0.0, r)] for r in rkl[i]])
        ref = qfep_ref.run(tmp, files, 2, kT, nskip, nbins, nmin, alpha, coeff, couplings=[(1, 2, A, mu, eta, r0)])
    Vs = [fep.state_energies(V, alpha) for V in Vraw]
    H = [fep.offdiag_coupling(r, A, mu, eta, r0) for r in rkl]
    res = fep.evb_pmf(Vs, lams, H, kT, coeff, nbins, nmin, nskip, rkl=rkl)
    rows = ref["part2"]
    assert len(rows) == int(np.sum(res["npts"] >= nmin))
    for r in rows:
        i = int(np.argmin(np.abs(lam1 - r["lam"])))
        k = r["bin"]
        assert r["pts"] == res["npts"][i, k]
        assert abs(r["dGg"] - res["dGg"][i, k]) <= 6e-3
        assert abs(r["c1"] - res["c1"][i, k]) <= 6e-4
    assert ref["part3"]
    for r in ref["part3"]:
        k = r["bin"]
        assert r["pts"] == res["ptsum"][k]
        assert abs(r["dGg"] - res["binned"][k]) <= 6e-3
        assert abs(r["c1"] - res["binned_c1"][k]) <= 6e-4
        assert abs(r["r"] - res["binned_r"][k]) <= 6e-3
    # a constant coupling is the mu = eta = 0 case
    Hc = [fep.offdiag_coupling(r, A, 0.0) for r in rkl]
    a = fep.evb_window_free_energy(Vs, lams, Hc, kT, nskip)
    b = fep.evb_window_free_energy(Vs, lams, A, kT, nskip)
    assert np.array_equal(a, b)
