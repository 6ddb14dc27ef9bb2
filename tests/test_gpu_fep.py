"""Qfep6's free-energy analysis on the device (qnb_free_energy,
qnb_batch_free_energy) against the host restatement q6_amd.fep, which
tests/test_qfep_reference.py pins to stock Qfep6 on the same kind of series:
Part 1 (Zwanzig dgf, dgr, dG), Part 2 (per window and energy-gap bin: points,
ground-state and diabatic free energies, c1^2) and Part 3 (bin averages).
Point counts must be equal; energies within 1e-9 kcal/mol (the device sums in
frame order without fused multiply-adds, the host with numpy's pairwise sums
and its own exp/log)."""
import numpy as np
import pytest

from q6_amd import abi, fep

pytestmark = pytest.mark.gpu
TOL = 1e-9


def series(seed, nw=7, npts=40):
    rng = np.random.default_rng(seed)
    lam1 = np.linspace(1, 0, nw)
    lams = np.stack([lam1, 1 - lam1], 1)
    V = [np.stack([rng.normal(-20 + 7 * i, 2.5, npts), rng.normal(10 - 5 * i, 2.5, npts)], 1) for i in range(nw)]
    rkl = [rng.uniform(2.0, 4.5, npts) for _ in range(nw)]
    return lams, V, rkl


def close(a, b, tol=TOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(a)
    assert np.all(np.abs(a[m] - b[m]) <= tol * np.maximum(1.0, np.abs(b[m]))), np.max(np.abs(a[m] - b[m]))


@pytest.mark.parametrize("seed", [1, 2])
def test_random_series_match_host(seed):
    from q6_amd.md import free_energy
    print(f"\n[gpu] device Qfep analysis, random series seed {seed}", flush=True)
    lams, V, rkl = series(seed)
    kT, nskip, nbins, nmin = 0.596, 5, 12, 3
    A, mu, eta, r0 = 35.0, 0.6, 0.05, 2.0
    alpha, coeff = np.array([0.0, 1.5]), np.array([1.0, -1.0])
    d = free_energy(V, lams, kT, rkl=rkl, nskip=nskip, nbins=nbins, nmin=nmin, alpha=alpha, coeff=coeff, A=A, mu=mu,
                    eta=eta, r0=r0)
    Vs = [fep.state_energies(v, alpha) for v in V]
    dgf, dgr, dG = fep.zwanzig(Vs, lams, kT, nskip)
    close(d["dgf"], dgf)
    close(d["dgr"], dgr)
    close(d["dG"], dG)
    H = [fep.offdiag_coupling(r, A, mu, eta, r0) for r in rkl]
    h = fep.evb_pmf(Vs, lams, H, kT, coeff, nbins, nmin, nskip, rkl=rkl)
    assert np.array_equal(d["npts"], h["npts"])
    assert np.array_equal(d["ptsum"], h["ptsum"])
    assert int((h["npts"] >= nmin).sum()) > nbins  # bins reported in several windows
    close(d["centres"], h["centres"])
    close(d["dGg"], h["dGg"])
    close(d["dGv"], h["dGv"])
    close(d["c1"], h["c1"])
    close(d["binned"], h["binned"])
    close(d["binned_c1"], h["binned_c1"])
    close(d["binned_r"], h["binned_r"])


def test_constant_gap_closed_form():
    """Every frame of every window has V2 - V1 = 3.25: each forward step
    between neighbouring windows perturbs by dlambda * 3.25 exactly, so dG of
    the last window is 3.25 (to the rounding of exp and log), and so is every
    reverse step's negative."""
    from q6_amd.md import free_energy
    nw, npts = 5, 30
    lam1 = np.linspace(1, 0, nw)
    lams = np.stack([lam1, 1 - lam1], 1)
    rng = np.random.default_rng(7)
    V = []
    for i in range(nw):
        e1 = rng.normal(-5.0, 3.0, npts)
        V.append(np.stack([e1, e1 + 3.25], 1))
    d = free_energy(V, lams, 0.596, nskip=0)
    assert abs(d["dG"][-1] - 3.25) <= 1e-12
    assert np.allclose(d["dgf"], 3.25 / (nw - 1), rtol=0, atol=1e-13)
    assert np.allclose(d["dgr"], -3.25 / (nw - 1), rtol=0, atol=1e-13)


def test_refuses_bad_input():
    from q6_amd.md import free_energy
    lams, V, _ = series(3, nw=3, npts=10)
    with pytest.raises(abi.QnbError):
        free_energy(V, lams, 0.596, nskip=10)  # nothing left after the skipped frames
    with pytest.raises(abi.QnbError):
        free_energy(V, lams, -1.0)
    with pytest.raises(abi.QnbError):  # three states: no closed-form ground state for Part 2
        free_energy([np.c_[v, v[:, :1]] for v in V], np.c_[lams, np.zeros(3)], 0.596, nbins=4)


def test_batch_frames_match_host():
    """Three windows of the 20 A system run on the device with an EVB
    off-diagonal sampled with the energies; the analysis of the recorded
    frames equals the host's on the same frames (state energies formed as
    pot_energy forms EQ(s)%total, H12 from the recorded r_kl)."""
    from q6_amd import synth
    from q6_amd.engine import pot_energy_summary
    from q6_amd.md import BatchMD
    s = synth.cfg1()
    v = synth.maxwell_velocities(s, 300.0, 4)
    lams = np.array([[1.0, 0.0], [0.5, 0.5], [0.0, 1.0]])
    print("\n[gpu] device Qfep analysis of batch frames", flush=True)
    b = BatchMD([s] * 3, list(lams), [v] * 3, precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=5, max_frames=40)
    for k in range(3):
        b.set_offdiag(k, [(0, 1)], [80.0], [0.0])
    b.run(150)
    b.sync()
    kT, nskip, nbins, nmin = 2.0, 4, 6, 2
    A, mu = 40.0, 0.3
    coeff = np.array([1.0, -1.0])
    d = b.free_energy(kT, nskip=nskip, nbins=nbins, nmin=nmin, coeff=coeff, A=A, mu=mu, r0=1.5)
    frames = [b.energies(k) for k in range(3)]
    b.close()
    Vs, rkl = [], []
    for fr in frames:
        Vw = np.zeros((len(fr), 2))
        for f, row in enumerate(fr):
            EQ = [dict(qq_el=row[abi.F_Q0 + 6 * st], qq_vdw=row[abi.F_Q0 + 6 * st + 1],
                       qp_el=row[abi.F_Q0 + 6 * st + 2], qp_vdw=row[abi.F_Q0 + 6 * st + 3],
                       qw_el=row[abi.F_Q0 + 6 * st + 4], qw_vdw=row[abi.F_Q0 + 6 * st + 5]) for st in range(2)]
            E = dict(pp_el=0.0, pp_vdw=0.0, pw_el=0.0, pw_vdw=0.0, ww_el=0.0, ww_vdw=0.0, lrf=0.0,
                     solvent_radial=0.0, water_pol=0.0)
            qb = [row[abi.F_QB0 + 4 * st:abi.F_QB0 + 4 * st + 4] for st in range(2)]
            _, eqs = pot_energy_summary(E, EQ, [0.5, 0.5], EQ_bonded=qb)
            Vw[f] = [eqs[0]["total"], eqs[1]["total"]]
        Vs.append(Vw)
        rkl.append(fr[:, abi.F_OFFD + 1])
    assert len(Vs[0]) >= 25 and all(np.all(r > 0.5) for r in rkl)
    dgf, dgr, dG = fep.zwanzig(Vs, lams, kT, nskip)
    assert np.all(np.isfinite(dG))
    close(d["dG"], dG)
    close(d["dgf"], dgf)
    close(d["dgr"], dgr)
    H = [fep.offdiag_coupling(r, A, mu, 0.0, 1.5) for r in rkl]
    h = fep.evb_pmf(Vs, lams, H, kT, coeff, nbins, nmin, nskip, rkl=rkl)
    assert np.array_equal(d["npts"], h["npts"])
    close(d["dGg"], h["dGg"])
    close(d["dGv"], h["dGv"])
    close(d["c1"], h["c1"])
    close(d["binned"], h["binned"])
