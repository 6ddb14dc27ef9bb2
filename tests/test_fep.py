"""Zwanzig/EVB reductions (q6_amd.fep) against closed forms and brute force."""
import numpy as np

from q6_amd import fep


def test_constant_gap_gives_exact_free_energy():
    # V2 - V1 = c on every frame: the free energy of 1 -> 2 is exactly c
    rng = np.random.default_rng(0)
    lam1 = np.linspace(0, 1, 11)
    lams = np.stack([1 - lam1, lam1], 1)
    Vs = []
    for w in range(11):
        v1 = rng.normal(0, 5, 200)
        Vs.append(np.stack([v1, v1 + 3.7], 1))
    dgf, dgr, dG = fep.zwanzig(Vs, lams, 0.596)
    assert np.allclose(dgf, 0.37, atol=1e-12)
    assert np.allclose(dgr, -0.37, atol=1e-12)
    assert abs(dG[-1] - 3.7) < 1e-12


def test_gaussian_gap_matches_second_cumulant():
    # Gaussian dV with variance s2: -kT ln<exp(-dV/kT)> = <dV> - s2/(2kT)
    rng = np.random.default_rng(1)
    kT, mu, sd = 0.6, 1.2, 0.4
    x = rng.normal(mu, sd, 400000)
    V = np.stack([np.zeros_like(x), x], 1)
    dgf, _, _ = fep.zwanzig([V, V], np.array([[1.0, 0.0], [0.0, 1.0]]), kT)
    assert abs(dgf[0] - (mu - sd * sd / (2 * kT))) < 5e-3


def test_evb_ground_state_closed_form():
    V = np.array([[0.0, 0.0], [10.0, -10.0], [3.0, 5.0]])
    eg, c1 = fep.evb_ground(V, 80.0)
    assert abs(eg[0] + 80.0) < 1e-12          # degenerate states: -H12
    # eigenvalue of [[V1, H],[H, V2]]
    for k in range(3):
        ev = np.linalg.eigvalsh(np.array([[V[k, 0], 80.0], [80.0, V[k, 1]]]))
        assert abs(eg[k] - ev[0]) < 1e-9
    assert abs(c1[0] - 0.5) < 1e-12


def test_zwanzig_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    nwin, npts = 6, 300
    lam1 = np.linspace(0, 1, nwin)
    lams = np.stack([1 - lam1, lam1], 1)
    base = [rng.normal(0, 1, (npts, 2)) for _ in range(nwin)]
    # state energies linear in two parameters p: V = base + p0*A + p1*B
    A = [rng.normal(0, 1, (npts, 2)) for _ in range(nwin)]
    B = [rng.normal(0, 1, (npts, 2)) for _ in range(nwin)]

    def dG(p):
        Vs = [base[w] + p[0] * A[w] + p[1] * B[w] for w in range(nwin)]
        return fep.zwanzig(Vs, lams, 0.6)[2][-1]
    p0 = np.array([0.3, -0.2])
    Vs = [base[w] + p0[0] * A[w] + p0[1] * B[w] for w in range(nwin)]
    dVs = [np.stack([A[w], B[w]], 2) for w in range(nwin)]
    g = fep.zwanzig_grad(Vs, dVs, lams, 0.6)
    h = 1e-5
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        fd = (dG(p0 + e) - dG(p0 - e)) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1, abs(fd))


def test_evb_ground_state_free_energy_of_uncoupled_states_is_zero():
    """H12 = 0 and V2 = V1 + c with c > 0: E_g = V1 on every frame, the
    mapping free energy is dG_i = lambda2_i c exactly, and the ground-state
    free energy dG_i - kT ln<exp(-(E_g - eps_i)/kT)> vanishes at every window."""
    rng = np.random.default_rng(5)
    nwin, c, kT = 7, 3.5, 0.596
    lam1 = np.linspace(1, 0, nwin)
    lams = np.stack([lam1, 1 - lam1], 1)
    Vs = []
    for _ in range(nwin):
        v1 = rng.normal(-40.0, 2.0, 50)
        Vs.append(np.stack([v1, v1 + c], 1))
    _, _, dG = fep.zwanzig(Vs, lams, kT)
    assert np.allclose(dG, lams[:, 1] * c, atol=1e-12)
    G = fep.evb_window_free_energy(Vs, lams, 0.0, kT)
    assert np.abs(G).max() < 1e-10


def test_pmf_single_bin_is_point_weighted_window_average():
    """With one bin every frame but the largest-gap one (which falls in the
    reference's unreported extra bin) is binned together, so Part 3's value
    is the point-weighted mean over windows of the single-bin Part 2 value."""
    rng = np.random.default_rng(6)
    nwin, kT, H12 = 4, 0.596, 7.0
    lam1 = np.linspace(1, 0, nwin)
    lams = np.stack([lam1, 1 - lam1], 1)
    Vs = [np.stack([rng.normal(-5 * i, 2, 40), rng.normal(5 * i - 20, 2, 40)], 1) for i in range(nwin)]
    coeff = np.array([1.0, -1.0])
    res = fep.evb_pmf(Vs, lams, H12, kT, coeff, nbins=1, nmin=1)
    gaps = [V @ coeff for V in Vs]
    top = max(range(nwin), key=lambda i: gaps[i].max())
    _, _, dG = fep.zwanzig(Vs, lams, kT)
    num = den = 0.0
    for i, V in enumerate(Vs):
        keep = np.ones(len(V), bool)
        if i == top:
            keep[np.argmax(gaps[i])] = False
        eg, _ = fep.evb_ground(V[keep], H12)
        dvg = eg - V[keep] @ lams[i]
        g = dG[i] - kT * np.log(np.mean(np.exp(-dvg / kT)))
        num += g * keep.sum()
        den += keep.sum()
    assert res["ptsum"][0] == den
    assert abs(res["binned"][0] - num / den) < 1e-9
