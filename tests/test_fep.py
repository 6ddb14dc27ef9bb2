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
