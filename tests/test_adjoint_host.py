"""Host side of the adjoint, pinned on the CPU against the oracle.

On a small synthetic sphere (two bead chains whose first 4 of 8 atoms are
Q-atoms in 2 states, so there are Q-Q, Q-bonded-neighbour and Q-solute 1-4
pairs, TIP3P):
  * the per-frame sums of the device kernel's definition (tests/adjoint_ref.py)
    give state energies equal to the oracle's EQ(s)%qq + %qp + %qw;
  * mapped through the tangent tables and the parameter chain rule
    (q6_amd.adjoint.energy_derivatives) they equal central finite differences
    of the oracle's state energies for every parameter: Q charges per state,
    per-type sigma and epsilon, alpha;
  * the reduced gradients (Zwanzig dG, mean gap, EVB window free energy,
    binned reaction free energy) equal finite differences of q6_amd.fep's
    estimators over the same frames.
"""
import numpy as np
import pytest

from adjoint_ref import slot_sums
from oracle.oracle import Oracle
from q6_amd import abi, adjoint, synth

QMASK = abi.T_QP | abi.T_QW | abi.T_QQ | abi.T_QQP
KT = synth.BOLTZ * 300.0


def oracle_state_energies(s, frames):
    o = Oracle(s)
    E = np.zeros((len(frames), s.n["nstates"]))
    for f, x in enumerate(frames):
        o.make_pair_lists(x)
        _, _, EQ = o.eval_terms(QMASK, x)
        E[f] = [sum(q.values()) for q in EQ]
    return E


@pytest.fixture(scope="module")
def case():
    s = synth.build(11.0, 4, 8, 2, 2, seed=5, chain_region=5.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0,
                    name="adjoint-small", q_per_chain=4)
    rng = np.random.default_rng(7)
    frames = np.stack([s.x] + [s.x + rng.normal(0, 0.05, s.x.shape) for _ in range(5)])
    space = adjoint.QParamSpace(s, alpha=[0.0, 4.0], H12=20.0)
    tang = synth.q_tangents(s.md["qctx"], s.md["qcrg"], s.md["qA"], s.md["qB"])
    out = np.stack([slot_sums(s, x, tang) for x in frames])
    V, dV = adjoint.energy_derivatives(out, space)
    return s, frames, space, out, V, dV


def test_enumerated_energies_equal_oracle(case):
    s, frames, space, out, V, dV = case
    E = oracle_state_energies(s, frames)
    assert np.allclose(V - space.alpha[None, :], E, rtol=1e-11, atol=1e-9)


def test_per_frame_parameter_derivatives_match_fd(case):
    s, frames, space, out, V, dV = case
    theta0 = space.vector()
    h = 1e-5
    try:
        for p in range(space.npar - 1):  # H12 does not enter the state energies
            th = theta0.copy()
            th[p] += h
            sp = space.with_vector(th)
            sp.apply(s)
            Vp = oracle_state_energies(s, frames) + sp.alpha[None, :]
            th[p] -= 2 * h
            sm = space.with_vector(th)
            sm.apply(s)
            Vm = oracle_state_energies(s, frames) + sm.alpha[None, :]
            fd = (Vp - Vm) / (2 * h)
            scale = max(1.0, np.abs(fd).max())
            assert np.abs(fd - dV[:, :, p]).max() <= 2e-6 * scale, (space.names()[p], fd, dV[:, :, p])
    finally:
        space.apply(s)
    assert np.abs(dV[:, :, space.i_H12]).max() == 0.0


def _windows(V, nwin):
    """Split the frames into windows with lambda_1 from 1 to 0."""
    lam1 = np.linspace(1.0, 0.0, nwin)
    return np.stack([lam1, 1 - lam1], 1)


def test_reduced_gradients_match_fd_of_estimators():
    """Synthetic state-energy series with a linear parameter dependence
    V(theta) = V0 + dV . theta: every reduced gradient equals central FD."""
    rng = np.random.default_rng(3)
    nwin, npts, npar = 5, 60, 4
    lam = _windows(None, nwin)
    V0 = [np.stack([rng.normal(-10 + 4 * i, 1.5, npts), rng.normal(-2 - 3 * i, 1.5, npts)], 1) for i in range(nwin)]
    dVs = [rng.normal(0, 1, (npts, 2, npar)) for _ in range(nwin)]
    H12, i_H = 6.0, npar - 1
    for d in dVs:
        d[:, :, i_H] = 0.0
    dH = np.zeros(npar)
    dH[i_H] = 1.0

    class Space:
        pass

    sp = Space()
    sp.H12 = H12
    sp.dH = lambda: dH
    coeff = np.array([1.0, -1.0])
    Vs = [v.copy() for v in V0]
    g = adjoint.gradients(Vs, dVs, sp, lam, KT, coeff, nskip=5, nbins=8, nmin=3)
    base = adjoint.analysis(Vs, lam, KT, H12, coeff, nskip=5, nbins=8, nmin=3)
    h = 1e-6
    for p in range(npar):
        res = []
        for sgn in (1, -1):
            th = np.zeros(npar)
            th[p] = sgn * h
            Vp = [v + d @ th for v, d in zip(V0, dVs)]
            res.append(adjoint.analysis(Vp, lam, KT, H12 + th[i_H], coeff, nskip=5, nbins=8, nmin=3,
                                        frame_bins=base["frame_bins"]))
        for key in ("dG", "gap", "G_evb", "pmf"):
            fd = (np.asarray(res[0][key]) - np.asarray(res[1][key])) / (2 * h)
            an = np.asarray(g[key])[..., p] if key != "dG" else g[key][p]
            ok = np.isfinite(fd)
            assert np.allclose(np.asarray(an)[ok], fd[ok], rtol=1e-6, atol=1e-6), (key, p, an, fd)
