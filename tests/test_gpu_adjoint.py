"""Adjoint kernel on the GPU.

  * FP64 kernel: per-frame sums equal the direct enumeration of their
    definition (tests/adjoint_ref.py, itself pinned by finite differences of
    the oracle on the CPU) and its state energies equal the plain FP64
    engine's EQ(s)%qq + %qp + %qw, to 1e-10;
  * FP32-pair kernel (the production backward): within the FP32 budget of
    the FP64 one;
  * batched adjoint over frames stored during mixed-precision MD of several
    lambda windows: gradients of the Zwanzig dG, the mean energy gap of every
    window, the EVB ground-state free energy of every window and the binned
    reaction free energy with respect to Q charges, per-type sigma/epsilon,
    alpha and H12 agree with central finite differences (step 1e-4) of the
    same estimators evaluated on the plain FP64 engine over the same frames;
  * an uneven batch (two systems) returns per instance what the
    single-system entry point returns on its frames.
"""
import numpy as np
import pytest

from adjoint_ref import slot_sums
from q6_amd import abi, adjoint, synth

pytestmark = pytest.mark.gpu

QMASK = abi.T_QP | abi.T_QW | abi.T_QQ | abi.T_QQP
KT = synth.BOLTZ * 300.0


def small_system():
    return synth.build(11.0, 4, 8, 2, 2, seed=5, chain_region=5.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0,
                       name="adjoint-small", q_per_chain=4)


def engine_state_energies(s, frames):
    from q6_amd.engine import NonbondedEngine
    eng = NonbondedEngine(s)
    E = np.zeros((len(frames), s.n["nstates"]))
    for f, x in enumerate(frames):
        eng.make_pair_lists(x)
        _, _, EQ = eng.eval_terms(QMASK, x)
        E[f] = [sum(q.values()) for q in EQ]
    eng.close()
    return E


@pytest.fixture(scope="module")
def frames_case():
    s = small_system()
    rng = np.random.default_rng(7)
    frames = np.stack([s.x] + [s.x + rng.normal(0, 0.05, s.x.shape) for _ in range(5)])
    return s, frames


def test_fp64_adjoint_matches_enumeration_and_engine(frames_case):
    from q6_amd.engine import NonbondedEngine
    s, frames = frames_case
    print("\n[gpu] adjoint fp64 vs enumeration", flush=True)
    tg, keep = adjoint.tangent_struct(s)
    eng = NonbondedEngine(s)
    out = eng.adjoint(tg, frames)
    eng.close()
    tang = synth.q_tangents(s.md["qctx"], s.md["qcrg"], s.md["qA"], s.md["qB"])
    ref = np.stack([slot_sums(s, x, tang) for x in frames])
    for c in range(abi.QNB_ADJ_NV):
        scale = max(np.abs(ref[..., c]).max(), 1e-30)
        assert np.abs(out[..., c] - ref[..., c]).max() <= 1e-10 * scale, c
    E = engine_state_energies(s, frames)
    Ek = (out[..., abi.ADJ_EEL] + out[..., abi.ADJ_EVDW]).sum(2)
    assert np.allclose(Ek, E, rtol=1e-10, atol=1e-9)


def test_mixed_adjoint_within_fp32_budget(frames_case):
    from q6_amd.engine import NonbondedEngine
    s, frames = frames_case
    print("\n[gpu] adjoint mixed vs fp64", flush=True)
    tg, keep = adjoint.tangent_struct(s)
    e64 = NonbondedEngine(s)
    e32 = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
    o64 = e64.adjoint(tg, frames)
    o32 = e32.adjoint(tg, frames)
    e64.close()
    e32.close()
    for c in range(abi.QNB_ADJ_NV):
        scale = np.abs(o64[..., c]).max()
        assert np.abs(o32[..., c] - o64[..., c]).max() <= 1e-5 * scale, c


@pytest.fixture(scope="module")
def md_case():
    """5 lambda windows of the small system, mixed precision, frames stored
    every 10 steps after a short relaxation."""
    from q6_amd.md import BatchMD
    s = small_system()
    nwin = 5
    lam1 = np.linspace(1.0, 0.0, nwin)
    lams = np.stack([lam1, 1 - lam1], 1)
    vs = [synth.maxwell_velocities(s, 300.0, 100 + w) for w in range(nwin)]
    print("\n[gpu] adjoint md: relax + 300 steps x 5 windows", flush=True)
    b = BatchMD([s] * nwin, list(lams), vs, precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=10,
                max_frames=64, tau_T_fs=10.0)
    b.minimize(100)
    b.run(100)
    b.store_frames(10, 40)
    b.run(300)
    b.sync()
    tg, keep = adjoint.tangent_struct(s)
    outs = b.adjoint([tg] * nwin)
    frames = [b.frames(k) for k in range(nwin)]
    b.close()
    return s, lams, frames, outs


def test_batched_adjoint_gradients_match_fd_on_fp64_path(md_case):
    s, lams, frames, outs = md_case
    nskip, nbins, nmin = 5, 6, 3
    coeff = np.array([1.0, -1.0])
    space = adjoint.QParamSpace(s, alpha=[0.0, 2.0], H12=20.0)
    assert all(len(f) == 30 for f in frames)
    VdV = [adjoint.energy_derivatives(o, space) for o in outs]
    Vs = [v for v, _ in VdV]
    dVs = [d for _, d in VdV]
    g = adjoint.gradients(Vs, dVs, space, lams, KT, coeff, nskip, nbins, nmin)
    base = adjoint.analysis(Vs, lams, KT, space.H12, coeff, nskip, nbins, nmin)
    theta0 = space.vector()
    h = 1e-4
    print("\n[gpu] adjoint fd over", space.npar, "parameters", flush=True)
    worst, fds, ans = {}, {}, {}
    try:
        for p in range(space.npar):
            res = []
            for sgn in (1.0, -1.0):
                th = theta0.copy()
                th[p] += sgn * h
                sp = space.with_vector(th)
                sp.apply(s)
                Vp = [engine_state_energies(s, fr) + sp.alpha[None, :] for fr in frames]
                res.append(adjoint.analysis(Vp, lams, KT, sp.H12, coeff, nskip, nbins, nmin,
                                            frame_bins=base["frame_bins"]))
            for key in ("dG", "gap", "G_evb", "pmf"):
                fd = (np.asarray(res[0][key]) - np.asarray(res[1][key])) / (2 * h)
                an = g[key][p] if key == "dG" else np.asarray(g[key])[..., p]
                fds.setdefault(key, []).append(np.atleast_1d(fd))
                ans.setdefault(key, []).append(np.atleast_1d(np.asarray(an, float)))
    finally:
        space.apply(s)
    # 1e-4 relative: every derivative against its finite difference, with a
    # floor at 1e-2 of the largest derivative of the same estimator (where the
    # finite difference itself carries noise of that order), and the whole
    # gradient of each estimator in norm
    for key in fds:
        fd, an = np.stack(fds[key], -1), np.stack(ans[key], -1)
        ok = np.isfinite(fd)
        fdk, ank = fd[ok], an[ok]
        big = np.max(np.abs(fdk))
        tol = 1e-4 * np.maximum(np.abs(fdk), 1e-2 * big)
        err = np.abs(ank - fdk)
        worst[key] = float((err / tol).max())
        assert np.all(err <= tol), (key, np.argmax(err / tol), ank, fdk)
        assert np.linalg.norm(ank - fdk) <= 1e-4 * np.linalg.norm(fdk), key
    print("[gpu] worst |adjoint - fd| / tol per estimator:", worst, flush=True)


def test_uneven_batch_matches_single_system_entry():
    from q6_amd.engine import NonbondedEngine
    from q6_amd.md import BatchMD
    s1 = small_system()
    s2 = synth.build(12.0, 5, 8, 1, 3, seed=9, chain_region=5.0, q_sigma=0.3, lambdas=[0.2, 0.3, 0.5], cut=7.0,
                     name="adjoint-3state", q_per_chain=6)
    sys = [s1, s2, s1]
    lams = [np.array([0.5, 0.5]), np.array([0.2, 0.3, 0.5]), np.array([1.0, 0.0])]
    vs = [synth.maxwell_velocities(s, 300.0, 40 + k) for k, s in enumerate(sys)]
    print("\n[gpu] adjoint uneven batch", flush=True)
    b = BatchMD(sys, lams, vs, precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=10, max_frames=16, tau_T_fs=10.0)
    b.minimize(50)
    b.store_frames(10, 4)
    b.run(60)
    b.sync()
    t1, k1 = adjoint.tangent_struct(s1)
    t2, k2 = adjoint.tangent_struct(s2)
    outs = b.adjoint([t1, t2, t1])
    frames = [b.frames(k) for k in range(3)]
    b.close()
    assert [len(f) for f in frames] == [4, 4, 4]
    for k, (s, tg) in enumerate([(s1, t1), (s2, t2), (s1, t1)]):
        eng = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
        single = eng.adjoint(tg, frames[k])
        eng.close()
        assert outs[k].shape == single.shape
        # same kernel; the batch instantiates it for the largest state count
        for c in range(abi.QNB_ADJ_NV):
            scale = max(np.abs(single[..., c]).max(), 1e-30)
            assert np.abs(outs[k][..., c] - single[..., c]).max() <= 1e-6 * scale, (k, c)
