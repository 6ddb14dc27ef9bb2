"""The GPU path at the sizes of the benchmark workloads, against the oracle.

  * the 30 A sphere (10,861 atoms, 400 solute groups, 30 Q-atoms, 2 states)
    and the largest sphere of the uneven batch (45 A, 37,268 atoms, 3
    states): FP64 lists equal the oracle's in order and every term is within
    1e-10; FP32 pair arithmetic within its stated budget;
  * three windows (lambda_1 = 0, 0.5, 1) of the 51-window batch: the
    energies recorded at step 0 equal the oracle's with each window's lambda;
  * three mixed-precision windows after a relaxation: the Zwanzig dG over 20
    stored frames from the recorded state energies is within 0.05 kcal/mol
    of the dG from the FP64 oracle's state energies of the same frames;
  * the adjoint on a 30 A frame: the FP64 kernel equals the direct
    enumeration, the FP32 kernel is within its budget of it;
  * an uneven batch (15, 25, 29 and 45 A, 2-3 states): every instance's
    step-0 energies equal the single-system engine's;
  * neighbour rows longer than the initial capacity (16 A cut-offs) regrow
    the lists of the single engine and of a batch without dropping a pair.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from q6_amd import abi, synth

pytestmark = pytest.mark.gpu

REL = 1e-10
TERMS = [("ww", abi.T_WW), ("qw", abi.T_QW), ("pw", abi.T_PW), ("pp", abi.T_PP), ("qp", abi.T_QP),
         ("lrf", abi.T_LRF), ("qq", abi.T_QQ), ("qqp", abi.T_QQP), ("solvent_radial", abi.T_RAD),
         ("water_pol", abi.T_POL)]
LISTS = (abi.LIST_PP, abi.LIST_PW, abi.LIST_WW, abi.LIST_QP, abi.LIST_QW)
QMASK = abi.T_QP | abi.T_QW | abi.T_QQ | abi.T_QQP
KT = synth.BOLTZ * 300.0


def flat_energies(E, EQ):
    out = [E[k] for k in ("ww_el", "ww_vdw", "pw_el", "pw_vdw", "pp_el", "pp_vdw", "lrf", "solvent_radial",
                          "water_pol")]
    for q in EQ:
        out += [q["qq_el"], q["qq_vdw"], q["qp_el"], q["qp_vdw"], q["qw_el"], q["qw_vdw"]]
    return np.array(out)


@pytest.fixture(scope="module")
def cfg2():
    print("\n[gpu] building the 30 A sphere", flush=True)
    s = synth.cfg2()
    o = Oracle(s)
    o.make_pair_lists(s.x)
    return s, o


def check_frame_fp64(s, o, name):
    from q6_amd.engine import NonbondedEngine
    eng = NonbondedEngine(s)
    eng.make_pair_lists(s.x)
    for which in LISTS:
        got, ref = eng.pairs(which), o.pairs(which)
        assert got.shape == ref.shape and np.array_equal(got, ref), (name, which, got.shape, ref.shape)
    lt, lo = eng.lrf(), o.lrf()
    for a, b in ((0, 3), (3, 4), (4, 7), (7, 16), (16, 43)):
        assert np.max(np.abs(lt[:, a:b] - lo[:, a:b])) <= REL * np.max(np.abs(lo[:, a:b])), (name, a, b)
    for term, mask in TERMS:
        d, E, EQ = eng.eval_terms(mask, s.x, Tfree=300.0)
        do, Eo, EQo = o.eval_terms(mask, s.x, Tfree=300.0)
        e, eo = flat_energies(E, EQ), flat_energies(Eo, EQo)
        assert np.max(np.abs(e - eo)) <= REL * max(np.max(np.abs(eo)), 1e-12), (name, term)
        assert np.max(np.abs(d - do)) <= REL * max(np.max(np.abs(do)), 1e-12), (name, term)
    eng.close()


def test_cfg2_frame_fp64(cfg2):
    s, o = cfg2
    print(f"[gpu] {synth.describe(s)}: FP64 lists and terms vs oracle", flush=True)
    check_frame_fp64(s, o, "cfg2")


# FP32 pair arithmetic with FP64 accumulation on a frame of the 30 A sphere:
# the largest force error is bounded relative to the largest force component,
# and every class energy's error relative to the largest class energy (a sum
# of 1e5-1e6 pair terms of either sign can be near zero, so a bound relative to
# the class itself would say nothing).
MIXED_FORCE_REL, MIXED_ENERGY_REL = 5e-6, 1e-5  # observed 1.6e-6 and 1.1e-7


def test_cfg2_frame_mixed_budget(cfg2):
    """Mixed precision against the oracle within the budget above."""
    from q6_amd.engine import NonbondedEngine
    s, o = cfg2
    print("[gpu] 30 A sphere: mixed precision vs oracle", flush=True)
    eng = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
    eng.make_pair_lists(s.x)
    d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=300.0)
    do, Eo, EQo = o.eval_terms(abi.T_ALL, s.x, Tfree=300.0)
    eng.close()
    ferr = np.max(np.abs(d - do)) / np.max(np.abs(do))
    e, eo = flat_energies(E, EQ), flat_energies(Eo, EQo)
    rel = np.abs(e - eo) / np.max(np.abs(eo))
    print(f"[gpu] mixed: max force error {ferr:.2e} of the largest force; energy errors / largest class "
          f"{' '.join('%.1e' % v for v in rel)}", flush=True)
    assert ferr < MIXED_FORCE_REL
    assert np.all(rel < MIXED_ENERGY_REL)


@pytest.fixture(scope="module")
def cfg5_largest():
    print("\n[gpu] building the 45 A sphere", flush=True)
    s = synth.cfg5_member(15)
    o = Oracle(s)
    o.make_pair_lists(s.x)
    return s, o


def test_cfg5_largest_frame_fp64(cfg5_largest):
    s, o = cfg5_largest
    print(f"[gpu] {synth.describe(s)}: FP64 lists and terms vs oracle", flush=True)
    assert s.n["nstates"] == 3
    check_frame_fp64(s, o, "cfg5[15]")


def test_cfg3_windows_step0_vs_oracle(cfg2):
    """All 51 windows in one batch, FP64 and mixed: the step-0 record of
    windows 0, 25 and 50 against the oracle with the window's lambda (the
    polarisation constant cstb follows lambda)."""
    from q6_amd.md import BatchMD, scaas_cstb
    s, _ = cfg2
    nw = 51
    lam1 = np.linspace(0.0, 1.0, nw)
    lams = [np.array([l, 1.0 - l]) for l in lam1]
    vels = [synth.maxwell_velocities(s, 300.0, 100 + w) for w in range(nw)]
    recs = {}
    for prec in (abi.QNB_PREC_FP64, abi.QNB_PREC_MIXED):
        print(f"[gpu] 51 windows, precision {prec}: one step", flush=True)
        b = BatchMD([s] * nw, lams, vels, precision=prec, nbcycle=10, ene_every=10, max_frames=2)
        b.run(1, graph=False)
        b.sync()
        recs[prec] = [b.energies(w)[0] for w in range(nw)]
        b.close()
    for w in (0, 25, 50):
        o = Oracle(s, **{"lambda": lams[w], "wshell_cstb": scaas_cstb(s, lams[w])})
        o.make_pair_lists(s.x)
        f = recs[abi.QNB_PREC_FP64][w]
        _, Eo, EQo = o.eval_terms(abi.T_ALL, s.x, Tfree=f[abi.F_TFREE], istep=0)
        eo = np.array([Eo[k] for k in ("pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf",
                                       "solvent_radial", "water_pol")]
                      + sum([[q[k] for k in ("qq_el", "qq_vdw", "qp_el", "qp_vdw", "qw_el", "qw_vdw")] for q in EQo],
                            []))
        got = f[:len(eo)]
        assert np.max(np.abs(got - eo)) <= REL * np.max(np.abs(eo)), (w, got, eo)
        gm = recs[abi.QNB_PREC_MIXED][w][:len(eo)]
        merr = np.abs(gm - eo) / np.max(np.abs(eo))
        print(f"[gpu] window {w}: mixed energy errors / largest class {' '.join('%.1e' % v for v in merr)}",
              flush=True)
        assert np.all(merr < MIXED_ENERGY_REL), (w, gm, eo)


def relaxed(s, steps=200):
    from q6_amd.md import BatchMD
    m = BatchMD([s], [np.asarray(s.par["lambda"])], [np.zeros((s.natom, 3))], nbcycle=10, ene_every=10, max_frames=2)
    m.minimize(steps, 0.02)
    m.sync()
    x, _ = m.state(0)
    m.close()
    return x


@pytest.fixture(scope="module")
def mixed_frames(cfg2):
    """Three windows (lambda_1 = 0, 0.5, 1) of mixed-precision MD from a
    relaxed start, 20 frames stored every 10 steps after 100 steps."""
    from q6_amd.md import BatchMD, q_state_energies
    s, _ = cfg2
    print("[gpu] 30 A sphere: relax, then 3 mixed windows x 300 steps", flush=True)
    x0 = relaxed(s)
    lams = [np.array([l, 1.0 - l]) for l in (0.0, 0.5, 1.0)]
    vels = [synth.maxwell_velocities(s, 300.0, 200 + w) for w in range(3)]
    b = BatchMD([s] * 3, lams, vels, precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=10, max_frames=40,
                tau_T_fs=10.0, positions=[x0] * 3)
    b.run(100)
    b.store_frames(10, 20)
    b.run(200)
    b.sync()
    frames = [b.frames(w) for w in range(3)]
    V = [q_state_energies(b.energies(w)[-20:], 2) for w in range(3)]
    b.close()
    return s, lams, frames, V


def test_cfg3_dG_mixed_vs_oracle(mixed_frames):
    from q6_amd import fep
    s, lams, frames, V = mixed_frames
    Vo = []
    for w in range(3):
        o = Oracle(s)
        o.make_pair_lists(frames[w][0])
        E = np.zeros((20, 2))
        for f in range(20):
            _, _, EQ = o.eval_terms(QMASK, frames[w][f])
            E[f] = [sum(q.values()) for q in EQ]
        Vo.append(E)
    lam = np.array(lams)
    _, _, dG_m = fep.zwanzig(V, lam, KT)
    _, _, dG_o = fep.zwanzig(Vo, lam, KT)
    derr = np.max(np.abs(dG_m - dG_o))
    eerr = max(np.max(np.abs(a - b)) for a, b in zip(V, Vo))
    print(f"[gpu] dG mixed {dG_m[-1]:.4f} vs FP64 oracle {dG_o[-1]:.4f}: max |ddG| {derr:.2e} kcal/mol, "
          f"state energies within {eerr:.2e}", flush=True)
    assert derr < 0.05
    assert eerr < 0.05


def test_cfg2_adjoint_vs_enumeration(mixed_frames):
    from adjoint_ref import slot_sums
    from q6_amd import adjoint
    from q6_amd.engine import NonbondedEngine
    s, _, frames, _ = mixed_frames
    fr = frames[1][:2]
    tg, keep = adjoint.tangent_struct(s)
    e64 = NonbondedEngine(s)
    e32 = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
    o64, o32 = e64.adjoint(tg, fr), e32.adjoint(tg, fr)
    e64.close()
    e32.close()
    tang = synth.q_tangents(s.md["qctx"], s.md["qcrg"], s.md["qA"], s.md["qB"])
    ref = slot_sums(s, fr[0], tang)
    for c in range(abi.QNB_ADJ_NV):
        scale = max(np.abs(ref[..., c]).max(), 1e-30)
        assert np.abs(o64[0, ..., c] - ref[..., c]).max() <= REL * scale, c
        assert np.abs(o32[..., c] - o64[..., c]).max() <= 1e-5 * np.abs(o64[..., c]).max(), c


def test_cfg5_uneven_batch_step0():
    """Four spheres of the uneven batch (2 and 3 states, 10-60 Q-atoms) in one
    FP64 batch: each instance's step-0 record equals the single-system
    engine's evaluation of its start."""
    from q6_amd.engine import NonbondedEngine
    from q6_amd.md import BatchMD
    members = [synth.cfg5_member(k) for k in (0, 5, 7, 15)]
    print("\n[gpu] uneven batch: " + ", ".join(synth.describe(m) for m in members), flush=True)
    lams = [np.asarray(m.par["lambda"]) for m in members]
    vels = [synth.maxwell_velocities(m, 300.0, 700 + k) for k, m in enumerate(members)]
    b = BatchMD(members, lams, vels, nbcycle=10, ene_every=10, max_frames=2)
    b.run(1, graph=False)
    b.sync()
    recs = [b.energies(k)[0] for k in range(len(members))]
    b.close()
    for k, m in enumerate(members):
        eng = NonbondedEngine(m)
        eng.make_pair_lists(m.x)
        d, E, EQ = eng.eval_terms(abi.T_ALL, m.x, Tfree=recs[k][abi.F_TFREE], istep=0)
        eng.close()
        ns = int(m.n["nstates"])
        ref = np.array([E[key] for key in ("pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf",
                                           "solvent_radial", "water_pol")]
                       + sum([[q[key] for key in ("qq_el", "qq_vdw", "qp_el", "qp_vdw", "qw_el", "qw_vdw")]
                              for q in EQ], []))
        got = recs[k][:9 + 6 * ns]
        assert np.max(np.abs(got - ref)) <= REL * np.max(np.abs(ref)), (k, got, ref)


def test_list_capacity_regrowth():
    """16 A cut-offs give rows of about 600 waters, beyond the initial row
    capacity (256): the single engine regrows its lists and the batch sizes
    them at creation; no pair is dropped in either."""
    from q6_amd.engine import NonbondedEngine
    from q6_amd.md import BatchMD
    s = synth.build(20.0, 1, 12, 1, 2, 1, chain_region=2.0, q_sigma=0.4, lambdas=[0.5, 0.5], cut=16.0,
                    water_fill=20.0, name="20 A sphere, 16 A cut-offs")
    print(f"\n[gpu] {synth.describe(s)}: list regrowth", flush=True)
    o = Oracle(s)
    o.make_pair_lists(s.x)
    eng = NonbondedEngine(s)
    eng.make_pair_lists(s.x)
    for which in LISTS:
        assert np.array_equal(eng.pairs(which), o.pairs(which)), which
    eng.close()
    nww = len(o.pairs(abi.LIST_WW))
    b = BatchMD([s], [np.array([0.5, 0.5])], [synth.maxwell_velocities(s, 300.0, 3)], nbcycle=10, ene_every=10,
                max_frames=2)
    counts = b.pair_counts(0)
    b.run(1, graph=False)
    b.sync()
    rec = b.energies(0)[0]
    b.close()
    assert counts["ww"] == 9 * nww
    _, Eo, _ = o.eval_terms(abi.T_WW, s.x)
    assert abs(rec[4] - Eo["ww_el"]) <= REL * abs(Eo["ww_el"])
