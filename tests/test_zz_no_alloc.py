"""The entry point that allocates nothing (qnb_adjoint_static) against the
allocating one it mirrors (qnb_adjoint): same kernel, every buffer beyond
the caller's inputs and outputs in a fixed per-GPU workspace.  Two calls in a
row on one handle, the second with a smaller input, each equal to
qnb_adjoint's output; a 30 A system whose frames go through the workspace in
more than one chunk; tangent tables larger than the workspace are refused
with a clear error and the handle stays usable."""
import numpy as np
import pytest

from q6_amd import abi, adjoint, synth

pytestmark = pytest.mark.gpu


def same(got, ref):
    for c in range(abi.QNB_ADJ_NV):
        scale = max(np.abs(ref[..., c]).max(), 1e-300)
        assert np.abs(got[..., c] - ref[..., c]).max() <= 1e-12 * scale, c


def frames_of(s, n, seed, sigma=0.05):
    rng = np.random.default_rng(seed)
    return np.stack([s.x + rng.normal(0, sigma, s.x.shape) for _ in range(n)])


def test_static_adjoint_two_calls_second_smaller():
    from q6_amd.engine import NonbondedEngine
    s = synth.build(11.0, 4, 8, 2, 2, seed=5, chain_region=5.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0,
                    name="adjoint-small", q_per_chain=4)
    print("\n[gpu] qnb_adjoint_static: small system, 6 then 3 frames", flush=True)
    tg, keep = adjoint.tangent_struct(s)
    for prec in (abi.QNB_PREC_FP64, abi.QNB_PREC_MIXED):
        eng = NonbondedEngine(s, precision=prec)
        for n, seed in ((6, 1), (3, 2)):
            fr = frames_of(s, n, seed)
            same(eng.adjoint_static(tg, fr), eng.adjoint(tg, fr))
        eng.close()


def test_static_adjoint_chunks_a_30A_system():
    from q6_amd.engine import NonbondedEngine
    s = synth.cfg2()
    print("\n[gpu] qnb_adjoint_static: 30 A sphere, 130 then 40 frames", flush=True)
    tg, keep = adjoint.tangent_struct(s)
    eng = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
    for n, seed in ((130, 3), (40, 4)):
        fr = frames_of(s, n, seed, 0.02)
        same(eng.adjoint_static(tg, fr), eng.adjoint(tg, fr))
    eng.close()


def test_static_adjoint_refuses_oversize_tables():
    from q6_amd.engine import NonbondedEngine
    s = synth.build(30.0, 400, 10, 12, 3, seed=11, chain_region=18.0, lambdas=[0.4, 0.3, 0.3], compact=3.0,
                    water_fill=30.0, name="120 Q-atoms x 3 states", **synth.PACK)
    print(f"\n[gpu] qnb_adjoint_static: {synth.describe(s)} (tables beyond the workspace)", flush=True)
    tg, keep = adjoint.tangent_struct(s)
    eng = NonbondedEngine(s, precision=abi.QNB_PREC_MIXED)
    fr = frames_of(s, 2, 5, 0.02)
    with pytest.raises(abi.QnbError, match="static workspace"):
        eng.adjoint_static(tg, fr)
    small = synth.build(11.0, 4, 8, 2, 2, seed=5, chain_region=5.0, q_sigma=0.4, lambdas=[0.3, 0.7], cut=7.0,
                        q_per_chain=4)
    eng2 = NonbondedEngine(small)
    t2, k2 = adjoint.tangent_struct(small)
    fr2 = frames_of(small, 2, 6)
    same(eng2.adjoint_static(t2, fr2), eng2.adjoint(t2, fr2))
    eng2.close()
    eng.close()
