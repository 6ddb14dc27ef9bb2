"""Batched device-resident MD on the GPU: the batch path uses the same
nonbonded kernels as the single-system engine (checked against the reference
in test_gpu_golden.py); here the batch plumbing is checked."""
import numpy as np
import pytest

from q6_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from q6_amd import synth
    s = synth.cfg1()
    v = synth.maxwell_velocities(s, 300.0, 4)
    return s, v


def test_first_frame_matches_single_engine(small):
    """Energies recorded at step 0 equal the single-system engine's at x0."""
    from q6_amd.engine import NonbondedEngine
    from q6_amd.md import BatchMD
    s, v = small
    print("\n[gpu] batch first frame", flush=True)
    b = BatchMD([s], [np.array([0.5, 0.5])], [v], nbcycle=10, ene_every=10, max_frames=4)
    b.run(10, graph=False)
    b.sync()
    f0 = b.energies(0)[0]
    b.close()
    eng = NonbondedEngine(s, **{"lambda": np.array([0.5, 0.5])})
    eng.make_pair_lists(s.x)
    d, E, EQ = eng.eval_terms(abi.T_ALL, s.x, Tfree=f0[abi.F_TFREE], istep=0)
    assert abs(f0[4] - E["ww_el"]) <= 1e-10 * abs(E["ww_el"])
    assert abs(f0[5] - E["ww_vdw"]) <= 1e-10 * abs(E["ww_vdw"])
    assert abs(f0[6] - E["lrf"]) <= 1e-10 * max(abs(E["lrf"]), 1.0)
    for st in range(2):
        qw = f0[abi.F_Q0 + 6 * st + 4]
        assert abs(qw - EQ[st]["qw_el"]) <= 1e-10 * max(abs(EQ[st]["qw_el"]), 1.0)
    eng.close()
    # harmonic bonds (with the switch-atom restraints) and angles at x0
    # (bond, angle: src/bondene.f90:41-246), each term counted once
    md, x = s.md, s.x
    ij = np.concatenate([np.asarray(md["bonds"]).reshape(-1, 2), np.asarray(md["restraints"]).reshape(-1, 2)])
    r0 = np.concatenate([np.broadcast_to(md["bond_r0"], len(md["bonds"])),
                         np.broadcast_to(md["restraint_r0"], len(md["restraints"]))])
    kb = np.concatenate([np.broadcast_to(md["bond_k"], len(md["bonds"])),
                         np.broadcast_to(md["restraint_k"], len(md["restraints"]))])
    r = np.linalg.norm(x[ij[:, 1]] - x[ij[:, 0]], axis=1)
    eb = float(np.sum(0.5 * kb * (r - r0) ** 2))
    ijk = np.asarray(md["angles"]).reshape(-1, 3)
    u, w = x[ijk[:, 0]] - x[ijk[:, 1]], x[ijk[:, 2]] - x[ijk[:, 1]]
    ct = np.clip(np.sum(u * w, 1) / (np.linalg.norm(u, axis=1) * np.linalg.norm(w, axis=1)), -1.0, 1.0)
    ea = float(np.sum(0.5 * md["angle_k"] * (np.arccos(ct) - md["angle_t0"]) ** 2))
    assert abs(f0[abi.F_BOND] - eb) <= 1e-10 * max(abs(eb), 1.0), (f0[abi.F_BOND], eb)
    assert abs(f0[abi.F_ANGLE] - ea) <= 1e-10 * max(abs(ea), 1.0), (f0[abi.F_ANGLE], ea)


def test_windows_are_independent_and_graph_matches_eager(small):
    """Two identical instances in one batch follow the same trajectory, and
    graph replay gives the trajectory eager launches give."""
    from q6_amd.md import BatchMD
    s, v = small
    lam = [np.array([0.5, 0.5])] * 2
    print("\n[gpu] batch independence", flush=True)
    b = BatchMD([s, s], lam, [v, v], nbcycle=10, ene_every=10, max_frames=8)
    b.run(40, graph=True)
    b.sync()
    x0, _ = b.state(0)
    x1, _ = b.state(1)
    assert np.max(np.abs(x0 - x1)) < 1e-9
    b.close()
    c = BatchMD([s], lam[:1], [v], nbcycle=10, ene_every=10, max_frames=8)
    c.run(40, graph=False)
    c.sync()
    xe, _ = c.state(0)
    assert np.max(np.abs(xe - x0)) < 1e-9
    c.close()


def test_md_stays_stable_and_rigid(small):
    """The seeded lattice start relaxes (random water orientations release
    potential energy); with a 10 fs Berendsen coupling the 300 K bath absorbs it
    within a few hundred steps.  Waters stay rigid (SHAKE, relative tolerance
    1e-4 on d^2) and every sampled energy stays finite."""
    from q6_amd.md import BatchMD
    s, v = small
    print("\n[gpu] batch stability", flush=True)
    b = BatchMD([s], [np.array([0.5, 0.5])], [v], precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=10,
                max_frames=60, tau_T_fs=10.0)
    b.run(500)
    b.sync()
    x, _ = b.state(0)
    e = b.energies(0)
    b.close()
    o = s.n["nat_solute"] + 3 * np.arange(s.n["nwat"])
    roh = np.linalg.norm(x[o + 1] - x[o], axis=1)
    assert np.all(np.abs(roh - 0.9572) < 2e-3)
    T = e[:, abi.F_TEMP]
    print(f"[gpu] T first {T[0]:.1f} max {T.max():.1f} last {T[-1]:.1f}", flush=True)
    assert np.all(np.isfinite(e[:, :abi.F_STEP]))
    assert 200.0 < T[-1] < 450.0


def test_offdiag_sampled_with_energies(small):
    """EVB off-diagonal elements recorded with the energies: r_kl is the
    distance of the two Q-atoms in the coordinates of the sampled step (the
    stored frames of the same steps), H = A exp(-mu r_kl); an instance
    without off-diagonals records zeros there."""
    from q6_amd.md import BatchMD
    s, v = small
    iqseq = s.a["iqseq"]
    pairs = [(0, 1), (2, len(iqseq) - 1)]
    A, mu = np.array([30.0, 12.5]), np.array([0.45, 1.2])
    print("\n[gpu] offdiag sampling", flush=True)
    b = BatchMD([s, s], [np.array([0.5, 0.5])] * 2, [v, v], precision=abi.QNB_PREC_MIXED, nbcycle=10,
                ene_every=10, max_frames=16)
    b.set_offdiag(0, pairs, A, mu)
    b.store_frames(10, 16)
    b.run(60)
    b.sync()
    e0, e1, fr = b.energies(0), b.energies(1), b.frames(0)
    with pytest.raises(abi.QnbError):
        b.set_offdiag(0, [(0, len(iqseq))], [1.0], [0.0])
    b.close()
    assert len(e0) == len(fr) == 6  # steps 0, 10, ..., 50
    for m, (k, l) in enumerate(pairs):
        r = np.linalg.norm(fr[:, iqseq[l]] - fr[:, iqseq[k]], axis=1)
        H = A[m] * np.exp(-mu[m] * r)
        assert np.allclose(e0[:, abi.F_OFFD + 2 * m + 1], r, rtol=1e-13, atol=0)
        assert np.allclose(e0[:, abi.F_OFFD + 2 * m], H, rtol=1e-13, atol=0)
    assert np.all(e0[:, abi.F_OFFD + 4:] == 0.0)
    assert np.all(e1[:, abi.F_OFFD:] == 0.0)


def golden_md_system(case):
    """A golden system with the MD data the batch needs: masses from the
    topology types; no bonds or angles of its own (the step under test takes
    the full gradient pot_energy left, d_total)."""
    from conftest import golden_path
    from q6_amd.system import from_golden
    s = from_golden(golden_path(case))
    s.md = dict(mass=s.ref["type_mass"][s.a["iac"]], bonds=np.zeros((0, 2), np.int32), bond_r0=1.0, bond_k=0.0,
                angles=np.zeros((0, 3), np.int32), angle_t0=0.0, angle_k=0.0,
                water_roh=float(np.sqrt(s.ref["const_dist2"][0])), water_rhh=float(np.sqrt(s.ref["const_dist2"][2])))
    return s


@pytest.mark.parametrize("case", ["lig_eq1", "lig_fep2"])
def test_one_md_step_matches_qdyn6(case):
    """One step of md_run after pot_energy (src/md.f90:1774-1797) on the
    device from Qdyn6's own start and gradient: leap-frog with the solute
    and solvent Berendsen factors, SHAKE of the waters, v = (x - x_old)/dt
    and temperature().  Positions, velocities, Temp, Tfree, the kinetic
    energy and each group's tfree and next factor within 1e-10 of stock
    Qdyn6's step (the initial factors too)."""
    from q6_amd.md import BatchMD
    print(f"\n[gpu] one MD step vs Qdyn6 {case}", flush=True)
    s = golden_md_system(case)
    assert int(s.ref["ntgroups"]) == 2
    dt_fs = float(s.ref["dt"]) / 0.020462
    b = BatchMD([s], [np.asarray(s.par["lambda"])], [s.ref["v"].T], dt_fs=dt_fs, T0=float(s.ref["Temp0"]),
                tau_T_fs=float(s.ref["tau_T"]) / 0.020462, nbcycle=10, ene_every=0, max_frames=2)
    th0 = b.thermostat(0)
    assert th0["ngroups"] == 2
    assert np.max(np.abs(th0["sfact"] - s.ref["md1_sfact_in"])) <= 1e-12
    b.integrate([s.ref["d_total"].T])
    b.sync()
    x, v = b.state(0)
    th = b.thermostat(0)
    b.close()
    xr, vr = s.ref["md1_x"].T, s.ref["md1_v"].T
    assert np.max(np.abs(x - xr)) <= 1e-10 * np.max(np.abs(xr))
    assert np.max(np.abs(v - vr)) <= 1e-10 * np.max(np.abs(vr)), np.max(np.abs(v - vr))
    t = s.ref["md1_temp"]
    got = [th["Temp"], th["Tfree"], th["Ekin"], th["tfree"][0], th["sfact"][0], th["tfree"][1], th["sfact"][1]]
    ref = [t[0], t[1], t[2], t[4], t[5], t[7], t[8]]
    assert np.allclose(got, ref, rtol=1e-10, atol=0), (got, ref)


def test_group_contributions_recorded():
    """Group-contribution arrays (calculate_exc) recorded on every sampled
    step of a batch: for each frame, array and state they equal the oracle's
    calculate_exc at the stored coordinates of that step, from the same
    recorded per-state energies (FP64 on both sides, within 1e-12)."""
    from oracle.oracle import Oracle
    from q6_amd import synth
    from q6_amd.md import BatchMD
    s = synth.cfg2()
    nsol, ns = s.n["nat_solute"], s.n["nstates"]
    iqseq = s.a["iqseq"]
    masks = np.zeros((3, nsol), np.uint8)
    masks[0, :10] = 1                           # a chain of Q-atoms
    masks[1, [5] + list(range(100, 110))] = 1   # one Q-atom, a plain chain
    masks[2, [iqseq[-1]] + list(range(200, 210))] = 1
    types = np.array([abi.GC_FULL, abi.GC_ELECTRO, abi.GC_VDW], np.int32)
    lams = [np.full(ns, 1.0 / ns), np.eye(ns)[0]]
    v = synth.maxwell_velocities(s, 300.0, 9)
    print("\n[gpu] batch group contributions", flush=True)
    b = BatchMD([s, s], lams, [v, v], precision=abi.QNB_PREC_MIXED, nbcycle=10, ene_every=5, max_frames=12)
    b.minimize(200, 0.02)  # the seeded start has close contacts
    b.set_groups(1, masks, types)
    b.store_frames(5, 12)
    b.run(30)
    b.sync()
    ge, en, fr = b.group_energies(1), b.energies(1), b.frames(1)
    assert b.group_energies(0).shape[1] == 0
    b.close()
    assert len(ge) == len(en) == len(fr) == 6
    o = Oracle(s)
    for f in range(len(fr)):
        row = en[f]
        eq_in = np.array([[*row[abi.F_Q0 + 6 * st:abi.F_Q0 + 6 * st + 6], *row[abi.F_QB0 + 4 * st:abi.F_QB0 + 4 * st + 4],
                           0.0] for st in range(ns)])
        ref = o.calculate_exc(fr[f], masks, types, eq_in)
        assert np.all(np.abs(ge[f] - ref) <= 1e-12 * np.maximum(np.abs(ref), 1.0)), (f, np.max(np.abs(ge[f] - ref)))
    assert np.any(ge[:, 0, :, 0] != en[:, [abi.F_Q0 + 6 * st for st in range(ns)]])  # Q-Q energy removed
