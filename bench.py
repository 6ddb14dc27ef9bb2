"""Benchmark: batched mixed-precision lambda windows vs the plain FP64 path.

Workload (the metric's headline case): 51 lambda windows (lambda_1 = 0, 0.02,
..., 1; lambda_2 = 1 - lambda_1) of the 30 A sphere system (400 ten-atom
solute charge groups, 30 Q-atoms in 2 EVB states, TIP3P), each window with its
own seeded Maxwell-Boltzmann velocities at 300 K, 1 fs steps, pair-list / LRF
update every 10 steps, all windows advanced in the same kernel launches with
FP32 pair arithmetic and FP64 accumulation, one HIP graph per list cycle.
Reported next to it: the plain FP64 GPU path running one window at a time
(51 windows run sequentially give that same aggregate ns/day), and the C
oracle (single thread) on a bounded sample of the same system.

A "step" is one MD step of every window: nonbonded engine (forces, per-state
Q energies, LRF Taylor term, SCAAS restraints; lists + LRF expansion every
10th step), solute bonds/angles, leap-frog, SHAKE of the waters, Berendsen.
ns/day = steps/s x dt(fs) x 0.0864, summed over windows.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X peaks (vendor figures): FP64 vector 78.6 TFLOP/s, FP32 vector 157.3 TFLOP/s, HBM 8 TB/s
PEAK_FP64, PEAK_FP32, PEAK_HBM_GBPS = 78.6, 157.3, 8000.0
# HBM bytes per launch of the dominant kernels, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (scaled by a copy kernel of known size) on this
# workload; the file gives the commands
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_51windows_10861atoms_mixed.md")
# algorithmic FLOPs per unit (div and sqrt count as one)
F_LRF = 89        # LRF atom-centre update as restated in the engine (20 independent components)
F_TEST = 8        # group-pair distance test of the list sweep
F_LJC, F_COUL = 33, 22
F_Q_SHARED, F_Q_STATE, F_QH_STATE = 12, 27, 9


def lrf_units(s, x):
    """Atom-centre updates and pair tests of one list update (host count from
    positions: every receiver group gets every non-Q source atom of the groups
    beyond its class cut-off)."""
    n = s.n
    ncs, ng = n["ncgp_solute"], n["ncgp"]
    ref = np.concatenate([s.a["cgp_iswitch"][:ncs], n["nat_solute"] + 3 * np.arange(n["nwat"])]).astype(int)
    p = x[ref]
    nonq = np.array([sum(1 for k in range(s.a["cgp_first"][g], s.a["cgp_last"][g] + 1)
                         if s.a["iqatom"][s.a["cgpatom"][k]] == 0) for g in range(ng)])
    sol = np.arange(ng) < ncs
    units = 0
    for a in range(ng):
        r2 = ((p - p[a]) ** 2).sum(1)
        cut = np.where(sol & sol[a], s.par["Rcpp"] ** 2, np.where(~sol & ~sol[a], s.par["Rcww"] ** 2, s.par["Rcpw"] ** 2))
        far = (r2 > cut) & (r2 <= s.par["RcLRF"] ** 2)
        far[a] = False
        units += int(nonq[far].sum())
    return units, ng * ng


def measured_traffic(nw, natom):
    """Bytes per launch from the counter summary, when it is for this workload."""
    try:
        for ln in open(TRAFFIC_FILE):
            if ln.strip().startswith("traffic_bytes_per_launch "):
                t = json.loads(ln.strip().split(" ", 1)[1])
                if t.get("workload") == f"{nw} lambda windows, {natom} atoms, mixed":
                    return t
    except OSError:
        pass
    return None


def pair_flops(counts):
    """Algorithmic FLOPs of the non-Q pairs (pp, pw, ww), each pair once: the
    work of the receiver-side pass k_wforces + k_sforces (which evaluates
    every pair from both ends)."""
    return (counts["ww"] // 9) * (F_LJC + 8 * F_COUL) + (counts["pp"] + counts["pw"]) * F_LJC


def q_flops(counts, ns):
    """Algorithmic FLOPs of the Q-atom pairs (qp, qw: per-state energies and
    forces; qq, qqp), each pair once."""
    return (counts["qp"] * (F_Q_SHARED + F_Q_STATE * ns)
            + (counts["qw"] // 3) * ((F_Q_SHARED + F_Q_STATE * ns) + 2 * (F_Q_SHARED + F_QH_STATE * ns))
            + (counts["qq"] + counts["qqp"]) * (F_Q_SHARED + F_Q_STATE))


def equilibrate(s, steps=300, min_steps=300):
    """Relax the seeded start (untimed set-up): steepest descent, then one FP64
    instance at 300 K with a 10 fs bath coupling; its final coordinates start
    every window."""
    from q6_amd import synth
    from q6_amd.md import BatchMD
    m = BatchMD([s], [np.asarray(s.par["lambda"])], [np.zeros((s.natom, 3))], nbcycle=10, ene_every=10, max_frames=2)
    m.minimize(min_steps, 0.02)
    m.sync()
    x, _ = m.state(0)
    m.close()
    b = BatchMD([s], [np.asarray(s.par["lambda"])], [synth.maxwell_velocities(s, 300.0, 1)], nbcycle=10,
                ene_every=10, max_frames=steps // 10 + 2, tau_T_fs=10.0, positions=[x])
    b.run(steps)
    b.sync()
    x, _ = b.state(0)
    b.close()
    return x


def run_md(systems, lambdas, vels, precision, steps, warmup, graph=True, positions=None, tune=True, extra=0):
    from q6_amd.md import BatchMD
    b = BatchMD(systems, lambdas, vels, precision=precision, nbcycle=10, ene_every=10,
                max_frames=(warmup + steps + extra) // 10 + 2, positions=positions)
    if tune:
        b.shapes = b.autotune(reps=10)
    if graph:
        b.prepare_graph()
    b.run(warmup, graph=graph)
    b.sync()
    t0 = time.perf_counter()
    b.run(steps, graph=graph)
    b.sync()
    dt = time.perf_counter() - t0
    return b, dt


def oracle_state_energies(s, frames_by_window, threads):
    """FP64 oracle state energies EQ(s)%qq + %qp + %qw of every stored frame
    (the Q lists are built once, as Qdyn6 does when Rcq exceeds the sphere),
    one oracle per thread over the windows."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import Oracle
    from q6_amd import abi
    qmask = abi.T_QP | abi.T_QW | abi.T_QQ | abi.T_QQP
    ns = int(s.n["nstates"])

    def one(frames):
        o = Oracle(s)
        o.make_pair_lists(frames[0])
        E = np.zeros((len(frames), ns))
        for f, x in enumerate(frames):
            _, _, EQ = o.eval_terms(qmask, x)
            E[f] = [sum(q.values()) for q in EQ]
        return E

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, frames_by_window))


def free_energy_precision(bm, s, lambdas, nfr, threads):
    """Runs nfr*10 more steps storing a frame and an energy record every 10;
    the per-window state energies the mixed run recorded and the FP64
    oracle's state energies of the same frames give two Zwanzig dG curves
    over the same frames: their difference is the precision error of dG.
    Then times the adjoint over the stored frames (one launch, all windows)."""
    from q6_amd import adjoint, fep
    from q6_amd.md import q_state_energies
    nw = len(lambdas)
    bm.store_frames(10, nfr)
    bm.run(10 * nfr)
    bm.sync()
    frames = [bm.frames(k) for k in range(nw)]
    V_mixed = [q_state_energies(bm.energies(k)[-nfr:], 2) for k in range(nw)]
    t0 = time.perf_counter()
    V64 = oracle_state_energies(s, frames, threads)
    t_orc = time.perf_counter() - t0
    kT = 0.001986 * 300.0
    lam = np.asarray(lambdas)
    _, _, dG_m = fep.zwanzig(V_mixed, lam, kT)
    _, _, dG_64 = fep.zwanzig(V64, lam, kT)
    tg, keep = adjoint.tangent_struct(s)
    t0 = time.perf_counter()
    outs = bm.adjoint([tg] * nw)
    t_adj = time.perf_counter() - t0
    space = adjoint.QParamSpace(s, H12=80.0)
    VdV = [adjoint.energy_derivatives(o, space) for o in outs]
    t1 = time.perf_counter()
    g = adjoint.gradients([v for v, _ in VdV], [d for _, d in VdV], space, lam, kT, [1.0, -1.0])
    t_red = time.perf_counter() - t1
    E_adj = [v for v, _ in VdV]
    return {"frames_per_window": nfr, "windows": nw, "comparator": "FP64 oracle (C restatement) on the same frames",
            "oracle_s": round(t_orc, 1), "oracle_threads": threads,
            "dG_end_mixed": round(float(dG_m[-1]), 4), "dG_end_fp64_oracle": round(float(dG_64[-1]), 4),
            "max_abs_dG_diff_kcal_mol": float(np.abs(dG_m - dG_64).max()),
            "max_abs_state_energy_diff_kcal_mol": float(max(np.abs(a - b).max() for a, b in zip(V_mixed, V64))),
            "adjoint": {"launch_s": round(t_adj, 4), "frames": nw * nfr, "frames_per_s": round(nw * nfr / t_adj, 1),
                        "params": space.npar, "host_reduction_s": round(t_red, 4),
                        "max_abs_E_adjoint_vs_oracle": float(max(np.abs(a - b).max() for a, b in zip(E_adj, V64))),
                        "grad_dG_norm": float(np.linalg.norm(g["dG"]))}}


def cpu_baseline(s, ncycles=1):
    """Oracle (C restatement, 1 thread) on one window: one list/LRF update and
    10 full evaluations per cycle = 10 steps of the hot path."""
    from oracle.oracle import Oracle
    from q6_amd import abi
    o = Oracle(s)
    t0 = time.perf_counter()
    for _ in range(ncycles):
        o.make_pair_lists(s.x)
        for k in range(10):
            o.eval_terms(abi.T_ALL, s.x, Tfree=300.0, istep=k)
    dt = time.perf_counter() - t0
    steps = 10 * ncycles
    return steps / dt, dt, steps


def reference_baseline(s, lambdas, nproc, nsteps):
    """Stock Qdyn6 (serial; oracle/_ref/Qdyn6, built from the reference
    sources) on nproc of the windows at once, one process per host core, each
    `nsteps` MD steps of the same system from the same start: the topology,
    FEP file and inputs come from q6_amd.qfiles, whose tables Qdyn6 reproduces
    exactly.  Returns (ns/day summed over the processes from each one's own
    main-loop timer, src/md.f90:1924-1926; the slowest main loop; the wall
    time), or None when the build is absent or a run fails."""
    import shutil
    import subprocess
    import tempfile
    from q6_amd import qfiles
    exe = os.path.join(ROOT, "oracle", "_ref", "Qdyn6")
    if not os.access(exe, os.X_OK):
        return None
    tmp = tempfile.mkdtemp(prefix="q6cpu_")
    try:
        qfiles.write_top(s, os.path.join(tmp, "sys.top"))
        qfiles.write_fep(s, os.path.join(tmp, "sys.fep"), offdiag=[(1, 2, 1, 2, 80.0, 0.0)])
        pick = np.unique(np.linspace(0, len(lambdas) - 1, nproc).round().astype(int))
        runs = []
        for k, w in enumerate(pick):
            d = os.path.join(tmp, "w%d" % k)
            os.makedirs(d)
            qfiles.write_inp(s, os.path.join(d, "run.inp"), "../sys.top", "../sys.fep", nsteps, lambdas=lambdas[w],
                             seed=101 + k)
            runs.append(d)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([exe, "run.inp"], cwd=d, stdout=open(os.path.join(d, "log"), "w"),
                                  stderr=subprocess.STDOUT) for d in runs]
        rcs = [p.wait() for p in procs]
        wall = time.perf_counter() - t0
        loops = []
        for d, rc in zip(runs, rcs):
            txt = open(os.path.join(d, "log")).read()
            t = [ln.split(":")[1].split("(")[0] for ln in txt.splitlines() if ln.startswith("Total time of main loop")]
            if rc != 0 or "ABNORMAL" in txt or not t:
                print("# Qdyn6 run failed (rc %d):\n%s" % (rc, txt[-1500:]), file=sys.stderr, flush=True)
                return None
            loops.append(float(t[0]))
        ns_day = sum(nsteps / max(t, 1e-3) for t in loops) * 1.0 * 0.0864
        return ns_day, max(loops), wall, len(runs)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def host_cores():
    """Host cores this process may run on (its CPU affinity), at most 16."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--windows", type=int, default=51)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="MD steps per stock-Qdyn6 process of the CPU baseline (about 1 s per step at 10,861 atoms)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--frames", type=int, default=200,
                    help="frames per window stored after the timed run for the dG precision check and the adjoint")
    ap.add_argument("--workload", choices=["windows", "uneven"], default="windows",
                    help="windows: 51 lambda windows of the 30 A system (headline); uneven: 16 spheres of 15-45 A")
    ap.add_argument("--calibration-copy", type=int, default=0, metavar="MIB",
                    help="end with one device-to-device copy of MIB MiB (MIB read + MIB written), the known-size "
                         "reference for scaling hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: lambda windows would shard across GPUs (whole windows per rank, "
              "no per-step exchange), but only --gpus 1 is built and measured.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be positive and --warmup non-negative")
    if args.workload == "uneven":
        return uneven(args)

    from q6_amd import abi, synth
    s = synth.cfg2()
    nw = args.windows
    lam1 = np.linspace(0.0, 1.0, nw) if nw > 1 else np.array([0.5])
    lambdas = [np.array([l, 1.0 - l]) for l in lam1]
    vels = [synth.maxwell_velocities(s, 300.0, 100 + w) for w in range(nw)]
    print(f"# {synth.describe(s)}", file=sys.stderr, flush=True)
    s.x = equilibrate(s)
    xs = [s.x] * nw

    # batched mixed precision, all windows in one launch per kernel, graphs
    bm, tm = run_md([s] * nw, lambdas, vels, abi.QNB_PREC_MIXED, args.steps, args.warmup, positions=xs,
                    extra=10 * args.frames)
    steps_per_s = args.steps / tm
    ns_day = nw * steps_per_s * 1.0 * 0.0864
    all_counts = [bm.pair_counts(k) for k in range(nw)]  # list lengths of each window's last list update
    pairs_step = sum(s.pair_count_per_step(c) for c in all_counts)
    print(f"# batched mixed: {args.steps} steps x {nw} windows in {tm:.3f} s -> {ns_day:.1f} ns/day aggregate; "
          f"pairs/step over the windows {pairs_step}", file=sys.stderr, flush=True)
    # kernel timings on the same stream (HIP events), same batch
    bm_shapes = bm.shapes
    t_k = {"k_lists_lrf": bm.time_kernel(bm.KERNEL_LISTS, 20), "k_wforces": bm.time_kernel(bm.KERNEL_WFORCES, 20),
           "k_sforces": bm.time_kernel(bm.KERNEL_SFORCES, 20), "k_qforces": bm.time_kernel(bm.KERNEL_QFORCES, 20)}
    e = bm.energies(0)
    temps = e[:, abi.F_TEMP]
    prec = free_energy_precision(bm, s, lambdas, args.frames, host_cores()) if args.frames > 0 else None
    if prec:
        print(f"# dG precision: max |dG_mixed - dG_fp64| = {prec['max_abs_dG_diff_kcal_mol']:.2e} kcal/mol; "
              f"adjoint {prec['adjoint']['frames']} frames in {prec['adjoint']['launch_s']:.3f} s",
              file=sys.stderr, flush=True)
    bm.close()

    # algorithmic FLOPs per launch over all windows; the list sweep runs once
    # per 10-step cycle, the force kernels every step
    units, tests = lrf_units(s, s.x)
    ns = int(s.n["nstates"])
    flop = {"k_lists_lrf": nw * (units * F_LRF + tests * F_TEST),
            "k_wforces+k_sforces": sum(pair_flops(c) for c in all_counts),
            "k_qforces": sum(q_flops(c, ns) for c in all_counts)}
    ms = {"k_lists_lrf": t_k["k_lists_lrf"], "k_wforces+k_sforces": t_k["k_wforces"] + t_k["k_sforces"],
          "k_qforces": t_k["k_qforces"]}
    per_step = {"k_lists_lrf": ms["k_lists_lrf"] / 10.0, "k_wforces+k_sforces": ms["k_wforces+k_sforces"],
                "k_qforces": ms["k_qforces"]}
    traffic = measured_traffic(nw, int(s.n["natom"]))

    def entry(k):
        a = flop[k] / (ms[k] * 1e-3) / 1e12
        tb = None
        if traffic:
            parts = k.split("+")
            if all(p in traffic for p in parts):
                tb = int(sum(traffic[p] for p in parts))
        return {"kernel": k, "achieved": round(a, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                "frac": round(a / PEAK_FP32, 4), "traffic": tb, "ms_per_launch": round(ms[k], 4),
                "ms_per_step": round(per_step[k], 4), "flop_per_launch": int(flop[k]),
                "hbm_GBps": None if tb is None else round(tb / (ms[k] * 1e-3) / 1e9, 1),
                "hbm_frac": None if tb is None else round(tb / (ms[k] * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)}
    # the dominant kernel is the one with the largest share of the step
    dom = max(per_step, key=per_step.get)
    roofline = dict(bound="valu", **entry(dom))
    roofline["kernels"] = {k: entry(k) for k in flop}
    roofline["ms_per_launch_each"] = {k: round(v, 4) for k, v in t_k.items()}

    plain = None
    if not args.no_plain:
        bp, tp = run_md([s], [lambdas[nw // 2]], [vels[nw // 2]], abi.QNB_PREC_FP64, args.steps, args.warmup,
                        positions=xs[:1])
        plain = args.steps / tp * 1.0 * 0.0864
        bp.close()
        print(f"# plain FP64, one window: {args.steps} steps in {tp:.3f} s -> {plain:.1f} ns/day", file=sys.stderr,
              flush=True)

    cpu = None
    if not args.no_cpu:
        nproc = min(host_cores(), nw)
        ref = reference_baseline(s, lambdas, nproc, args.cpu_steps)
        if ref is not None:
            v, tloop, wall, nrun = ref
            cpu = {"value": round(v, 4), "unit": "ns/day", "cores": nrun, "kind": "reference",
                   "sample": f"stock Qdyn6 (serial, built from the reference sources) on {nrun} of the {nw} windows "
                             f"at once, one process per core, {args.cpu_steps} MD steps each of the same system from "
                             f"the same start; ns/day from each run's main-loop timer (slowest {tloop:.1f} s, "
                             f"{wall:.1f} s wall with set-up)"}
        else:
            sps, tc, nsteps = cpu_baseline(s, ncycles=5)
            cpu = {"value": round(sps * 1.0 * 0.0864, 4), "unit": "ns/day", "cores": 1, "kind": "port",
                   "sample": f"oracle (C restatement, 1 thread) on one window of the same system: 1 list/LRF "
                             f"update + {nsteps} full nonbonded+restraint evaluations ({tc:.1f} s); integration "
                             f"not included (stock Qdyn6 build absent or failed)"}
        print(f"# cpu baseline ({cpu['kind']}): {cpu['value']} ns/day on {cpu['cores']} cores", file=sys.stderr,
              flush=True)

    out = {
        "metric": "ns/day + pair-interactions/s, batched mixed-prec lambda-windows vs plain FP64, 1 GPU",
        "value": round(ns_day, 2), "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(tm / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32-pair/fp64-accumulate", "data": "synthetic (seeded sphere, random init)",
        "config": {"workload": f"{nw} lambda windows of a 30 A sphere: {s.n['natom']} atoms ({s.n['nat_solute']} solute "
                               f"in 400 ten-atom groups, {s.n['nwat']} TIP3P), 30 Q-atoms x 2 EVB states, "
                               f"1 fs, lists/LRF every 10 steps",
                   "windows": nw, "natom": int(s.n["natom"]), "precision": "mixed", "graph": True,
                   "launch_shape_waves_per_block": list(bm_shapes)},
        "pair_interactions_per_s": round(pairs_step * steps_per_s, 1),
        "plain_fp64_ns_day": None if plain is None else round(plain, 2),
        "speedup_vs_plain_fp64": None if plain is None else round(ns_day / plain, 2),
        "temperature_range_K": [round(float(temps.min()), 1), round(float(temps.max()), 1)] if len(temps) else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "free_energy_precision": prec,
    }
    print(json.dumps(out))
    if args.calibration_copy > 0:
        calibration_copy(args.calibration_copy)


def calibration_copy(mib):
    """One hipMemcpy device-to-device of mib MiB (the runtime's copy kernel:
    mib MiB read, mib MiB written) through the HIP runtime the engine loaded."""
    import ctypes as C
    path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln)
    hip = C.CDLL(path)
    n = C.c_size_t(mib << 20)
    a, c = C.c_void_p(), C.c_void_p()
    for ptr in (a, c):
        if hip.hipMalloc(C.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
    ok = hip.hipMemset(a, 1, n) == 0 and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(c, a, n, 3) == 0 and hip.hipDeviceSynchronize() == 0  # 3: device to device
    hip.hipFree(a)
    hip.hipFree(c)
    if not ok:
        raise RuntimeError("calibration copy failed")
    print(f"# calibration copy: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)


def uneven(args):
    """16 independent spheres of 15-45 A in one batch (mixed, autotuned
    shapes) against the same 16 run one at a time on the plain FP64 path."""
    from q6_amd import abi, synth
    systems = synth.cfg5()
    vels = [synth.maxwell_velocities(s, 300.0, 700 + k) for k, s in enumerate(systems)]
    lams = [np.asarray(s.par["lambda"]) for s in systems]
    xs = [equilibrate(s, steps=100, min_steps=200) for s in systems]
    b, t = run_md(systems, lams, vels, abi.QNB_PREC_MIXED, args.steps, args.warmup, positions=xs)
    shapes = b.shapes
    b.close()
    agg = len(systems) * args.steps / t * 0.0864
    seq_time = 0.0
    if not args.no_plain:
        for k, s in enumerate(systems):
            bp, tp = run_md([s], [lams[k]], [vels[k]], abi.QNB_PREC_FP64, args.steps, args.warmup, positions=[xs[k]])
            bp.close()
            seq_time += tp
    seq = len(systems) * args.steps / seq_time * 0.0864 if seq_time else None
    out = {"metric": "ns/day + pair-interactions/s, batched mixed-prec lambda-windows vs plain FP64, 1 GPU",
           "value": round(agg, 2), "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(t / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32-pair/fp64-accumulate", "data": "synthetic (seeded spheres)",
           "config": {"workload": "16 independent spheres, radii 15-45 A, " +
                                  f"{min(s.natom for s in systems)}-{max(s.natom for s in systems)} atoms, "
                                  "2-3 states, one batch", "launch_shape_waves_per_block": list(shapes)},
           "sequential_plain_fp64_ns_day": None if seq is None else round(seq, 2),
           "speedup_vs_sequential_plain_fp64": None if seq is None else round(agg / seq, 2),
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
