"""Batched device-resident MD (lambda windows, uneven systems) over the C ABI.

`BatchMD` advances several instances in the same kernel launches
(qnb_batch_*): every lambda window of an FEP run, or independent systems of
different size.  Energies are sampled on the device; `energies()` returns the
per-instance time series, including the per-state Q-atom energies the Zwanzig
and EVB reductions need (q6_amd.fep).
"""
import ctypes as C

import numpy as np

from . import abi
from .synth import BOLTZ

_pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
_pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731


def scaas_cstb(s, lam):
    """Polarisation-shell constants cstb of a synthetic system at the state
    weights lam (wat_sphere crgQtot and wat_shells, src/simprep.f90:4640-4650,
    4863-4883; q6_amd.synth.shell_cstb)."""
    qcrg = s.md.get("qcrg") if s.md else None
    if qcrg is None or qcrg.size == 0:
        return np.asarray(s.par["wshell_cstb"])
    from .synth import shell_cstb
    return shell_cstb(qcrg, lam, s.par["wshell_rout"], s.par["wshell_dr"])


class BatchMD:
    def __init__(self, systems, lambdas, velocities, precision=abi.QNB_PREC_FP64, dt_fs=1.0, T0=300.0,
                 tau_T_fs=100.0, nbcycle=10, ene_every=10, max_frames=1000, positions=None,
                 separate_scaling=True, water_geometry=None):
        """systems[k]: QSystem of instance k (repeat the same object for the
        windows of one system: they share its device topology); lambdas[k]:
        state weights; velocities[k]: (natom, 3) initial velocities.
        separate_scaling: solute and solvent as separate Berendsen groups
        (Qdyn6's default); water_geometry: (r_OH, r_HH) of the SHAKE
        constraints when they differ from the system's."""
        self.lib = abi.load()
        n = len(systems)
        self.n = n
        self.systems = systems
        self._keep = []
        sys_structs = {}
        sp, pp, mp, xp, vp = ((C.c_void_p * n)() for _ in range(5))
        for k, s in enumerate(systems):
            key = id(s)
            if key not in sys_structs:
                st, _ = s.to_structs(precision=precision)
                sys_structs[key] = st
                self._keep.append(st)
            self._sys_struct = getattr(self, "_sys_struct", {})
            self._sys_struct[k] = sys_structs[key]
            lam = np.asarray(lambdas[k], float)
            _, par = s.to_structs(precision=precision, **{"lambda": lam, "wshell_cstb": scaas_cstb(s, lam),
                                                          "itdis": 100.0 / dt_fs})
            par.precision = precision
            md = s.md
            # harmonic bonds, then the harmonic distance restraints (same form)
            bonds = np.ascontiguousarray(md["bonds"], np.int32).reshape(-1, 2)
            rst = np.asarray(md.get("restraints", np.zeros((0, 2), np.int32)), np.int32).reshape(-1, 2)
            r0 = np.concatenate([np.broadcast_to(md["bond_r0"], len(bonds)),
                                 np.broadcast_to(md.get("restraint_r0", 0.0), len(rst))])
            kb = np.concatenate([np.broadcast_to(md["bond_k"], len(bonds)),
                                 np.broadcast_to(md.get("restraint_k", 0.0), len(rst))])
            bonds = np.concatenate([bonds, rst])
            angles = np.ascontiguousarray(md["angles"], np.int32).reshape(-1, 3)
            arrs = dict(bond_ij=bonds.ravel(), bond_r0=r0.astype(np.float64), bond_k=kb.astype(np.float64),
                        angle_ijk=angles.ravel(),
                        angle_t0=np.full(len(angles), md["angle_t0"]), angle_k=np.full(len(angles), md["angle_k"]),
                        winv=1.0 / np.asarray(md["mass"], float))
            arrs = {a: np.ascontiguousarray(v) for a, v in arrs.items()}
            roh, rhh = water_geometry if water_geometry is not None else (md["water_roh"], md["water_rhh"])
            m = abi.QnbMdParams(dt_fs=dt_fs, T0=T0, tau_T_fs=tau_T_fs, nbcycle=nbcycle, ene_every=ene_every,
                                max_frames=max_frames, nbonds=len(bonds), nangles=len(angles),
                                water_roh=roh, water_rhh=rhh, separate_scaling=int(bool(separate_scaling)))
            m.bond_ij, m.angle_ijk = _pi(arrs["bond_ij"]), _pi(arrs["angle_ijk"])
            for a in ("bond_r0", "bond_k", "angle_t0", "angle_k", "winv"):
                setattr(m, a, _pd(arrs[a]))
            x0 = np.ascontiguousarray(positions[k] if positions is not None else s.x, np.float64)
            v0 = np.ascontiguousarray(velocities[k], np.float64)
            self._keep += [par, m, arrs, x0, v0]
            sp[k] = C.addressof(sys_structs[key])
            pp[k] = C.addressof(par)
            mp[k] = C.addressof(m)
            xp[k] = x0.ctypes.data
            vp[k] = v0.ctypes.data
        h = C.c_void_p()
        abi.check(self.lib, self.lib.qnb_batch_create(n, sp, pp, mp, xp, vp, C.byref(h)), "qnb_batch_create")
        self.h = h
        self.max_frames = max_frames
        self.dt_fs = dt_fs

    def close(self):
        if getattr(self, "h", None):
            self.lib.qnb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, nsteps, graph=True):
        abi.check(self.lib, self.lib.qnb_batch_run(self.h, nsteps, int(graph)), "qnb_batch_run")

    def integrate(self, grads):
        """One leap-frog + SHAKE + temperature step of every instance from the
        given gradients (grads[k]: (natom, 3))."""
        gs = [np.ascontiguousarray(g, np.float64) for g in grads]
        arr = (C.c_void_p * self.n)(*[g.ctypes.data for g in gs])
        abi.check(self.lib, self.lib.qnb_batch_integrate(self.h, arr), "qnb_batch_integrate")

    def thermostat(self, k):
        """Instance k's thermostat state after the last step: dict with Temp,
        Tfree, Ekin and per temperature group the free temperature `tfree` and
        the Berendsen factor `sfact` of the next step."""
        out = np.zeros(8)
        abi.check(self.lib, self.lib.qnb_batch_thermostat(self.h, k, _pd(out)), "qnb_batch_thermostat")
        ng = int(out[7])
        return {"Temp": out[0], "Tfree": out[1], "Ekin": out[2], "ngroups": ng,
                "tfree": out[3:3 + ng].copy(), "sfact": out[5:5 + ng].copy()}

    def prepare_graph(self):
        """Capture the list-cycle graph now (no step is taken)."""
        abi.check(self.lib, self.lib.qnb_batch_run(self.h, 0, 1), "qnb_batch_run")

    def minimize(self, nsteps, max_disp=0.02):
        """Steepest-descent relaxation of the start structures (set-up aid)."""
        abi.check(self.lib, self.lib.qnb_batch_minimize(self.h, nsteps, max_disp), "qnb_batch_minimize")

    def autotune(self, reps=3):
        """Pick the launch shapes of the two heavy kernels by timing (returns
        waves per block for the list/LRF sweep and for the force kernel)."""
        out = np.zeros(2, np.int32)
        abi.check(self.lib, self.lib.qnb_batch_autotune(self.h, reps, _pi(out)), "qnb_batch_autotune")
        return int(out[0]), int(out[1])

    def sync(self):
        abi.check(self.lib, self.lib.qnb_batch_sync(self.h), "qnb_batch_sync")

    def state(self, k):
        nat = self.systems[k].natom
        x = np.zeros((nat, 3))
        v = np.zeros((nat, 3))
        abi.check(self.lib, self.lib.qnb_batch_get_state(self.h, k, _pd(x), _pd(v)), "qnb_batch_get_state")
        return x, v

    def energies(self, k):
        """(nframes, QNB_FRAME_LEN) history of instance k."""
        nf = C.c_int32()
        abi.check(self.lib, self.lib.qnb_batch_energy_history(self.h, k, None, 0, C.byref(nf)), "history")
        out = np.zeros((max(nf.value, 1), abi.QNB_FRAME_LEN))
        abi.check(self.lib, self.lib.qnb_batch_energy_history(self.h, k, _pd(out), nf.value, C.byref(nf)), "history")
        return out[:nf.value]

    def pair_counts(self, k):
        out = np.zeros(7, np.int64)
        abi.check(self.lib, self.lib.qnb_batch_pair_counts(self.h, k, out.ctypes.data_as(C.POINTER(C.c_int64))),
                  "qnb_batch_pair_counts")
        keys = ("pp", "pw", "ww", "qp", "qw", "qq", "qqp")
        return dict(zip(keys, (int(v) for v in out)))

    def set_offdiag(self, k, pairs, A, mu):
        """EVB off-diagonals of instance k sampled with the energies: pairs of
        0-based Q-atom numbers (k, l); H = A exp(-mu r_kl) (offdiag)."""
        kl = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1))
        A = np.ascontiguousarray(A, np.float64)
        mu = np.ascontiguousarray(mu, np.float64)
        abi.check(self.lib, self.lib.qnb_batch_set_offdiag(self.h, k, len(A), _pi(kl), _pd(A), _pd(mu)),
                  "qnb_batch_set_offdiag")

    def set_qbonded(self, k, tables, urey=False):
        """Q-atom bonded terms of instance k (abi.qbonded_struct layout; None
        removes them): their lambda-weighted forces act in every step and the
        per-state energies are sampled at abi.F_QB0 + 4 s."""
        if tables is None:
            abi.check(self.lib, self.lib.qnb_batch_set_qbonded(self.h, k, None), "qnb_batch_set_qbonded")
            return
        q, keep = abi.qbonded_struct(tables, urey)
        abi.check(self.lib, self.lib.qnb_batch_set_qbonded(self.h, k, C.byref(q)), "qnb_batch_set_qbonded")

    def store_frames(self, every, max_frames):
        """Keep a coordinate snapshot every `every` steps (max_frames per
        instance) on the device, for the adjoint."""
        abi.check(self.lib, self.lib.qnb_batch_store_frames(self.h, every, max_frames), "qnb_batch_store_frames")

    def frames(self, k):
        nf = C.c_int32()
        abi.check(self.lib, self.lib.qnb_batch_frames(self.h, k, None, 0, C.byref(nf)), "qnb_batch_frames")
        out = np.zeros((max(nf.value, 1), self.systems[k].natom, 3))
        abi.check(self.lib, self.lib.qnb_batch_frames(self.h, k, _pd(out), nf.value, C.byref(nf)), "qnb_batch_frames")
        return out[:nf.value]

    def adjoint(self, tangents):
        """Adjoint over every instance's stored frames (one launch); tangents[k]
        is instance k's qnb_qtangent.  Returns per instance an array
        (nframes, nstates, nqat, QNB_ADJ_NV)."""
        tp = (C.c_void_p * self.n)(*[C.addressof(t) for t in tangents])
        n = C.c_int64()
        abi.check(self.lib, self.lib.qnb_batch_adjoint(self.h, tp, None, 0, C.byref(n)), "qnb_batch_adjoint")
        buf = np.zeros(max(n.value, 1))
        abi.check(self.lib, self.lib.qnb_batch_adjoint(self.h, tp, _pd(buf), n.value, C.byref(n)), "qnb_batch_adjoint")
        res, off = [], 0
        for k, s in enumerate(self.systems):
            ns, nq = s.n["nstates"], s.n["nqat"]
            per = ns * nq * abi.QNB_ADJ_NV
            nf = (self._nframes(k))
            res.append(buf[off:off + nf * per].reshape(nf, ns, nq, abi.QNB_ADJ_NV))
            off += nf * per
        return res

    def _nframes(self, k):
        nf = C.c_int32()
        abi.check(self.lib, self.lib.qnb_batch_frames(self.h, k, None, 0, C.byref(nf)), "qnb_batch_frames")
        return nf.value

    def set_groups(self, k, masks, types):
        """Group-contribution arrays of instance k (qnb_batch_set_groups):
        masks (narr, nat_solute) 0/1, types (narr,) abi.GC_*; computed on
        every sampled step from that step's energy record."""
        masks = np.ascontiguousarray(np.asarray(masks, np.uint8).reshape(len(types), self._sys_struct[k].nat_solute))
        types = np.ascontiguousarray(types, np.int32)
        self._gc = getattr(self, "_gc", {})
        self._gc[k] = (masks, types)
        abi.check(self.lib, self.lib.qnb_batch_set_groups(self.h, k, C.byref(self._sys_struct[k]), len(types),
                                                           masks.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                           types.ctypes.data_as(C.POINTER(C.c_int32))),
                  "qnb_batch_set_groups")

    def group_energies(self, k):
        """(nframes, narr, nstates, 7) group energies of instance k's frames."""
        narr = len(self._gc[k][1]) if getattr(self, "_gc", None) and k in self._gc else 0
        ns = int(self.systems[k].n["nstates"])
        nf = C.c_int32()
        abi.check(self.lib, self.lib.qnb_batch_group_energies(self.h, k, None, 0, C.byref(nf)), "group_energies")
        out = np.zeros((max(nf.value, 1), narr, ns, abi.GC_NV))
        abi.check(self.lib, self.lib.qnb_batch_group_energies(self.h, k, _pd(out), nf.value, C.byref(nf)),
                  "group_energies")
        return out[:nf.value]

    def free_energy(self, kT, nskip=0, nbins=0, nmin=0, alpha=None, coeff=None, A=0.0, mu=0.0, eta=0.0, r0=0.0):
        """Qfep6's analysis on the GPU (qnb_batch_free_energy) of the frames
        every instance has recorded, one window per instance: Zwanzig dG, and
        with nbins > 0 the EVB reaction free energy along the energy gap."""
        ns = int(self.systems[0].n["nstates"])

        def call(p, out):
            abi.check(self.lib, self.lib.qnb_batch_free_energy(self.h, p, out), "qnb_batch_free_energy")
        return _fep_call(call, self.n, ns, kT, nskip, nbins, nmin, alpha, coeff, A, mu, eta, r0)

    # kernels qnb_batch_time_kernel times
    KERNEL_LISTS, KERNEL_FORCES, KERNEL_WFORCES, KERNEL_SFORCES, KERNEL_QFORCES = 0, 1, 2, 3, 4

    def time_kernel(self, which, reps):
        """Average launch duration (ms) of one kernel of the step over all
        instances, timed with HIP events on the batch's stream."""
        ms = C.c_double()
        abi.check(self.lib, self.lib.qnb_batch_time_kernel(self.h, which, reps, C.byref(ms)), "time_kernel")
        return ms.value


def _fep_call(call, n, ns, kT, nskip, nbins, nmin, alpha, coeff, A, mu, eta, r0):
    """Fills qnb_fep_params, allocates the result arrays, runs `call(p, out)`
    and returns the results as a dict (dgf[n-1] forward, dgr[n-1] reverse
    from window i+1 to i, as q6_amd.fep.zwanzig returns them)."""
    p = abi.QnbFepParams(kT=float(kT), nskip=int(nskip), nbins=int(nbins), nmin=int(nmin), A=float(A),
                         mu=float(mu), eta=float(eta), r0=float(r0))
    al = np.zeros(abi.QNB_MAX_STATES) if alpha is None else np.asarray(alpha, float)
    co = np.zeros(abi.QNB_MAX_STATES) if coeff is None else np.asarray(coeff, float)
    for k in range(len(al)):
        p.alpha[k] = float(al[k])
    for k in range(len(co)):
        p.coeff[k] = float(co[k])
    B = max(int(nbins), 0)
    r = dict(dgf=np.zeros(n), dgr=np.zeros(n), dG=np.zeros(n), gap_range=np.zeros(2), dGg=np.zeros((n, B)),
             dGv=np.zeros((n, B, ns)), c1=np.zeros((n, B)), c2=np.zeros((n, B)), r=np.zeros((n, B)),
             npts=np.zeros((n, B), np.int32), binned=np.zeros(B), binned_c1=np.zeros(B), binned_c2=np.zeros(B),
             binned_r=np.zeros(B), ptsum=np.zeros(B, np.int32))
    out = abi.QnbFepOut(dgf=_pd(r["dgf"]), dgr=_pd(r["dgr"]), dG=_pd(r["dG"]), gap_range=_pd(r["gap_range"]))
    if B:
        for f, k in (("dGg", "dGg"), ("dGv", "dGv"), ("c1", "c1"), ("c2", "c2"), ("r", "r"), ("bin_dGg", "binned"),
                     ("bin_c1", "binned_c1"), ("bin_c2", "binned_c2"), ("bin_r", "binned_r")):
            setattr(out, f, _pd(r[k]))
        out.npts = _pi(r["npts"])
        out.ptsum = _pi(r["ptsum"])
    call(C.byref(p), C.byref(out))
    r["dgf"], r["dgr"] = r["dgf"][:n - 1], r["dgr"][1:]
    if B:
        lo, hi = r["gap_range"]
        r["centres"] = lo + (np.arange(B) + 0.5) * ((hi - lo) / B)
    return r


def free_energy(Vs, lambdas, kT, rkl=None, nskip=0, nbins=0, nmin=0, alpha=None, coeff=None, A=0.0, mu=0.0, eta=0.0,
                r0=0.0):
    """Qfep6's analysis on the GPU (qnb_free_energy) of state-energy series
    the caller holds: Vs[w] (nframes_w, nstates) EQ(s)%total per frame of
    window w, lambdas (nwin, nstates), rkl[w] (nframes_w,) the first
    off-diagonal's distance (or None).  Returns the dict of _fep_call."""
    lib = abi.load()
    lam = np.ascontiguousarray(lambdas, np.float64)
    n, ns = lam.shape
    Vs = [np.ascontiguousarray(v, np.float64) for v in Vs]
    nf = np.array([len(v) for v in Vs], np.int32)
    Vp = (C.POINTER(C.c_double) * n)(*[_pd(v) for v in Vs])
    rk = None
    if rkl is not None:
        rkl = [np.ascontiguousarray(v, np.float64) for v in rkl]
        rk = (C.POINTER(C.c_double) * n)(*[_pd(v) for v in rkl])

    def call(p, out):
        abi.check(lib, lib.qnb_free_energy(n, ns, _pi(nf), Vp, rk, _pd(lam), p, out), "qnb_free_energy")
    return _fep_call(call, n, ns, kT, nskip, nbins, nmin, alpha, coeff, A, mu, eta, r0)


def q_state_energies(frames, nstates):
    """Per-state energies EQ(s)%total of the FEP/EVB analysis
    (src/potene.f90:226-231): the Q bonded terms plus EQ(s)%qx = qq + qp + qw
    (el + vdw); the engine has no Q restraints."""
    out = np.zeros((len(frames), nstates))
    for s in range(nstates):
        b = abi.F_Q0 + 6 * s
        out[:, s] = frames[:, b:b + 6].sum(1) + frames[:, abi.F_QB0 + 4 * s:abi.F_QB0 + 4 * s + 4].sum(1)
    return out


__all__ = ["BatchMD", "free_energy", "q_state_energies", "scaas_cstb", "BOLTZ"]
