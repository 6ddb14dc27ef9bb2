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
    """Polarisation-shell constant cstb for the lambda-weighted Q charge
    (wat_sphere crgQtot, src/simprep.f90:4640-4650; wat_shells
    src/simprep.f90:4883)."""
    qcrg = s.md.get("qcrg")
    if qcrg is None or qcrg.size == 0:
        return np.asarray(s.par["wshell_cstb"])
    from .synth import RHO_WAT, SQRT332, TIP3P_HOH, TIP3P_ROH, TIP3P_Q
    import math
    crgQtot = float((qcrg * SQRT332 * np.asarray(lam)[None, :]).sum())
    mu_w = -TIP3P_Q[0] * SQRT332 * TIP3P_ROH * math.cos(math.radians(TIP3P_HOH) / 2)
    out = []
    for rout, dr in zip(s.par["wshell_rout"], s.par["wshell_dr"]):
        ri = rout - dr
        rshell = (0.5 * (rout ** 3 + ri ** 3)) ** (1.0 / 3.0)
        out.append(crgQtot * (1.0 - 1.0 / 80.0) / (RHO_WAT * mu_w * 4.0 * math.pi * rshell ** 2))
    return np.array(out)


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

    # kernels qnb_batch_time_kernel times
    KERNEL_LISTS, KERNEL_FORCES, KERNEL_WFORCES, KERNEL_SFORCES, KERNEL_QFORCES = 0, 1, 2, 3, 4

    def time_kernel(self, which, reps):
        """Average launch duration (ms) of one kernel of the step over all
        instances, timed with HIP events on the batch's stream."""
        ms = C.c_double()
        abi.check(self.lib, self.lib.qnb_batch_time_kernel(self.h, which, reps, C.byref(ms)), "time_kernel")
        return ms.value


def q_state_energies(frames, nstates):
    """Per-state energies EQ(s)%total of the FEP/EVB analysis
    (src/potene.f90:226-231): the Q bonded terms plus EQ(s)%qx = qq + qp + qw
    (el + vdw); the engine has no Q restraints."""
    out = np.zeros((len(frames), nstates))
    for s in range(nstates):
        b = abi.F_Q0 + 6 * s
        out[:, s] = frames[:, b:b + 6].sum(1) + frames[:, abi.F_QB0 + 4 * s:abi.F_QB0 + 4 * s + 4].sum(1)
    return out


__all__ = ["BatchMD", "q_state_energies", "scaas_cstb", "BOLTZ"]
