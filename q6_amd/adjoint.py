"""Gradients of Q-atom energy gaps and FEP/EVB free energies with respect to
the parameters EVB calibration fits.

The device side (qnb_adjoint / qnb_batch_adjoint, include/qnb.h) returns, per
stored frame, state s and Q-atom iq, the derivative of the state energy
E_s = EQ(s)%qq + %qp + %qw (src/potene.f90:224-229) with respect to iq's
scaled charge and its LJ factors A, B (normal and 1-4 pairs), plus iq's share
of E_s.  This module maps those per-frame sums onto a parameter vector

    theta = [ q(iq, s) for s, iq ]   Q charges in e (unscaled), state-major
          + [ sigma_t ], [ eps_t ]   per Q-atom LJ type (geometric rule,
                                     A = 2 sqrt(eps) sigma^6, B = 2 sqrt(eps) sigma^3)
          + [ alpha_s, s >= 2 ]      gas-phase shifts (alpha_1 = 0, src/qfep.f90:447)
          + [ H12 ]                  EVB coupling (constant off-diagonal)

and reduces them with the Qfep6 estimators of q6_amd.fep, for fixed frames
(the derivative of the estimator itself, which is what central finite
differences over the same stored frames measure).
"""
import ctypes as C

import numpy as np

from . import abi, fep, synth
from .synth import SQRT332


class QParamSpace:
    """The parameter vector of a synthetic system's Q-atoms (see module doc)."""

    def __init__(self, s, alpha=None, H12=80.0):
        md = s.md
        self.nq, self.ns = s.n["nqat"], s.n["nstates"]
        self.q = np.array(md["qcrg"], float).reshape(self.nq, self.ns)
        self.sig = np.array(md["q_sig"], float)
        self.eps = np.array(md["q_eps"], float)
        self.type_of = np.asarray(md["q_type_of"])
        self.nt = len(self.sig)
        a = np.zeros(self.ns) if alpha is None else np.array(alpha, float)
        a[0] = 0.0
        self.alpha = a
        self.H12 = float(H12)
        nq, ns, nt = self.nq, self.ns, self.nt
        self.sl_q = slice(0, nq * ns)
        self.sl_sig = slice(nq * ns, nq * ns + nt)
        self.sl_eps = slice(nq * ns + nt, nq * ns + 2 * nt)
        self.sl_alpha = slice(nq * ns + 2 * nt, nq * ns + 2 * nt + ns - 1)
        self.i_H12 = nq * ns + 2 * nt + ns - 1
        self.npar = self.i_H12 + 1

    def q_index(self, iq, s):
        return s * self.nq + iq

    def vector(self):
        return np.concatenate([self.q.T.ravel(), self.sig, self.eps, self.alpha[1:], [self.H12]])

    def with_vector(self, v):
        """A copy carrying the parameters v."""
        o = QParamSpace.__new__(QParamSpace)
        o.__dict__.update(self.__dict__)
        v = np.asarray(v, float)
        o.q = v[self.sl_q].reshape(self.ns, self.nq).T.copy()
        o.sig = v[self.sl_sig].copy()
        o.eps = v[self.sl_eps].copy()
        o.alpha = np.concatenate([[0.0], v[self.sl_alpha]])
        o.H12 = float(v[self.i_H12])
        return o

    def names(self):
        n = [f"q[{iq},{s}]" for s in range(self.ns) for iq in range(self.nq)]
        n += [f"sigma[{t}]" for t in range(self.nt)] + [f"eps[{t}]" for t in range(self.nt)]
        n += [f"alpha[{s}]" for s in range(1, self.ns)] + ["H12"]
        return n

    def dH(self):
        """Derivative of H12 with respect to theta."""
        e = np.zeros(self.npar)
        e[self.i_H12] = 1.0
        return e

    def lj_factors(self, s):
        return synth.q_lj_factors(s, self.sig, self.eps)

    def apply(self, s):
        """Rebuild the system's Q tables for these parameters (in place)."""
        qA, qB = self.lj_factors(s)
        return synth.set_q_params(s, self.q, qA, qB)


def tangent_struct(s):
    """qnb_qtangent of a synthetic system's current Q tables; returns the
    struct and the arrays it points to (keep both alive)."""
    md = s.md
    t = synth.q_tangents(md["qctx"], md["qcrg"], md["qA"], md["qB"])
    keep = {}
    st = abi.QnbQTangent()
    for name in abi.QTANGENT_FIELDS:
        if name.endswith("is14"):
            a = np.ascontiguousarray(t[name], np.uint8)
            setattr(st, name, a.ctypes.data_as(C.POINTER(C.c_uint8)))
        else:
            a = np.ascontiguousarray(t[name], np.float64)
            setattr(st, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        keep[name] = a
    return st, keep


def energy_derivatives(out, space):
    """out: (nframes, ns, nq, QNB_ADJ_NV) adjoint records of one instance.
    Returns V (nframes, ns) = E_s + alpha_s and dV (nframes, ns, npar)."""
    out = np.asarray(out, np.float64)
    nf, ns, nq = out.shape[:3]
    E = (out[..., abi.ADJ_EEL] + out[..., abi.ADJ_EVDW]).sum(2)
    V = E + space.alpha[None, :]
    dV = np.zeros((nf, ns, space.npar))
    for s in range(ns):
        dV[:, s, space.q_index(0, s):space.q_index(0, s) + nq] = SQRT332 * out[:, s, :, abi.ADJ_DQ]
    gA = out[..., abi.ADJ_DA] + out[..., abi.ADJ_DA14]  # dE_s/dA_iq (every LJ code shares A_iq)
    gB = out[..., abi.ADJ_DB] + out[..., abi.ADJ_DB14]
    sig, eps = space.sig[space.type_of], space.eps[space.type_of]
    A = 2.0 * np.sqrt(eps) * sig ** 6
    B = 2.0 * np.sqrt(eps) * sig ** 3
    dsig = gA * (6.0 * A / sig) + gB * (3.0 * B / sig)  # (nf, ns, nq)
    deps = gA * (A / (2.0 * eps)) + gB * (B / (2.0 * eps))
    for t in range(space.nt):
        m = space.type_of == t
        dV[:, :, space.sl_sig.start + t] = dsig[:, :, m].sum(2)
        dV[:, :, space.sl_eps.start + t] = deps[:, :, m].sum(2)
    for s in range(1, ns):
        dV[:, s, space.sl_alpha.start + s - 1] = 1.0
    return V, dV


def analysis(Vs, lambdas, kT, H12, coeff, nskip=0, nbins=None, nmin=1, frame_bins=None):
    """The quantities differentiated below, from state energies alone."""
    _, _, dG = fep.zwanzig(Vs, lambdas, kT, nskip)
    gap = np.array([(np.asarray(V)[nskip:] @ coeff).mean() for V in Vs])
    r = dict(dG=dG[-1], gap=gap, G_evb=fep.evb_window_free_energy(Vs, lambdas, H12, kT, nskip))
    if nbins:
        p = fep.evb_pmf(Vs, lambdas, H12, kT, coeff, nbins, nmin, nskip, frame_bins)
        r["pmf"], r["frame_bins"] = p["binned"], p["frame_bins"]
    return r


def gradients(Vs, dVs, space, lambdas, kT, coeff, nskip=0, nbins=None, nmin=1):
    """d(analysis)/d(theta) for fixed frames: Zwanzig dG of the last window,
    the mean gap of every window, the per-window EVB ground-state free energy
    and (with nbins) the binned reaction free energy with bin membership held
    fixed."""
    coeff = np.asarray(coeff, float)
    dH = space.dH()
    g = dict(dG=fep.zwanzig_grad(Vs, dVs, lambdas, kT, nskip),
             gap=np.array([np.einsum("psk,s->k", np.asarray(d)[nskip:], coeff) / (len(d) - nskip) for d in dVs]),
             G_evb=fep.evb_window_free_energy_grad(Vs, dVs, dH, lambdas, space.H12, kT, nskip))
    if nbins:
        g["pmf"] = fep.evb_pmf_grad(Vs, dVs, dH, lambdas, space.H12, kT, coeff, nbins, nmin, nskip)
    return g


__all__ = ["QParamSpace", "tangent_struct", "energy_derivatives", "analysis", "gradients"]
