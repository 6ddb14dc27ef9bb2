"""Free-energy reductions of per-window state energies (Qfep6 semantics).

Restates the analysis Qfep6 applies to Qdyn6 energy files
(src/qfep.f90:402-648), on the per-window energy series the batched engine
samples on the device (q6_amd.md.BatchMD.energies):

  * state energies V_s = EQ(s)%total + alpha_s, alpha_1 = 0 (qfep.f90:448-452);
  * EVB ground state of a 2-state system
    E_g = (V1 + V2)/2 - sqrt((V1 - V2)^2 + 4 H12^2)/2 and the weight
    c1 = 1/(1 + ((E_g - V1)/H12)^2) (qfep.f90:492-500);
  * energy gap sum_s c_s V_s (qfep.f90:506-508);
  * Zwanzig: forward dgf_i = -kT ln< exp(-(V_{lambda_{i+1}} - V_{lambda_i})/kT) >_i
    over the frames of window i after nskip, reverse dgr_i from window i
    towards i-1, and dG_i = dG_{i-1} + (dgf_{i-1} - dgr_i)/2
    (qfep.f90:595-647).
All sums are FP64; the exponential average is taken in the reference's order
(plain sum of exp, then divide, then log).
"""
import numpy as np


def state_energies(V, alpha=None):
    """V: (npts, nstates) EQ(s)%total per frame; adds the alpha shifts."""
    V = np.asarray(V, np.float64)
    if alpha is None:
        return V.copy()
    a = np.asarray(alpha, np.float64).copy()
    a[0] = 0.0  # alfa(1) = zero (qfep.f90:447)
    return V + a[None, :]


def evb_ground(V, H12):
    """Ground-state energy and state-1 weight of a 2-state EVB system."""
    V = np.asarray(V, np.float64)
    v1, v2 = V[:, 0], V[:, 1]
    eg = 0.5 * (v1 + v2) - 0.5 * np.sqrt((v1 - v2) ** 2 + 4.0 * np.asarray(H12) ** 2)
    c1 = 1.0 / (1.0 + ((eg - v1) / H12) ** 2)
    return eg, c1


def gap(V, coeff):
    return np.asarray(V, np.float64) @ np.asarray(coeff, np.float64)


def zwanzig(Vs, lambdas, kT, nskip=0):
    """Vs: list over windows of (npts, nstates) state energies (alpha applied);
    lambdas: (nwin, nstates).  Returns dgf[nwin-1], dgr[nwin-1] (dgr[i] is the
    reverse step from window i+1 to i), dG[nwin] (dG[0] = 0)."""
    lam = np.asarray(lambdas, np.float64)
    n = len(Vs)
    dgf = np.zeros(max(n - 1, 0))
    dgr = np.zeros(max(n - 1, 0))
    for i in range(n - 1):
        V = np.asarray(Vs[i], np.float64)[nskip:]
        dv = V @ lam[i + 1] - V @ lam[i]
        dgf[i] = -kT * np.log(np.sum(np.exp(-dv / kT)) / len(V))
    for i in range(n - 1, 0, -1):
        V = np.asarray(Vs[i], np.float64)[nskip:]
        dv = V @ lam[i - 1] - V @ lam[i]
        dgr[i - 1] = -kT * np.log(np.sum(np.exp(-dv / kT)) / len(V))
    dG = np.zeros(n)
    for i in range(1, n):
        dG[i] = dG[i - 1] + 0.5 * (dgf[i - 1] - dgr[i - 1])
    return dgf, dgr, dG


def zwanzig_grad(Vs, dVs, lambdas, kT, nskip=0):
    """Derivative of the Zwanzig dG (last window) with respect to parameters,
    given per-frame derivatives dVs[i] (npts, nstates, nparam) of the state
    energies of window i, with the frames held fixed (the perturbation
    estimator's own derivative):
      d dgf_i = < w dDV > / < w >,  w = exp(-DV/kT),  DV = V.(lam_{i+1} - lam_i)
    and likewise for the reverse steps."""
    lam = np.asarray(lambdas, np.float64)
    n = len(Vs)
    npar = np.asarray(dVs[0]).shape[2]
    ddgf = np.zeros((max(n - 1, 0), npar))
    ddgr = np.zeros((max(n - 1, 0), npar))
    for i in range(n - 1):
        V = np.asarray(Vs[i], np.float64)[nskip:]
        dV = np.asarray(dVs[i], np.float64)[nskip:]
        dl = lam[i + 1] - lam[i]
        dv = V @ dl
        w = np.exp(-(dv - dv.min()) / kT)
        ddgf[i] = (w[:, None] * np.einsum("psk,s->pk", dV, dl)).sum(0) / w.sum()
    for i in range(n - 1, 0, -1):
        V = np.asarray(Vs[i], np.float64)[nskip:]
        dV = np.asarray(dVs[i], np.float64)[nskip:]
        dl = lam[i - 1] - lam[i]
        dv = V @ dl
        w = np.exp(-(dv - dv.min()) / kT)
        ddgr[i - 1] = (w[:, None] * np.einsum("psk,s->pk", dV, dl)).sum(0) / w.sum()
    return 0.5 * (ddgf.sum(0) - ddgr.sum(0))
