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
    with np.errstate(divide="ignore", invalid="ignore"):  # H12 = 0: c1 undefined, as in the reference
        c1 = 1.0 / (1.0 + ((eg - v1) / H12) ** 2)
    return eg, c1


def offdiag_coupling(rkl, A, mu, eta=0.0, r0=0.0):
    """Distance-dependent EVB coupling of Qfep6 (src/qfep.f90:476-480):
    H12 = A exp(-mu (r - r0)) exp(-eta (r - r0)^2), r the off-diagonal
    distance r_kl each frame of the energy file carries."""
    d = np.asarray(rkl, np.float64) - r0
    return A * np.exp(-mu * d) * np.exp(-eta * d ** 2)


def _window_h12(H12, i, nskip):
    """H12 of window i: a constant, or per window an array over its frames."""
    if np.ndim(H12) == 0:
        return float(H12)
    return np.asarray(H12[i], np.float64)[nskip:]


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


def zwanzig_grad(Vs, dVs, lambdas, kT, nskip=0, per_window=False):
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
    if per_window:  # d dG_i for every window (dG_0 = 0)
        return np.concatenate([np.zeros((1, npar)), np.cumsum(0.5 * (ddgf - ddgr), axis=0)])
    return 0.5 * (ddgf.sum(0) - ddgr.sum(0))


def _evb_terms(V, H12):
    """E_g and its derivatives with respect to V1, V2 and H12 (2 states)."""
    v1, v2 = V[:, 0], V[:, 1]
    root = np.sqrt((v1 - v2) ** 2 + 4.0 * H12 ** 2)
    eg = 0.5 * (v1 + v2) - 0.5 * root
    d1 = 0.5 - 0.5 * (v1 - v2) / root
    d2 = 0.5 + 0.5 * (v1 - v2) / root
    dh = -2.0 * H12 / root
    return eg, d1, d2, dh


def _ground_dv(V, dV, lam, H12, dH):
    """dvg = E_g - sum_s lam_s V_s and its parameter derivative per frame."""
    eg, d1, d2, dh = _evb_terms(V, H12)
    dvg = eg - V @ lam
    ddvg = d1[:, None] * dV[:, 0, :] + d2[:, None] * dV[:, 1, :] + dh[:, None] * dH[None, :] \
        - np.einsum("psk,s->pk", dV, lam)
    return dvg, ddvg


def evb_window_free_energy(Vs, lambdas, H12, kT, nskip=0):
    """Ground-state free energy at every window, Qfep6 Part 2 with the whole
    window in one bin (src/qfep.f90:716-795): dGg_i = dG_i - kT ln <exp(-(E_g
    - eps_i)/kT)>_i, eps_i = sum_s lambda_s V_s, taken about the mean of
    E_g - eps_i as the reference does.  H12 is a constant or, per window, an
    array over its frames (offdiag_coupling)."""
    lam = np.asarray(lambdas, np.float64)
    _, _, dG = zwanzig(Vs, lam, kT, nskip)
    out = np.zeros(len(Vs))
    for i, V in enumerate(Vs):
        V = np.asarray(V, np.float64)[nskip:]
        eg, _ = evb_ground(V, _window_h12(H12, i, nskip))
        dvg = eg - V @ lam[i]
        av = dvg.sum() / len(dvg)
        out[i] = dG[i] + (-kT * np.log(np.sum(np.exp(-(dvg - av) / kT)) / len(dvg)) + av)
    return out


def evb_window_free_energy_grad(Vs, dVs, dH, lambdas, H12, kT, nskip=0):
    """Derivative of evb_window_free_energy (nwin, nparam) for fixed frames;
    dH[nparam] is the derivative of H12 with respect to the parameters."""
    lam = np.asarray(lambdas, np.float64)
    dH = np.asarray(dH, np.float64)
    gdG = zwanzig_grad(Vs, dVs, lam, kT, nskip, per_window=True)
    out = np.zeros_like(gdG)
    for i, (V, dV) in enumerate(zip(Vs, dVs)):
        V = np.asarray(V, np.float64)[nskip:]
        dV = np.asarray(dV, np.float64)[nskip:]
        dvg, ddvg = _ground_dv(V, dV, lam[i], H12, dH)
        w = np.exp(-(dvg - dvg.min()) / kT)
        out[i] = gdG[i] + (w[:, None] * ddvg).sum(0) / w.sum()
    return out


def evb_pmf(Vs, lambdas, H12, kT, coeff, nbins, nmin, nskip=0, frame_bins=None, rkl=None):
    """Reaction free energy along the energy gap, Qfep6 Parts 2 and 3
    (src/qfep.f90:700-827), 2-state EVB.  The gap sum_s coeff_s V_s of every
    frame after nskip is binned between the smallest and largest gap of all
    windows (bin = int((gap - gapmin)/xint), 0-based; the largest gap lands in
    an extra bin that is never reported, as in the reference).  Per window and
    bin with at least nmin points: dGg = dG_i - kT ln<exp(-(E_g -
    eps_i)/kT)>_bin; Part 3 averages dGg over windows weighted by points.
    frame_bins (list of int arrays) overrides the binning, e.g. to hold it
    fixed under a parameter perturbation.  H12 is a constant or, per window,
    an array over its frames (offdiag_coupling); rkl (per window, over its
    frames) adds the bin averages of the off-diagonal distance Part 3 prints
    as <r_xy> (qfep.f90:748, 780).  Returns a dict: gap centres,
    dGg[nwin, nbins] (ground state), dGv[nwin, nbins, nstates] (diabatic
    curves), c1[nwin, nbins] (mean state-1 weight), npts[nwin, nbins],
    binned[nbins], binned_c1[nbins] and binned_r[nbins] (Part 3; binned_r
    NaN without rkl), ptsum[nbins], frame_bins; entries below nmin are NaN."""
    lam = np.asarray(lambdas, np.float64)
    coeff = np.asarray(coeff, np.float64)
    _, _, dG = zwanzig(Vs, lam, kT, nskip)
    gaps = [np.asarray(V, np.float64)[nskip:] @ coeff for V in Vs]
    gapmin = min(g.min() for g in gaps)
    gapmax = max(g.max() for g in gaps)
    xint = (gapmax - gapmin) / nbins
    if frame_bins is None:
        frame_bins = [np.floor((g - gapmin) / xint).astype(np.int64) for g in gaps]
    n = len(Vs)
    ns = lam.shape[1]
    dGg = np.full((n, nbins), np.nan)
    dGv = np.full((n, nbins, ns), np.nan)
    c1 = np.full((n, nbins), np.nan)
    npts = np.zeros((n, nbins), np.int64)
    num = np.zeros(nbins)
    numc = np.zeros(nbins)
    numr = np.zeros(nbins)
    ptsum = np.zeros(nbins, np.int64)
    for i, V in enumerate(Vs):
        V = np.asarray(V, np.float64)[nskip:]
        eg, w1 = evb_ground(V, _window_h12(H12, i, nskip))
        r = np.asarray(rkl[i], np.float64)[nskip:] if rkl is not None else None
        veff = V @ lam[i]
        dvg = eg - veff
        dvv = V - veff[:, None]
        b = frame_bins[i]
        for k in range(nbins):
            sel = b == k
            c = int(sel.sum())
            npts[i, k] = c
            if c == 0 or c < nmin:
                continue
            av = dvg[sel].sum() / c
            dGg[i, k] = dG[i] + (-kT * np.log(np.sum(np.exp(-(dvg[sel] - av) / kT)) / c) + av)
            avv = dvv[sel].sum(0) / c
            dGv[i, k] = dG[i] + (-kT * np.log(np.exp(-(dvv[sel] - avv) / kT).sum(0) / c) + avv)
            c1[i, k] = w1[sel].sum() / c
            num[k] += dGg[i, k] * c
            numc[k] += c1[i, k] * c
            if r is not None:
                numr[k] += (r[sel].sum() / c) * c  # bin mean x points, as Part 3 accumulates
            ptsum[k] += c
    ok = ptsum >= max(nmin, 1)
    binned = np.where(ok, num / np.maximum(ptsum, 1), np.nan)
    binned_c1 = np.where(ok, numc / np.maximum(ptsum, 1), np.nan)
    binned_r = np.where(ok & (rkl is not None), numr / np.maximum(ptsum, 1), np.nan)
    centres = gapmin + (np.arange(nbins) + 0.5) * xint
    return dict(centres=centres, dGg=dGg, dGv=dGv, c1=c1, npts=npts, binned=binned, binned_c1=binned_c1,
                binned_r=binned_r, ptsum=ptsum, frame_bins=frame_bins)


def evb_pmf_grad(Vs, dVs, dH, lambdas, H12, kT, coeff, nbins, nmin, nskip=0, frame_bins=None):
    """Derivative of evb_pmf's Part 3 curve (nbins, nparam) for fixed frames
    and fixed bin membership; rows of unreported bins are NaN."""
    lam = np.asarray(lambdas, np.float64)
    dH = np.asarray(dH, np.float64)
    res = evb_pmf(Vs, lam, H12, kT, coeff, nbins, nmin, nskip, frame_bins)
    gdG = zwanzig_grad(Vs, dVs, lam, kT, nskip, per_window=True)
    npar = gdG.shape[1]
    num = np.zeros((nbins, npar))
    for i, (V, dV) in enumerate(zip(Vs, dVs)):
        V = np.asarray(V, np.float64)[nskip:]
        dV = np.asarray(dV, np.float64)[nskip:]
        dvg, ddvg = _ground_dv(V, dV, lam[i], H12, dH)
        b = res["frame_bins"][i]
        for k in range(nbins):
            c = res["npts"][i, k]
            if c == 0 or c < nmin:
                continue
            sel = b == k
            w = np.exp(-(dvg[sel] - dvg[sel].min()) / kT)
            num[k] += c * (gdG[i] + (w[:, None] * ddvg[sel]).sum(0) / w.sum())
    out = np.full((nbins, npar), np.nan)
    ok = res["ptsum"] >= max(nmin, 1)
    out[ok] = num[ok] / res["ptsum"][ok, None]
    return out
