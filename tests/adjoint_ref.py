"""Direct enumeration of the adjoint's per-frame sums (test helper).

For one frame, every Q pair the reference's Q lists hold (nbqplist/nbqwlist
membership from Rcq and xpcent, the static nbqq/nbqqp lists,
src/nonbondene.f90:3231-3944) is visited in plain numpy FP64 and the sums
the device kernel returns are formed from their definitions: for a table
entry T = p_iq * f (p the Q-side parameter, f its tangent) with radial
function g(r), the entry contributes f * g(r) to dE_s/dp_iq, with
g = 1/r (charge), 1/r^12 (A) and -1/r^6 (B) (nbe_qx, src/nonbonded.f90:200-222).
The tests pin these sums by finite differences of the oracle's energies.
"""
import numpy as np

ADJ_DQ, ADJ_DA, ADJ_DB, ADJ_DA14, ADJ_DB14, ADJ_EEL, ADJ_EVDW = range(7)


def _F(a, shape):
    return np.asarray(a).reshape(shape, order="F")


def slot_sums(s, x, tang):
    n, a, p = s.n, s.a, s.par
    nq, ns, nsol, nwat, sa = n["nqat"], n["nstates"], n["nat_solute"], n["nwat"], n["solv_atom"]
    out = np.zeros((ns, nq, 7))
    if nq == 0:
        return out
    iqseq, qconn3, excl = a["iqseq"], a["qconn3"], a["excl"]
    xp, rcq2 = np.asarray(p["xpcent"], float), p["Rcq"] ** 2
    spc = n.get("solvent_spc", 0)
    T = {k: _F(a[k], (max(nsol, 1), nq, ns)) for k in ("qp_A", "qp_B", "qp_el", "qp_sc", "qp_set")}
    G = {k: _F(tang[k], (max(nsol, 1), nq, ns)) for k in ("qp_dq", "qp_dA", "qp_dB", "qp_is14")}
    W = {k: _F(a[k], (nq, sa, ns)) for k in ("qw_A", "qw_B", "qw_el", "qw_sc")} if nwat else {}
    GW = {k: _F(tang[k], (nq, sa, ns)) for k in ("qw_dq", "qw_dA", "qw_dB")} if nwat else {}
    Q = {k: _F(a[k], (nq, nq, ns)) for k in ("qq_A", "qq_B", "qq_el", "qq_sc", "qq_set", "qq_soft")}
    GQ = {k: _F(tang[k], (nq, nq, ns)) for k in ("qq_dq_i", "qq_dA_i", "qq_dB_i", "qq_dq_j", "qq_dA_j", "qq_dB_j",
                                                "qq_is14")}

    def add(s_, iq, r, A, B, el, sc, tq, tA, tB, is14, lj=True, energy=True):
        rin = 1.0 / r
        out[s_, iq, ADJ_DQ] += tq * rin
        if energy:
            out[s_, iq, ADJ_EEL] += el * rin
        if lj:
            r6 = 1.0 / (r ** 6 + sc)
            out[s_, iq, ADJ_DA14 if is14 else ADJ_DA] += tA * r6 * r6
            out[s_, iq, ADJ_DB14 if is14 else ADJ_DB] -= tB * r6
            if energy:
                out[s_, iq, ADJ_EVDW] += A * r6 * r6 - B * r6

    # Q-list membership
    waters = [w for w in range(nwat) if not excl[nsol + sa * w]
              and ((x[nsol + sa * w] - xp) ** 2).sum() <= rcq2]
    grp = np.zeros(nsol, int)
    for g in range(n["ncgp_solute"]):
        for k in range(a["cgp_first"][g], a["cgp_last"][g] + 1):
            grp[a["cgpatom"][k]] = g
    sol = [j for j in range(nsol) if not qconn3[j] and not excl[a["cgp_iswitch"][grp[j]]]
           and ((x[a["cgp_iswitch"][grp[j]]] - xp) ** 2).sum() <= rcq2]
    for iq in range(nq):
        i = iqseq[iq]
        for w in waters:
            for k in range(sa):
                j = nsol + sa * w + k
                r = np.sqrt(((x[j] - x[i]) ** 2).sum())
                for s_ in range(ns):
                    add(s_, iq, r, W["qw_A"][iq, k, s_], W["qw_B"][iq, k, s_], W["qw_el"][iq, k, s_],
                        W["qw_sc"][iq, k, s_], GW["qw_dq"][iq, k, s_], GW["qw_dA"][iq, k, s_], GW["qw_dB"][iq, k, s_],
                        False, lj=(k == 0 or not spc))
        for j in sol:
            r = np.sqrt(((x[j] - x[i]) ** 2).sum())
            for s_ in range(ns):
                add(s_, iq, r, T["qp_A"][j, iq, s_], T["qp_B"][j, iq, s_], T["qp_el"][j, iq, s_],
                    T["qp_sc"][j, iq, s_], G["qp_dq"][j, iq, s_], G["qp_dA"][j, iq, s_], G["qp_dB"][j, iq, s_],
                    bool(G["qp_is14"][j, iq, s_]))
        # Q-bonded neighbours: non-Q atoms 1-2/1-3 bonded to some Q-atom
        for j in range(nsol):
            if a["iqatom"][j] or not qconn3[j]:
                continue
            for s_ in range(ns):
                if not T["qp_set"][j, iq, s_]:
                    continue
                r = np.sqrt(((x[j] - x[i]) ** 2).sum())
                add(s_, iq, r, T["qp_A"][j, iq, s_], T["qp_B"][j, iq, s_], T["qp_el"][j, iq, s_],
                    T["qp_sc"][j, iq, s_], G["qp_dq"][j, iq, s_], G["qp_dA"][j, iq, s_], G["qp_dB"][j, iq, s_],
                    bool(G["qp_is14"][j, iq, s_]))
    for s_ in range(ns):
        for iq in range(nq - 1):
            for jq in range(iq + 1, nq):
                if not Q["qq_set"][iq, jq, s_]:
                    continue
                assert not Q["qq_soft"][iq, jq, s_]
                r = np.sqrt(((x[iqseq[jq]] - x[iqseq[iq]]) ** 2).sum())
                args = (Q["qq_A"][iq, jq, s_], Q["qq_B"][iq, jq, s_], Q["qq_el"][iq, jq, s_], Q["qq_sc"][iq, jq, s_])
                is14 = bool(GQ["qq_is14"][iq, jq, s_])
                add(s_, iq, r, *args, GQ["qq_dq_i"][iq, jq, s_], GQ["qq_dA_i"][iq, jq, s_], GQ["qq_dB_i"][iq, jq, s_],
                    is14)
                add(s_, jq, r, *args, GQ["qq_dq_j"][iq, jq, s_], GQ["qq_dA_j"][iq, jq, s_], GQ["qq_dB_j"][iq, jq, s_],
                    is14, energy=False)
    return out
