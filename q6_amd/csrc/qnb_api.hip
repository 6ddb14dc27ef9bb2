// qnb_api.hip -- C ABI of the engine (include/qnb.h): handle, validation,
// device tables, launches and the reference-order export of pair lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <memory>

#include "qnb_host.h"

namespace qnb {
template <typename R, bool ANY>
__global__ void k_lists_lrf(const DevSys*);
__global__ void k_qlists(const DevSys*);
template <typename R, int NS, bool E, int CAP>
__global__ void k_wforces(const DevSys*, uint32_t, int, int, int);
template <typename R, int NS>
__global__ void k_gforces(const DevSys*, uint32_t, int, int, int);
template <typename R, int NS, bool E, int CAP>
__global__ void k_sforces(const DevSys*, uint32_t, int, int, int);
template <typename R, int NS>
__global__ void k_qforces(const DevSys*, uint32_t, int, double*, int);
template <typename R, int NS>
__global__ void k_qq(const DevSys*, uint32_t);
__global__ void k_pack_xq(const DevSys*);
__global__ void k_pw_first(const DevSys*);
__global__ void k_watpol_theta(const DevSys*, const int*);
__global__ void k_watpol_rank(const DevSys*);
__global__ void k_watpol_force(const DevSys*);
__global__ void k_watpol_finish(const DevSys*);

// SCAAS polarisation restraint: theta and shells, ranks, forces (blocks of
// WP_THREADS waters, grid.y the instance), then the per-system sums
void launch_watpol(hipStream_t st, int nblk, int B, const DevSys* dsys, const int* istep) {
  const dim3 g(nblk, B), t(WP_THREADS);
  hipLaunchKernelGGL(k_watpol_theta, g, t, 0, st, dsys, istep);
  hipLaunchKernelGGL(k_watpol_rank, dim3(QNB_MAXSH, B), dim3(WP_RANK_THREADS), 0, st, dsys);
  hipLaunchKernelGGL(k_watpol_force, g, t, 0, st, dsys);
  hipLaunchKernelGGL(k_watpol_finish, dim3(B), dim3(64), 0, st, dsys);
}
__global__ void k_reduce_energies(const DevSys*, int, int);
__global__ void k_atom_terms(const DevSys*, uint32_t, const double*);

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

const char* last_error() { return g_err.c_str(); }

int validate_system(const qnb_system* s, const qnb_params* p) {
  if (!s || !p) return fail(QNB_EINVAL, "null system or params");
  if (s->natom <= 0 || s->nat_solute < 0 || s->nat_solute > s->natom) return fail(QNB_EINVAL, "bad atom counts");
  if (s->nstates < 1 || s->nstates > QNB_MAX_STATES) return fail(QNB_EINVAL, "nstates must be 1..%d", QNB_MAX_STATES);
  if (s->nwat > 0 && (s->solv_atom < 1 || s->solv_atom > QNB_WAVE))
    return fail(QNB_EUNSUPPORTED, "solvent molecules of 1..%d atoms are supported (solv_atom=%d)", QNB_WAVE,
                s->solv_atom);
  if (s->n_solv_int < 0 || (s->n_solv_int > 0 && (!s->solv_int_ij || !s->solv_int_par)))
    return fail(QNB_EINVAL, "solvent internal pairs: count and arrays");
  for (int k = 0; k < s->n_solv_int; k++)
    for (int e = 0; e < 2; e++)
      if (s->solv_int_ij[2 * k + e] < 0 || s->solv_int_ij[2 * k + e] >= s->solv_atom)
        return fail(QNB_EINVAL, "solvent internal pair %d: atom offset %d out of range", k, s->solv_int_ij[2 * k + e]);
  if (s->nat_solute + s->nwat * s->solv_atom != s->natom) return fail(QNB_EINVAL, "natom != nat_solute + nwat*solv_atom");
  if (s->ncgp != s->ncgp_solute + s->nwat) return fail(QNB_EINVAL, "each solvent molecule must be one charge group");
  if (p->nwpolr_shell < 0 || p->nwpolr_shell > QNB_MAX_SHELLS) return fail(QNB_EINVAL, "nwpolr_shell out of range");
  if (s->ntypes <= 0) return fail(QNB_EINVAL, "no atom types");
  for (int i = 0; i < s->natom; i++)
    if (s->iac[i] < 0 || s->iac[i] >= s->ntypes) return fail(QNB_EINVAL, "atom %d has type %d out of range", i, s->iac[i]);
  for (int g = 0; g < s->ncgp; g++) {
    if (s->cgp_first[g] > s->cgp_last[g]) return fail(QNB_EINVAL, "empty charge group %d", g);
    if (s->cgp_last[g] - s->cgp_first[g] + 1 > FC_CAP)
      return fail(QNB_EUNSUPPORTED, "charge group %d has more than %d atoms", g, FC_CAP);
    if (g >= s->ncgp_solute) {
      int w = g - s->ncgp_solute, o = s->nat_solute + s->solv_atom * w;
      if (s->cgp_last[g] - s->cgp_first[g] + 1 != s->solv_atom || s->cgp_iswitch[g] != o)
        return fail(QNB_EINVAL, "solvent group %d is not molecule %d", g, w);
      for (int k = 0; k < s->solv_atom; k++)
        if (s->cgpatom[s->cgp_first[g] + k] != o + k) return fail(QNB_EINVAL, "solvent group %d atom order", g);
    } else {
      for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++)
        if (s->cgpatom[k] < 0 || s->cgpatom[k] >= s->nat_solute)
          return fail(QNB_EINVAL, "solute group %d holds a non-solute atom", g);
    }
  }
  for (int i = s->nat_solute; i < s->natom; i++)
    if (s->iqatom[i] != 0) return fail(QNB_EUNSUPPORTED, "Q-atoms in the solvent are not supported");
  for (int q = 0; q < s->nqat; q++)
    if (s->iqseq[q] < 0 || s->iqseq[q] >= s->nat_solute) return fail(QNB_EINVAL, "Q-atom %d is not a solute atom", q);
  return 0;
}

void fill_params(DevSys& S, const qnb_system* s, const qnb_params* p) {
  const int ns = s->nstates;
  S.natom = s->natom;
  S.nat_solute = s->nat_solute;
  S.nwat = s->nwat;
  S.sa = s->solv_atom;
  S.ncgp = s->ncgp;
  S.ncgp_solute = s->ncgp_solute;
  S.nqat = s->nqat;
  S.nstates = ns;
  S.ntypes = s->ntypes;
  S.ivdw_rule = s->ivdw_rule;
  S.any_atom = s->any_atom_cutoff ? 1 : 0;
  S.spc = s->solvent_spc;
  S.pw_coul_h = 0;
  if (s->nwat > 0 && s->nat_solute > 0) {
    bool zero = true;
    for (size_t t = (size_t)s->nat_solute; t < (size_t)s->nat_solute * s->solv_atom && zero; t++)
      zero = s->pw_A[t] == 0.0 && s->pw_B[t] == 0.0;
    S.pw_coul_h = zero ? 1 : 0;
  }
  S.use_lrf = p->use_LRF;
  S.prec = p->precision;
  S.quirk = p->ref_quirks;
  S.rcpp2 = p->Rcpp * p->Rcpp;
  S.rcpw2 = p->Rcpw * p->Rcpw;
  S.rcww2 = p->Rcww * p->Rcww;
  S.rcq2 = p->Rcq * p->Rcq;
  S.rclrf2 = p->RcLRF * p->RcLRF;
  S.rexcl2 = p->rexcl_o * p->rexcl_o;
  for (int k = 0; k < QNB_MAXST; k++) S.lambda[k] = k < ns ? p->lambda[k] : 0.0;
  S.el14 = s->el14_scale;
  for (int c = 0; c < 3; c++) {
    S.xp[c] = p->xpcent[c];
    S.xw[c] = p->xwcent[c];
  }
  S.rwat = p->rwat;
  S.fk_wsphere = p->fk_wsphere;
  S.Dwmz = p->Dwmz;
  S.awmz = p->awmz;
  S.Boltz = p->Boltz;
  S.fkwpol = p->fkwpol;
  S.itdis = p->itdis;
  S.wpol = p->wpol_restr;
  S.nshell = p->nwpolr_shell;
  for (int k = 0; k < QNB_MAXSH; k++) {
    S.sh_rout[k] = k < S.nshell ? p->wshell_rout[k] : 0.0;
    S.sh_dr[k] = k < S.nshell ? p->wshell_dr[k] : 0.0;
    S.sh_cstb[k] = k < S.nshell ? p->wshell_cstb[k] : 0.0;
  }
}

int topo_build(DevSys& S, const qnb_system* s, HostTopo& T, DevAlloc& M) {
  const int nat = s->natom, nsol = s->nat_solute, nq = s->nqat, ns = s->nstates, sa = s->solv_atom;
  T.ntypes = s->ntypes;
  T.ivdw_rule = s->ivdw_rule;
  T.gfirst.assign(s->cgp_first, s->cgp_first + s->ncgp);
  T.glast.assign(s->cgp_last, s->cgp_last + s->ncgp);
  T.gswitch.assign(s->cgp_iswitch, s->cgp_iswitch + s->ncgp);
  const int ncgpat = s->cgp_last[s->ncgp - 1] + 1;
  T.cgpatom.assign(s->cgpatom, s->cgpatom + ncgpat);
  T.iqatom.assign(s->iqatom, s->iqatom + nat);
  T.iac.assign(s->iac, s->iac + nat);
  T.excl.assign(s->excl, s->excl + nat);
  T.crg.assign(s->crg, s->crg + nat);
  T.avdw.assign(s->avdw, s->avdw + 3 * s->ntypes);
  T.bvdw.assign(s->bvdw, s->bvdw + 3 * s->ntypes);
  T.ljcod.assign(s->ljcod, s->ljcod + s->ntypes * s->ntypes);
  std::vector<int> atom_group(nat, -1);
  for (int g = 0; g < s->ncgp; g++)
    for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++) atom_group[s->cgpatom[k]] = g;
  // Qdyn6 leaves excluded solute atoms out of every charge group when a
  // topology is cut to a sphere; no pair list reaches them, and the per-atom
  // pass, not a pair kernel, is the first writer of their gradient
  for (int i = 0; i < nat; i++)
    if (atom_group[i] < 0 && !(i < nsol && s->excl[i] && s->iqatom[i] == 0))
      return fail(QNB_EINVAL, "atom %d is in no charge group", i);
  // charge groups are consecutive runs of cgpatom, as the topology reader
  // lays them out
  std::vector<int> gslot(ncgpat);
  for (int g = 0; g < s->ncgp; g++) {
    if (s->cgp_first[g] != (g == 0 ? 0 : s->cgp_last[g - 1] + 1) || s->cgp_last[g] < s->cgp_first[g])
      return fail(QNB_EINVAL, "charge group %d is not the run after group %d in cgpatom", g, g - 1);
    for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++) gslot[k] = g & 63;
  }
  S.ncgpat = ncgpat;
  const int nsol1 = std::max(nsol, 1);
  T.exmask.assign(nsol1, 0u);
  T.m14mask.assign(nsol1, 0u);
  for (int i = 0; i < nsol; i++)
    for (int k = 1; k <= 25; k++) {
      if (s->listex[i * 25 + k - 1]) T.exmask[i] |= 1u << (k - 1);
      if (s->list14[i * 25 + k - 1]) T.m14mask[i] |= 1u << (k - 1);
    }
  {
    std::vector<std::vector<std::pair<int, int>>> lx(nsol1);
    for (int m = 0; m < s->nexlong; m++) {
      int a = s->listexlong[2 * m], b = s->listexlong[2 * m + 1];
      lx[std::min(a, b)].push_back({std::max(a, b), 0});
    }
    for (int m = 0; m < s->n14long; m++) {
      int a = s->list14long[2 * m], b = s->list14long[2 * m + 1];
      lx[std::min(a, b)].push_back({std::max(a, b), 1});
    }
    T.lx_off.assign(nsol1 + 1, 0);
    for (int i = 0; i < nsol1; i++) {
      // bit 31 of the exclusion mask: the atom has partners beyond the 25-atom window
      if (!lx[i].empty()) T.exmask[i] |= 1u << 31;
      T.lx_off[i + 1] = T.lx_off[i] + (int)lx[i].size();
      for (auto& e : lx[i]) {
        T.lx_j.push_back(e.first);
        T.lx_kind.push_back((uint8_t)e.second);
      }
    }
  }
  if (nq > 0) T.iqseq.assign(s->iqseq, s->iqseq + nq);
  if (nsol > 0 && nq > 0) T.qconn3.assign(s->qconn3, s->qconn3 + nsol);
  else T.qconn3.assign(nsol1, 0);
  // static Q lists (nbqqlist, nonbondene.f90:3231-3282), per state in reference order
  std::vector<int> qq_iq, qq_jq, qq_st, qqp_iq, qqp_j, qqp_st;
  std::vector<uint8_t> qq_soft;
  std::vector<double> qq_par, qqp_par;
  for (int st = 0; st < ns && nq > 0; st++) {
    for (int iq = 0; iq < nq - 1; iq++)
      for (int jq = iq + 1; jq < nq; jq++) {
        size_t k = (size_t)iq + nq * ((size_t)jq + (size_t)nq * st);
        if (!s->qq_set[k]) continue;
        qq_iq.push_back(iq);
        qq_jq.push_back(jq);
        qq_st.push_back(st);
        qq_soft.push_back(s->qq_soft[k]);
        qq_par.insert(qq_par.end(), {s->qq_A[k], s->qq_B[k], s->qq_el[k], s->qq_sc[k]});
        if (st == 0) {
          T.qq_iq.push_back(iq);
          T.qq_jq.push_back(jq);
        }
      }
    for (int ja = 0; ja < nsol; ja++) {
      if (s->iqatom[ja] != 0 || !s->qconn3[ja]) continue;
      for (int iq = 0; iq < nq; iq++) {
        size_t k = (size_t)ja + nsol * ((size_t)iq + (size_t)nq * st);
        if (!s->qp_set[k]) continue;
        qqp_iq.push_back(iq);
        qqp_j.push_back(ja);
        qqp_st.push_back(st);
        qqp_par.insert(qqp_par.end(), {s->qp_A[k], s->qp_B[k], s->qp_el[k], s->qp_sc[k]});
        if (st == 0) {
          T.qqp_iq.push_back(iq);
          T.qqp_j.push_back(ja);
        }
      }
    }
  }
  int rc = 0;
#define UP(field, src, n)                                    \
  do {                                                       \
    if ((rc = M.upload(&S.field, (src), (size_t)(n)))) return rc; \
  } while (0)
  UP(crg, s->crg, nat);
  UP(iac, s->iac, nat);
  UP(iqatom, s->iqatom, nat);
  UP(excl, s->excl, nat);
  UP(heavy, s->heavy, nat);
  UP(gfirst, s->cgp_first, s->ncgp);
  UP(glast, s->cgp_last, s->ncgp);
  UP(gswitch, s->cgp_iswitch, s->ncgp);
  UP(cgpatom, s->cgpatom, ncgpat);
  UP(atom_group, atom_group.data(), nat);
  UP(gslot, gslot.data(), ncgpat);

  UP(avdw, s->avdw, 3 * s->ntypes);
  UP(bvdw, s->bvdw, 3 * s->ntypes);
  UP(ljcod, s->ljcod, s->ntypes * s->ntypes);
  UP(exmask, T.exmask.data(), T.exmask.size());
  UP(m14mask, T.m14mask.data(), T.m14mask.size());
  {
    // fast-path data of the solute receivers (DevSys::gspan ... ppf)
    const int ncs1 = std::max(s->ncgp_solute, 1);
    std::vector<int> span(2 * ncs1, 0);
    std::vector<uint8_t> glong(ncs1, 0);
    for (int g = 0; g < s->ncgp_solute; g++) {
      int lo = nat, hi = -1;
      for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++) {
        const int i = s->cgpatom[k];
        lo = std::min(lo, i);
        hi = std::max(hi, i);
        if (i < nsol && (T.exmask[i] >> 31)) glong[g] = 1;
      }
      span[2 * g] = lo;
      span[2 * g + 1] = hi;
    }
    UP(gspan, span.data(), span.size());
    UP(glong, glong.data(), glong.size());
    bool code1 = s->ivdw_rule == 1;
    for (int k = 0; k < s->ntypes * s->ntypes && code1; k++) code1 = s->ljcod[k] == 1;
    S.ppfast = code1 ? 1 : 0;
    std::vector<double> pd(4 * (size_t)nsol1, 0.0);
    std::vector<float> pf(4 * (size_t)nsol1, 0.0f);
    for (int i = 0; i < nsol; i++) {
      const int t = s->iac[i];
      pd[4 * i] = s->avdw[3 * t];
      pd[4 * i + 1] = s->bvdw[3 * t];
      pd[4 * i + 2] = s->crg[i];
      for (int c = 0; c < 3; c++) pf[4 * i + c] = (float)pd[4 * i + c];
    }
    UP(ppd, pd.data(), pd.size());
    UP(ppf, pf.data(), pf.size());
  }
  UP(lx_off, T.lx_off.data(), T.lx_off.size());
  UP(lx_j, T.lx_j.data(), T.lx_j.size());
  UP(lx_kind, T.lx_kind.data(), T.lx_kind.size());
  if (nsol > 0 && s->nwat > 0) {
    UP(pwA, s->pw_A, (size_t)nsol * sa);
    UP(pwB, s->pw_B, (size_t)nsol * sa);
    UP(pwEl, s->pw_el, (size_t)nsol * sa);
    if (S.prec == QNB_PREC_MIXED && sa == 3) {
      std::vector<float> f((size_t)nsol * PW_REC, 0.0f);
      for (int i = 0; i < nsol; i++)
        for (int k = 0; k < 3; k++) {
          const size_t t = (size_t)i + (size_t)nsol * k;
          f[(size_t)PW_REC * i + 3 * k] = (float)s->pw_A[t];
          f[(size_t)PW_REC * i + 3 * k + 1] = (float)s->pw_B[t];
          f[(size_t)PW_REC * i + 3 * k + 2] = (float)s->pw_el[t];
        }
      UP(pw_f, f.data(), f.size());
    }
  }
  if (s->nwat > 0) {
    UP(wwA, s->ww_A, sa * sa);
    UP(wwB, s->ww_B, sa * sa);
    UP(wwEl, s->ww_el, sa * sa);
    S.n_sint = s->n_solv_int;
    if (S.n_sint > 0) {
      UP(sint_ij, s->solv_int_ij, 2 * (size_t)S.n_sint);
      UP(sint_par, s->solv_int_par, 3 * (size_t)S.n_sint);
    }
  }
  UP(qconn3, T.qconn3.data(), T.qconn3.size());
  if (nq > 0) {
    UP(iqseq, s->iqseq, nq);
    if (nsol > 0) {
      size_t n = (size_t)nsol * nq * ns;
      UP(qpA, s->qp_A, n);
      UP(qpB, s->qp_B, n);
      UP(qpEl, s->qp_el, n);
      UP(qpSc, s->qp_sc, n);
      if (S.prec == QNB_PREC_MIXED) {
        std::vector<float> f(n);
        const double* src[4] = {s->qp_A, s->qp_B, s->qp_el, s->qp_sc};
        const float** dst[4] = {&S.qpA_f, &S.qpB_f, &S.qpEl_f, &S.qpSc_f};
        for (int k = 0; k < 4; k++) {
          for (size_t m = 0; m < n; m++) f[m] = (float)src[k][m];
          if ((rc = M.upload(dst[k], f.data(), n))) return rc;
        }
      }
    }
    if (s->nwat > 0) {
      size_t n = (size_t)nq * sa * ns;
      UP(qwA, s->qw_A, n);
      UP(qwB, s->qw_B, n);
      UP(qwEl, s->qw_el, n);
      UP(qwSc, s->qw_sc, n);
    }
    S.nqq = (int)qq_iq.size();
    UP(qq_iq, qq_iq.data(), qq_iq.size());
    UP(qq_jq, qq_jq.data(), qq_jq.size());
    UP(qq_st, qq_st.data(), qq_st.size());
    UP(qq_soft, qq_soft.data(), qq_soft.size());
    UP(qq_par, qq_par.data(), qq_par.size());
    S.nqqp = (int)qqp_iq.size();
    UP(qqp_iq, qqp_iq.data(), qqp_iq.size());
    UP(qqp_j, qqp_j.data(), qqp_j.size());
    UP(qqp_st, qqp_st.data(), qqp_st.size());
    UP(qqp_par, qqp_par.data(), qqp_par.size());
    // atom-centred view of both lists for k_qq: entities are the Q-atoms and
    // the distinct solute partners of the Q-bonded-neighbour pairs; entity e
    // is atom qe_atom[e] and owns items qe_off[e]..qe_off[e+1]-1 of qe_item,
    // each {first atom, second atom, pair, (state << 2) | (kind << 1) | side},
    // kind 0 a Q-Q pair and 1 a Q-neighbour pair, side 0 the pair's first
    // atom (which counts its energy)
    std::vector<int> ent_atom(nq), cnt(nq, 0);
    std::vector<int> ent_of_j(nsol1, -1);
    for (int iq = 0; iq < nq; iq++) ent_atom[iq] = s->iqseq[iq];
    for (size_t p = 0; p < qqp_j.size(); p++)
      if (ent_of_j[qqp_j[p]] < 0) {
        ent_of_j[qqp_j[p]] = (int)ent_atom.size();
        ent_atom.push_back(qqp_j[p]);
        cnt.push_back(0);
      }
    for (size_t p = 0; p < qq_iq.size(); p++) cnt[qq_iq[p]]++, cnt[qq_jq[p]]++;
    for (size_t p = 0; p < qqp_iq.size(); p++) cnt[qqp_iq[p]]++, cnt[ent_of_j[qqp_j[p]]]++;
    const int nent = (int)ent_atom.size();
    std::vector<int> off(nent + 1, 0);
    for (int e = 0; e < nent; e++) off[e + 1] = off[e] + cnt[e];
    std::vector<int> item(4 * (size_t)std::max(off[nent], 1)), pos(off.begin(), off.end() - 1);
    auto put = [&](int e, int a, int b, int p, int w) {
      const int q = pos[e]++;
      item[4 * q] = a;
      item[4 * q + 1] = b;
      item[4 * q + 2] = p;
      item[4 * q + 3] = w;
    };
    for (size_t p = 0; p < qq_iq.size(); p++) {
      const int a = s->iqseq[qq_iq[p]], b = s->iqseq[qq_jq[p]], st = qq_st[p] << 2;
      put(qq_iq[p], a, b, (int)p, st);
      put(qq_jq[p], a, b, (int)p, st | 1);
    }
    for (size_t p = 0; p < qqp_iq.size(); p++) {
      const int a = s->iqseq[qqp_iq[p]], b = qqp_j[p], st = qqp_st[p] << 2;
      put(qqp_iq[p], a, b, (int)p, st | 2);
      put(ent_of_j[qqp_j[p]], a, b, (int)p, st | 3);
    }
    S.nqe = nent;
    UP(qe_atom, ent_atom.data(), ent_atom.size());
    UP(qe_off, off.data(), off.size());
    UP(qe_item, item.data(), item.size());
  }
#undef UP
  return 0;
}

int instance_alloc(DevSys& S, DevAlloc& M, const qnb_params* p, int nslots) {
  int rc = 0;
#define AL(field, n)                                          \
  do {                                                        \
    if ((rc = M.alloc(&S.field, (size_t)(n)))) return rc;     \
  } while (0)
  const int nat = S.natom, nsol1 = std::max(S.nat_solute, 1), nw1 = std::max(S.nwat, 1);
  AL(x, 3 * nat);
  if (p->precision == QNB_PREC_MIXED) AL(xqg, 4 * (size_t)S.ncgpat);
  else AL(xqgd, 4 * (size_t)S.ncgpat);
  AL(gref, 4 * (size_t)S.ncgp);
  AL(d, 3 * nat);
  AL(lrf, (size_t)QNB_NLRF * S.ncgp);
  if (S.cap <= 0) S.cap = 256;
  AL(nbr, (size_t)S.cap * S.ncgp);
  AL(nnbr, S.ncgp);
  AL(pw_first, std::max(S.ncgp_solute, 1));
  AL(need, 1);
  AL(qw_in, nw1);
  AL(qp_in, std::max(S.ncgp_solute, 1));
  AL(qw_list, nw1);
  AL(qp_list, nsol1);
  AL(nq_lists, 2);
  S.nslots = nslots;
  AL(epart, (size_t)QNB_NE * nslots);
  AL(ered, QNB_NE);
  AL(eatom, 2 * (size_t)nat);
  AL(sh_state, 3 * QNB_MAXSH);
  AL(w_theta, nw1);
  AL(w_shell, nw1);
  AL(w_rank, nw1);
  AL(sh_list, (size_t)QNB_MAXSH * nw1);
  AL(sh_lth, (size_t)QNB_MAXSH * nw1);
  AL(sh_n, QNB_MAXSH);
  AL(wp_part, (size_t)(1 + QNB_MAXSH) * ((nw1 + WP_THREADS - 1) / WP_THREADS));

#undef AL
  double sh0[3 * QNB_MAXSH] = {0};
  for (int k = 0; k < S.nshell; k++) sh0[k] = p->wshell_theta_corr[k];
  HIPCHK(hipMemcpy(S.sh_state, sh0, sizeof sh0, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(S.nq_lists, 0, 2 * sizeof(int)));
  HIPCHK(hipMemset(S.need, 0, sizeof(int)));
  HIPCHK(hipMemset(S.sh_n, 0, sizeof(int) * QNB_MAXSH));
  HIPCHK(hipMemset(S.nnbr, 0, sizeof(int) * S.ncgp));
  HIPCHK(hipMemset(S.ered, 0, sizeof(double) * QNB_NE));
  return 0;
}

}  // namespace qnb

using namespace qnb;

// ---------------------------------------------------------------------------
// Fixed device workspace of qnb_adjoint_static (callers that cannot
// allocate): the tangent tables, a chunk of frames, the kernel's parameter
// record and the chunk's output are staged here.  Each GPU holds its own copy
// for the life of the process.  In host code `scratch` only names the symbol:
// its device address on each GPU comes from hipGetSymbolAddress, looked up
// once per device when a handle is created there (qnb_create), and only that
// address is used.  Every call on one GPU shares the array; calls that overlap
// on one GPU race.
static __device__ __align__(256) unsigned char scratch[QNB_ADJOINT_STATIC_BYTES];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device `dev` (cached per device)
static int scratch_lookup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(QNB_EDEVICE, "device %d beyond the workspace table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(QNB_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
  g_scratch_ptr[dev] = (unsigned char*)ptr;
  return 0;
}

struct qnb_handle {
  int device = 0;  // the device current at qnb_create
  bool small_groups = true;  // every solute group fits the water kernel's small staging
  DevSys hs;  // host copy of the device record (device pointers)
  DevSys* dsys = nullptr;
  DevAlloc mem;
  HostTopo T;
  hipStream_t stream = nullptr;
  double* d_tfree = nullptr;
  int* d_istep = nullptr;
  bool qlists_built = false;
  bool lists_built = false;
  QbDev qb{};                  // Q bonded terms (host copy of the device record)
  QbDev* dqb = nullptr;
  std::vector<void*> qb_owned;
  ExcDev exc{};                         // group-contribution arrays (host copy of device pointers)
  std::unique_ptr<DevAlloc> exc_mem;    // their lists, freed when the arrays are replaced
};

// host twin of the device pp_param filter, for list export
bool qnb::host_pp_param(const HostTopo& T, int i, int j) {
  int lo = std::min(i, j), hi = std::max(i, j), dk = hi - lo;
  bool is14 = false;
  if (dk <= 25) {
    uint32_t bit = 1u << (dk - 1);
    if (T.exmask[lo] & bit) return false;
    is14 = (T.m14mask[lo] & bit) != 0;
  } else {
    bool ex = false;
    for (int m = T.lx_off[lo]; m < T.lx_off[lo + 1]; m++)
      if (T.lx_j[m] == hi) {
        if (T.lx_kind[m] == 0) ex = true;
        else is14 = true;
      }
    if (ex) return false;
  }
  int ti = T.iac[i], tj = T.iac[j];
  int vdw = is14 ? 3 : T.ljcod[ti + tj * T.ntypes];
  int k = vdw - 1;
  double A, B;
  if (T.ivdw_rule == 1) {
    A = T.avdw[ti * 3 + k] * T.avdw[tj * 3 + k];
    B = T.bvdw[ti * 3 + k] * T.bvdw[tj * 3 + k];
  } else {
    double tA = T.avdw[ti * 3 + k] + T.avdw[tj * 3 + k];
    tA = tA * tA;
    tA = tA * tA * tA;
    double tB = T.bvdw[ti * 3 + k] * T.bvdw[tj * 3 + k];
    A = (tA * tA) * tB;
    B = 2.0 * tA * tB;
  }
  double el = T.crg[i] * T.crg[j];
  return std::fabs(A) > 1.0e-10 || std::fabs(B) > 1.0e-10 || std::fabs(el) > 1.0e-10;
}

extern "C" int32_t qnb_abi_version(void) { return QNB_ABI_VERSION; }
extern "C" const char* qnb_last_error(void) { return qnb::last_error(); }
extern "C" int32_t qnb_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int32_t qnb_destroy(qnb_handle* h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

extern "C" int32_t qnb_create(const qnb_system* s, const qnb_params* p, qnb_handle** out) {
  if (!out) return fail(QNB_EINVAL, "null out");
  *out = nullptr;
  int rc = validate_system(s, p);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(QNB_EDEVICE, "no HIP device");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(QNB_EDEVICE, "no current HIP device");
  if ((rc = scratch_lookup(dev))) return rc;
  qnb_handle* h = new qnb_handle();
  h->device = dev;
  h->small_groups = max_solute_group(s->cgp_first, s->cgp_last, s->ncgp_solute) <= WF_CAP_SMALL;
  DevSys& S = h->hs;
  memset(&S, 0, sizeof S);
  fill_params(S, s, p);
  if ((rc = topo_build(S, s, h->T, h->mem)) || (rc = instance_alloc(S, h->mem, p, s->ncgp + 2)) ||
      (rc = h->mem.alloc(&h->d_tfree, 1)) || (rc = h->mem.alloc(&h->d_istep, 1)) || (rc = h->mem.alloc(&h->dsys, 1))) {
    delete h;
    return rc;
  }
  if (hipStreamCreate(&h->stream) != hipSuccess || hipMemcpy(h->dsys, &S, sizeof S, hipMemcpyHostToDevice) != hipSuccess) {
    delete h;
    return fail(QNB_EDEVICE, "stream or system record set-up failed");
  }
  *out = h;
  return 0;
}

static int upload_x(qnb_handle* h, const double* x) {
  if (!x) return fail(QNB_EINVAL, "null coordinates");
  HIPCHK(hipMemcpyAsync(h->hs.x, x, sizeof(double) * 3 * h->hs.natom, hipMemcpyHostToDevice, h->stream));
  return 0;
}

template <typename R>
static void launch_lists(qnb_handle* h) {
  const int wpb = 4;
  dim3 grid((h->hs.ncgp + wpb - 1) / wpb, 1);
  hipLaunchKernelGGL(k_pack_xq, dim3((h->hs.ncgpat + 255) / 256, 1), dim3(256), 0, h->stream, h->dsys);
  if (pw_quirk(h->hs))
    hipLaunchKernelGGL(k_pw_first, dim3((h->hs.ncgp_solute + 3) / 4, 1), dim3(256), 0, h->stream, h->dsys);
  if (h->hs.any_atom) hipLaunchKernelGGL((k_lists_lrf<R, true>), grid, dim3(64 * wpb), 0, h->stream, h->dsys);
  else hipLaunchKernelGGL((k_lists_lrf<R, false>), grid, dim3(64 * wpb), 0, h->stream, h->dsys);
}

extern "C" int32_t qnb_make_pair_lists(qnb_handle* h, const double* x) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  int rc = upload_x(h, x);
  if (rc) return rc;
  for (int attempt = 0; attempt < 4; attempt++) {
    HIPCHK(hipMemsetAsync(h->hs.need, 0, sizeof(int), h->stream));
    if (h->hs.prec == QNB_PREC_MIXED) launch_lists<float>(h);
    else launch_lists<double>(h);
    HIPCHK(hipGetLastError());
    int need = 0;
    HIPCHK(hipMemcpyAsync(&need, h->hs.need, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (need <= h->hs.cap) break;
    // reallocate_nonbondlist_* (qalloc.f90:370-605): grow and rebuild
    int ncap = ((need * 5 / 4) + 63) / 64 * 64;
    h->mem.release(h->hs.nbr);
    h->hs.nbr = nullptr;
    rc = h->mem.alloc(&h->hs.nbr, (size_t)ncap * h->hs.ncgp);
    if (rc) return rc;
    h->hs.cap = ncap;
    HIPCHK(hipMemcpy(h->dsys, &h->hs, sizeof(DevSys), hipMemcpyHostToDevice));
    if (attempt == 3) return fail(QNB_EOVERFLOW, "neighbour list capacity");
  }
  // Q lists: rebuilt every update unless Rcq exceeds the exclusion radius,
  // then built once (the `save`d list_done of nbqplist/nbqwlist)
  if (h->hs.nqat > 0 && !(h->qlists_built && h->hs.rcq2 > h->hs.rexcl2)) {
    hipLaunchKernelGGL(k_qlists, dim3(1), dim3(1024), 0, h->stream, h->dsys);
    HIPCHK(hipGetLastError());
    h->qlists_built = true;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  h->lists_built = true;
  return 0;
}

template <typename R, int NS>
static void launch_eval(qnb_handle* h, uint32_t mask) {
  const DevSys& S = h->hs;
  const int wpb = 4;
  if (S.nwat > 0 && S.sa != 3) {
    const int nb = (S.nwat + FC_WPB - 1) / FC_WPB;
    hipLaunchKernelGGL((k_gforces<R, NS>), dim3(nb, 1), dim3(64 * FC_WPB), 0, h->stream, h->dsys, mask, 0, nb, 1);
  } else if (S.nwat > 0) {
    const int nb = (S.nwat + wpb - 1) / wpb;
    if (h->small_groups)
      hipLaunchKernelGGL((k_wforces<R, NS, true, WF_CAP_SMALL>), dim3(nb, 1), dim3(64 * wpb),
                         (wf_wave_bytes<R, WF_CAP_SMALL>() * wpb), h->stream, h->dsys, mask, 0, nb, 1);
    else
      hipLaunchKernelGGL((k_wforces<R, NS, true, FC_CAP>), dim3(nb, 1), dim3(64 * wpb),
                         (wf_wave_bytes<R, FC_CAP>() * wpb), h->stream, h->dsys, mask, 0, nb, 1);
  }
  if (S.ncgp_solute > 0) {
    const int nb = (S.ncgp_solute + wpb - 1) / wpb;
    if (h->small_groups)
      hipLaunchKernelGGL((k_sforces<R, NS, true, WF_CAP_SMALL>), dim3(nb, 1), dim3(64 * wpb),
                         (sf_wave_bytes<R, WF_CAP_SMALL>() * wpb), h->stream, h->dsys, mask, 0, nb, 1);
    else
      hipLaunchKernelGGL((k_sforces<R, NS, true, FC_CAP>), dim3(nb, 1), dim3(64 * wpb), (sf_wave_bytes<R, FC_CAP>() * wpb),
                         h->stream, h->dsys, mask, 0, nb, 1);
  }
  hipLaunchKernelGGL(k_atom_terms, dim3((S.natom + 255) / 256, 1), dim3(256), 0, h->stream, h->dsys, mask,
                     (const double*)h->d_tfree);
  if (S.nqat > 0 && (mask & (QNB_T_QW | QNB_T_QP))) {
    // partners per Q-atom: at most nwat + nat_solute; about QF_PART / 64 per lane
    int maxp = S.nwat + S.nat_solute;
    int chunks = std::max(1, std::min((maxp + QF_PART - 1) / QF_PART, 64));
    hipLaunchKernelGGL((k_qforces<R, NS>), dim3(qforces_blocks(S.nqat, chunks, wpb), 1), dim3(64 * wpb), 0,
                       h->stream, h->dsys, mask, chunks, (double*)nullptr, 0);
  }
  const int with_qq = S.nqat > 0 && (mask & (QNB_T_QQ | QNB_T_QQP));
  const int with_pol = S.nwat > 0 && S.wpol && S.nshell > 0 && (mask & QNB_T_POL);
  if (with_qq) hipLaunchKernelGGL((k_qq<R, NS>), dim3(1), dim3(QQ_THREADS), 0, h->stream, h->dsys, mask);
  if (with_pol) launch_watpol(h->stream, (S.nwat + WP_THREADS - 1) / WP_THREADS, 1, h->dsys, (const int*)h->d_istep);
  hipLaunchKernelGGL(k_reduce_energies, dim3(1), dim3(1024), 0, h->stream, h->dsys, with_qq, with_pol);
}

template <typename R>
static void launch_eval_r(qnb_handle* h, uint32_t mask) {
  int ns = h->hs.nstates;
  if (ns <= 1) launch_eval<R, 1>(h, mask);
  else if (ns <= 2) launch_eval<R, 2>(h, mask);
  else if (ns <= 4) launch_eval<R, 4>(h, mask);
  else launch_eval<R, 8>(h, mask);
}

extern "C" int32_t qnb_eval_terms(qnb_handle* h, uint32_t mask, const double* x, double* d, double Tfree, int32_t istep,
                       qnb_energies* E, qnb_qenergies* EQ) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  if (!h->lists_built) return fail(QNB_EINVAL, "make_pair_lists must run before the first evaluation");
  if (!d) return fail(QNB_EINVAL, "null gradient");
  int rc = upload_x(h, x);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h->d_tfree, &Tfree, sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->d_istep, &istep, sizeof(int), hipMemcpyHostToDevice, h->stream));
  if (h->hs.prec == QNB_PREC_MIXED) launch_eval_r<float>(h, mask);
  else launch_eval_r<double>(h, mask);
  HIPCHK(hipGetLastError());
  const int nat = h->hs.natom;
  std::vector<double> dd(3 * (size_t)nat);
  double er[QNB_NE];
  HIPCHK(hipMemcpyAsync(dd.data(), h->hs.d, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(er, h->hs.ered, sizeof er, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < dd.size(); k++) d[k] += dd[k];
  bool finite = true;
  for (int k = 0; k < QNB_NE; k++) finite = finite && std::isfinite(er[k]);
  if (E) {
    if (mask & QNB_T_PP) { E->pp_el = er[E_PP_EL]; E->pp_vdw = er[E_PP_VDW]; }
    if (mask & QNB_T_PW) { E->pw_el = er[E_PW_EL]; E->pw_vdw = er[E_PW_VDW]; }
    if (mask & (QNB_T_WW | QNB_T_SINT)) { E->ww_el = er[E_WW_EL]; E->ww_vdw = er[E_WW_VDW]; }
    if (mask & QNB_T_LRF) E->lrf = er[E_LRF];
    if (mask & QNB_T_RAD) E->solvent_radial = er[E_RAD];
    if (mask & QNB_T_POL) E->water_pol = er[E_POL];
  }
  if (EQ)
    for (int s = 0; s < h->hs.nstates; s++) {
      const double* q = er + E_Q0 + 6 * s;
      if (mask & QNB_T_QQ) { EQ[s].qq_el = q[0]; EQ[s].qq_vdw = q[1]; }
      // nonbond_qp and nonbond_qqp both add into EQ(s)%qp: with QP selected the
      // field is set (Q-solute plus, when selected, Q-bonded-neighbour terms);
      // with QQP alone the Q-bonded-neighbour energy is added to it
      if (mask & QNB_T_QP) {
        EQ[s].qp_el = q[2];
        EQ[s].qp_vdw = q[3];
      } else if (mask & QNB_T_QQP) {
        EQ[s].qp_el += q[2];
        EQ[s].qp_vdw += q[3];
      }
      if (mask & QNB_T_QW) { EQ[s].qw_el = q[4]; EQ[s].qw_vdw = q[5]; }
    }
  if (!finite) return fail(QNB_ENONFINITE, "non-finite energy");
  return 0;
}

extern "C" int32_t qnb_set_qbonded(qnb_handle* h, const qnb_qbonded* qb) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  qbonded_free(h->mem, h->qb_owned);
  int rc = qbonded_build(qb, h->hs.nstates, h->hs.natom, h->hs.lambda, h->mem, h->qb_owned, h->qb);
  if (rc) {
    qbonded_free(h->mem, h->qb_owned);
    h->qb = QbDev{};
  }
  if (!h->dqb) {
    int rc2 = h->mem.alloc(&h->dqb, 1);
    if (rc2) return rc2;
  }
  HIPCHK(hipMemcpy(h->dqb, &h->qb, sizeof(QbDev), hipMemcpyHostToDevice));
  return rc;
}

extern "C" int32_t qnb_eval_qbonded(qnb_handle* h, const double* x, double* d, double* eq) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  if (!d || !eq) return fail(QNB_EINVAL, "null gradient or energies");
  const int ns = h->hs.nstates, nat = h->hs.natom;
  for (int k = 0; k < 4 * ns; k++) eq[k] = 0.0;
  if (!h->dqb || h->qb.nb + h->qb.na + h->qb.nt + h->qb.ni == 0) return 0;
  int rc = upload_x(h, x);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(h->hs.d, 0, sizeof(double) * 3 * nat, h->stream));
  launch_qbonded(h->stream, 1, h->dsys, h->dqb);
  HIPCHK(hipGetLastError());
  std::vector<double> dd(3 * (size_t)nat);
  HIPCHK(hipMemcpyAsync(dd.data(), h->hs.d, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(eq, h->qb.e, sizeof(double) * 4 * ns, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < dd.size(); k++) d[k] += dd[k];
  for (int k = 0; k < 4 * ns; k++)
    if (!std::isfinite(eq[k])) return fail(QNB_ENONFINITE, "non-finite Q bonded energy");
  return 0;
}

extern "C" int32_t qnb_set_groups(qnb_handle* h, const qnb_system* sys, int32_t narr, const uint8_t* mask,
                                  const int32_t* type) {
  if (!h || !sys) return fail(QNB_EINVAL, "null handle or system");
  if (sys->natom != h->hs.natom || sys->nat_solute != h->hs.nat_solute || sys->nqat != h->hs.nqat ||
      sys->nstates != h->hs.nstates)
    return fail(QNB_EINVAL, "groups: not the system the engine was created from");
  ExcHost X;
  if (int rc = exc_build(sys, narr, mask, type, X)) return rc;
  auto mem = std::make_unique<DevAlloc>();
  ExcDev D{};
  if (int rc = exc_upload(X, D, *mem)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  h->exc = D;
  h->exc_mem = std::move(mem);
  return 0;
}

extern "C" int32_t qnb_calculate_exc(qnb_handle* h, const double* x, const double* eq_in, double* out) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  const int n = h->exc.narr * h->exc.ns;
  if (n == 0) return 0;
  if (!x || !eq_in || !out) return fail(QNB_EINVAL, "null coordinates, energies or output");
  if (int rc = upload_x(h, x)) return rc;
  DevAlloc tmp;
  double *de = nullptr, *dout = nullptr;
  if (int rc = tmp.alloc(&de, (size_t)11 * h->exc.ns)) return rc;
  if (int rc = tmp.alloc(&dout, (size_t)QNB_GC_NV * n)) return rc;
  HIPCHK(hipMemcpyAsync(de, eq_in, sizeof(double) * 11 * h->exc.ns, hipMemcpyHostToDevice, h->stream));
  launch_exc_eval(h->stream, h->exc, h->hs.x, de, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * QNB_GC_NV * n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int32_t qnb_nonbonded(qnb_handle* h, const double* x, double* d, qnb_energies* E, qnb_qenergies* EQ) {
  return qnb_eval_terms(h, QNB_T_NONBONDED, x, d, 0.0, 0, E, EQ);
}

extern "C" int32_t qnb_qq(qnb_handle* h, const double* x, double* d, qnb_qenergies* EQ) {
  if (!EQ) return fail(QNB_EINVAL, "null EQ");
  return qnb_eval_terms(h, QNB_T_QQ | QNB_T_QQP, x, d, 0.0, 0, nullptr, EQ);
}

extern "C" int32_t qnb_restraints(qnb_handle* h, const double* x, double* d, double Tfree, int32_t istep, qnb_energies* E) {
  return qnb_eval_terms(h, QNB_T_RAD | QNB_T_POL, x, d, Tfree, istep, E, nullptr);
}

extern "C" int32_t qnb_adjoint(qnb_handle* h, const qnb_qtangent* tan, int32_t nframes, const double* frames,
                               double* out) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  if (nframes < 0 || (nframes > 0 && (!frames || !out))) return fail(QNB_EINVAL, "bad frames or output");
  const DevSys& S = h->hs;
  if (S.nqat <= 0 || nframes == 0) return 0;
  DevAlloc tmp;
  AdjDev A{};
  int rc = adjoint_upload(tmp, S, tan, A);
  if (rc) return rc;
  const int chunk = 256;
  const size_t fl = 3 * (size_t)S.natom, ol = (size_t)S.nstates * S.nqat * QNB_ADJ_NV;
  double *dfr = nullptr, *dout = nullptr;
  AdjDev* dadj = nullptr;
  if ((rc = tmp.alloc(&dfr, fl * std::min(nframes, chunk)))) return rc;
  if ((rc = tmp.alloc(&dout, ol * std::min(nframes, chunk)))) return rc;
  if ((rc = tmp.alloc(&dadj, 1))) return rc;
  for (int f0 = 0; f0 < nframes; f0 += chunk) {
    const int nf = std::min(chunk, nframes - f0);
    A.frames = dfr;
    A.nframes = nf;
    A.out = dout;
    HIPCHK(hipMemcpyAsync(dfr, frames + fl * f0, sizeof(double) * fl * nf, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dadj, &A, sizeof A, hipMemcpyHostToDevice, h->stream));
    launch_adjoint(h->stream, S.prec, S.nstates, (long)nf * S.nqat, 1, h->dsys, dadj);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out + ol * f0, dout, sizeof(double) * ol * nf, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return 0;
}

// qnb_adjoint without device allocation: the same kernel, with every buffer
// it needs beyond the caller's inputs and outputs in the fixed workspace,
// laid out as [parameter record | tangent tables | frames | output], each
// 256-byte aligned; frames go through in chunks that fit.
extern "C" int32_t qnb_adjoint_static(qnb_handle* h, const qnb_qtangent* tan, int32_t nframes, const double* frames,
                                      double* out) {
  if (!h) return fail(QNB_EINVAL, "null handle");
  if (nframes < 0 || (nframes > 0 && (!frames || !out))) return fail(QNB_EINVAL, "bad frames or output");
  if (!tan) return fail(QNB_EINVAL, "null tangent tables");
  const DevSys& S = h->hs;
  if (S.nqat <= 0 || nframes == 0) return 0;
  unsigned char* base = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch_ptr[h->device] : nullptr;
  if (!base) return fail(QNB_EDEVICE, "no workspace address for device %d", h->device);
  const size_t nq = S.nqat, ns = S.nstates, nsol = S.nat_solute, sa = S.sa;
  const size_t np = nsol * nq * ns, nw = S.nwat > 0 ? nq * sa * ns : 0, nqq = nq * nq * ns;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  // the tangent tables (layouts of the Q tables, qnb_qtangent)
  struct Tab {
    const void* src;
    size_t bytes;
    const char* name;
  };
  const Tab tabs[] = {{tan->qp_dq, 8 * np, "qp_dq"},     {tan->qp_dA, 8 * np, "qp_dA"},
                      {tan->qp_dB, 8 * np, "qp_dB"},     {tan->qp_is14, np, "qp_is14"},
                      {tan->qw_dq, 8 * nw, "qw_dq"},     {tan->qw_dA, 8 * nw, "qw_dA"},
                      {tan->qw_dB, 8 * nw, "qw_dB"},     {tan->qq_dq_i, 8 * nqq, "qq_dq_i"},
                      {tan->qq_dA_i, 8 * nqq, "qq_dA_i"}, {tan->qq_dB_i, 8 * nqq, "qq_dB_i"},
                      {tan->qq_dq_j, 8 * nqq, "qq_dq_j"}, {tan->qq_dA_j, 8 * nqq, "qq_dA_j"},
                      {tan->qq_dB_j, 8 * nqq, "qq_dB_j"}, {tan->qq_is14, nqq, "qq_is14"}};
  size_t off = al(sizeof(AdjDev));
  size_t toff[14];
  for (int t = 0; t < 14; t++) {
    if (tabs[t].bytes > 0 && !tabs[t].src) return fail(QNB_EINVAL, "tangent table %s missing", tabs[t].name);
    toff[t] = off;
    off += al(tabs[t].bytes);
  }
  const size_t fl = 3 * (size_t)S.natom, ol = ns * nq * QNB_ADJ_NV;
  const size_t per = al(8 * fl) + al(8 * ol);  // one frame and its output (upper bound per frame)
  if (off + per > QNB_ADJOINT_STATIC_BYTES)
    return fail(QNB_ENOMEM,
                "qnb_adjoint_static: tangent tables (%zu bytes) plus one frame and its output (%zu bytes) exceed the "
                "static workspace of %u bytes; use qnb_adjoint",
                off, per, (unsigned)QNB_ADJOINT_STATIC_BYTES);
  const size_t room = QNB_ADJOINT_STATIC_BYTES - off;
  int chunk = (int)std::min<size_t>((size_t)nframes, std::min<size_t>(room / (8 * fl + 8 * ol), 256));
  while (chunk > 1 && al(8 * fl * chunk) + al(8 * ol * chunk) > room) chunk--;
  double* dfr = (double*)(base + off);
  double* dout = (double*)(base + off + al(8 * fl * chunk));
  AdjDev A{};
  const void** dst[14] = {(const void**)&A.qp_dq,   (const void**)&A.qp_dA,   (const void**)&A.qp_dB,
                          (const void**)&A.qp_is14, (const void**)&A.qw_dq,   (const void**)&A.qw_dA,
                          (const void**)&A.qw_dB,   (const void**)&A.qq_dq_i, (const void**)&A.qq_dA_i,
                          (const void**)&A.qq_dB_i, (const void**)&A.qq_dq_j, (const void**)&A.qq_dA_j,
                          (const void**)&A.qq_dB_j, (const void**)&A.qq_is14};
  for (int t = 0; t < 14; t++) {
    *dst[t] = base + toff[t];
    if (tabs[t].bytes) HIPCHK(hipMemcpyAsync(base + toff[t], tabs[t].src, tabs[t].bytes, hipMemcpyHostToDevice, h->stream));
  }
  AdjDev* dadj = (AdjDev*)base;
  for (int f0 = 0; f0 < nframes; f0 += chunk) {
    const int nf = std::min(chunk, nframes - f0);
    A.frames = dfr;
    A.nframes = nf;
    A.out = dout;
    HIPCHK(hipMemcpyAsync(dfr, frames + fl * f0, sizeof(double) * fl * nf, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dadj, &A, sizeof A, hipMemcpyHostToDevice, h->stream));
    launch_adjoint(h->stream, S.prec, S.nstates, (long)nf * S.nqat, 1, h->dsys, dadj);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out + ol * f0, dout, sizeof(double) * ol * nf, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return 0;
}

extern "C" int32_t qnb_get_lrf(qnb_handle* h, double* lrf) {
  if (!h || !lrf) return fail(QNB_EINVAL, "null argument");
  HIPCHK(hipMemcpy(lrf, h->hs.lrf, sizeof(double) * QNB_NLRF * h->hs.ncgp, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int32_t qnb_get_shell_state(qnb_handle* h, double* st) {
  if (!h || !st) return fail(QNB_EINVAL, "null argument");
  double buf[3 * QNB_MAXSH];
  HIPCHK(hipMemcpy(buf, h->hs.sh_state, sizeof buf, hipMemcpyDeviceToHost));
  int n = h->hs.nshell;
  for (int k = 0; k < n; k++) {
    st[k] = buf[k];
    st[n + k] = buf[QNB_MAXSH + k];
    st[2 * n + k] = buf[2 * QNB_MAXSH + k];
  }
  return 0;
}

extern "C" int32_t qnb_set_shell_state(qnb_handle* h, const double* st) {
  if (!h || !st) return fail(QNB_EINVAL, "null argument");
  double buf[3 * QNB_MAXSH] = {0};
  int n = h->hs.nshell;
  for (int k = 0; k < n; k++) {
    buf[k] = st[k];
    buf[QNB_MAXSH + k] = st[n + k];
    buf[2 * QNB_MAXSH + k] = st[2 * n + k];
  }
  HIPCHK(hipMemcpy(h->hs.sh_state, buf, sizeof buf, hipMemcpyHostToDevice));
  return 0;
}

static inline bool owns_h(int a, int b) { return (a < b && ((a + b) & 1) == 0) || (a > b && ((a + b) & 1) == 1); }

// Expands the device's group-level lists into the reference's atom-pair lists,
// in the reference's append order (make_pair_lists, nonbondene.f90:749-837).
extern "C" int32_t qnb_get_pairs(qnb_handle* h, int32_t which, int32_t* ab, int64_t cap, int64_t* n) {
  if (!h || !n) return fail(QNB_EINVAL, "null argument");
  if (!h->lists_built) return fail(QNB_EINVAL, "no lists built yet");
  const DevSys& S = h->hs;
  std::vector<int32_t> out;
  std::vector<int> nn(S.ncgp), nbr((size_t)S.cap * S.ncgp);
  HIPCHK(hipMemcpy(nn.data(), S.nnbr, sizeof(int) * S.ncgp, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nbr.data(), S.nbr, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
  auto row = [&](int g) { return nbr.data() + (size_t)g * S.cap; };
  const int ns = S.ncgp_solute;
  switch (which) {
    case QNB_LIST_PP:
      for (int ig = 0; ig < ns; ig++) {
        if (h->T.excl[h->T.gswitch[ig]]) continue;
        std::vector<int> partners;
        for (int k = 0; k < nn[ig]; k++) {
          int b = row(ig)[k];
          if (b < ns && owns_h(ig, b)) partners.push_back(b);
        }
        partners.push_back(ig);
        std::sort(partners.begin(), partners.end());
        for (int jg : partners)
          for (int ia = h->T.gfirst[ig]; ia <= h->T.glast[ig]; ia++) {
            int i = h->T.cgpatom[ia];
            if (h->T.iqatom[i]) continue;
            for (int ja = h->T.gfirst[jg]; ja <= h->T.glast[jg]; ja++) {
              int j = h->T.cgpatom[ja];
              if (h->T.iqatom[j]) continue;
              if (ig == jg && i >= j) continue;
              if (!host_pp_param(h->T, i, j)) continue;
              out.push_back(i);
              out.push_back(j);
            }
          }
      }
      break;
    case QNB_LIST_PW:
      for (int ig = 0; ig < ns; ig++) {
        if (h->T.excl[h->T.gswitch[ig]]) continue;
        for (int k = 0; k < nn[ig]; k++) {
          int b = row(ig)[k];
          if (b < ns) continue;
          int o = S.nat_solute + S.sa * (b - ns);
          for (int ia = h->T.gfirst[ig]; ia <= h->T.glast[ig]; ia++) {
            int i = h->T.cgpatom[ia];
            if (h->T.iqatom[i]) continue;
            for (int j = 0; j < S.sa; j++) {
              out.push_back(i);
              out.push_back(o + j);
            }
          }
        }
      }
      break;
    case QNB_LIST_WW:
      for (int a = ns; a < S.ncgp; a++)
        for (int k = 0; k < nn[a]; k++) {
          int b = row(a)[k];
          if (b >= ns && owns_h(a, b)) {
            out.push_back(a - ns);
            out.push_back(b - ns);
          }
        }
      break;
    case QNB_LIST_QP:
    case QNB_LIST_QW: {
      int nql[2] = {0, 0};
      HIPCHK(hipMemcpy(nql, S.nq_lists, sizeof nql, hipMemcpyDeviceToHost));
      int cnt = which == QNB_LIST_QW ? nql[0] : nql[1];
      std::vector<int> lst(std::max(cnt, 1));
      if (cnt) HIPCHK(hipMemcpy(lst.data(), which == QNB_LIST_QW ? S.qw_list : S.qp_list, sizeof(int) * cnt,
                                hipMemcpyDeviceToHost));
      for (int m = 0; m < cnt; m++)
        for (int iq = 0; iq < S.nqat; iq++) {
          out.push_back(iq);
          out.push_back(lst[m]);
        }
      break;
    }
    case QNB_LIST_QQ:
      for (size_t m = 0; m < h->T.qq_iq.size(); m++) {
        out.push_back(h->T.qq_iq[m]);
        out.push_back(h->T.qq_jq[m]);
      }
      break;
    case QNB_LIST_QQP:
      for (size_t m = 0; m < h->T.qqp_iq.size(); m++) {
        out.push_back(h->T.qqp_iq[m]);
        out.push_back(h->T.qqp_j[m]);
      }
      break;
    default:
      return fail(QNB_EINVAL, "unknown list %d", which);
  }
  *n = (int64_t)(out.size() / 2);
  if (ab && *n <= cap) memcpy(ab, out.data(), sizeof(int32_t) * out.size());
  return 0;
}

