/*
 * qnb_oracle.c -- CPU restatement of Qdyn6's spherical-boundary nonbonded path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is held
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The engine (q6_amd/csrc) never calls it and has no CPU path.
 *
 * It restates, in plain C and in the reference's loop and summation order, the
 * routines below (paths relative to the reference's src/).  Every function
 * cites the lines it follows.  It is built with -O2 -ffp-contract=off (no FMA
 * contraction, as the reference's IEEE build) so that list membership, which
 * hinges on `r2 <= Rc^2` of q_dist4 (math.f90:254-262), is reproduced bit for
 * bit.
 *
 * Parity status (what pins each function; the tests that do it are named):
 *   list builders, LRF, all nonbond_* loops, lrf_taylor, restrain_solvent,
 *   watpol: golden vectors of the reference itself (stock Qdyn6 built from
 *   /root/reference by oracle/ref/Makefile, dumped by oracle/ref/qnb_dump.f90)
 *   -- tests/test_oracle_golden.py; pair kernels additionally by closed forms
 *   and finite differences -- tests/test_oracle_math.py.
 */
#include "qnb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define QREAL_EPS 1.0e-10 /* sizes.f90:50 */
#define NLRF 43

typedef struct {
  double x, y, z;
} vec3;

static inline vec3 v3(const double *x, int i) {
  vec3 r = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  return r;
}
/* qvec_square (math.f90:206-212): a%x**2 + a%y**2 + a%z**2, left to right */
static inline double sq(vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* q_dotprod (math.f90:275-284) */
static inline double dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* q_dist4 (math.f90:254-262): |b - a|^2 */
static inline double dist4(vec3 a, vec3 b) {
  vec3 t = {b.x - a.x, b.y - a.y, b.z - a.z};
  return sq(t);
}

typedef struct {
  int32_t *i, *j;
  double *par; /* stride npar per entry (per state for Q lists) */
  int64_t n, cap;
  int npar;
} plist;

struct orc_handle {
  qnb_system s;
  qnb_params p;
  int32_t *iwhich_cgp;
  plist pp, pw, ww, qp, qw; /* qp/qw: par = [s][4] per entry */
  plist qq[QNB_MAX_STATES], qqp[QNB_MAX_STATES];
  uint8_t *qq_soft[QNB_MAX_STATES];
  int qp_done, qw_done; /* the `save`d list_done flags of nbqplist/nbqwlist */
  double *lrf;          /* [ncgp][43] */
  /* polarisation-restraint state (wshell, src/globals.f90:165-169) */
  double theta_corr[QNB_MAX_SHELLS], avtheta[QNB_MAX_SHELLS], avn_insh[QNB_MAX_SHELLS];
  int32_t n_insh[QNB_MAX_SHELLS];
  double *theta, *tdum;
  int32_t *list_sh, *nsort; /* [shell][nwat] */
};

static void pl_init(plist *l, int npar) {
  memset(l, 0, sizeof(*l));
  l->npar = npar;
}
static void pl_free(plist *l) {
  free(l->i);
  free(l->j);
  free(l->par);
  memset(l, 0, sizeof(*l));
}
static double *pl_push(plist *l, int32_t i, int32_t j) {
  if (l->n == l->cap) {
    /* reallocate_nonbondlist_* (qalloc.f90:370-605): grow in place */
    int64_t nc = l->cap ? 2 * l->cap : 1024;
    l->i = (int32_t *)realloc(l->i, nc * sizeof(int32_t));
    l->j = (int32_t *)realloc(l->j, nc * sizeof(int32_t));
    l->par = (double *)realloc(l->par, nc * l->npar * sizeof(double));
    l->cap = nc;
  }
  l->i[l->n] = i;
  l->j[l->n] = j;
  return l->par + (l->n++) * l->npar;
}

/* ---------------------------------------------------------------- set-up */

orc_handle *orc_create(const qnb_system *s, const qnb_params *p) {
  orc_handle *h = (orc_handle *)calloc(1, sizeof(orc_handle));
  h->s = *s;
  h->p = *p;
  h->iwhich_cgp = (int32_t *)malloc(sizeof(int32_t) * (s->natom > 0 ? s->natom : 1));
  for (int a = 0; a < s->natom; a++) h->iwhich_cgp[a] = -1;
  for (int g = 0; g < s->ncgp; g++)
    for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++) h->iwhich_cgp[s->cgpatom[k]] = g;
  pl_init(&h->pp, 3);
  pl_init(&h->pw, 3);
  pl_init(&h->ww, 3);
  pl_init(&h->qp, 4 * s->nstates);
  pl_init(&h->qw, 4 * s->nstates);
  for (int st = 0; st < s->nstates; st++) {
    pl_init(&h->qq[st], 4);
    pl_init(&h->qqp[st], 4);
  }
  h->lrf = (double *)calloc((size_t)NLRF * (s->ncgp > 0 ? s->ncgp : 1), sizeof(double));
  for (int k = 0; k < p->nwpolr_shell; k++) h->theta_corr[k] = p->wshell_theta_corr[k];
  int nw = s->nwat > 0 ? s->nwat : 1;
  h->theta = (double *)calloc(nw, sizeof(double));
  h->tdum = (double *)calloc(nw, sizeof(double));
  h->list_sh = (int32_t *)calloc((size_t)QNB_MAX_SHELLS * nw, sizeof(int32_t));
  h->nsort = (int32_t *)calloc((size_t)QNB_MAX_SHELLS * nw, sizeof(int32_t));
  orc_nbqqlist(h);
  return h;
}

void orc_destroy(orc_handle *h) {
  if (!h) return;
  free(h->iwhich_cgp);
  pl_free(&h->pp);
  pl_free(&h->pw);
  pl_free(&h->ww);
  pl_free(&h->qp);
  pl_free(&h->qw);
  for (int st = 0; st < QNB_MAX_STATES; st++) {
    pl_free(&h->qq[st]);
    pl_free(&h->qqp[st]);
    free(h->qq_soft[st]);
  }
  free(h->lrf);
  free(h->theta);
  free(h->tdum);
  free(h->list_sh);
  free(h->nsort);
  free(h);
}

/* nbqqlist (nonbondene.f90:3231-3282): static Q-Q and Q-"bonded neighbour"
 * lists, one per state, in the reference's append order. */
void orc_nbqqlist(orc_handle *h) {
  const qnb_system *s = &h->s;
  int nq = s->nqat, ns = s->nstates, nsol = s->nat_solute;
  for (int st = 0; st < ns; st++) {
    h->qq[st].n = 0;
    h->qqp[st].n = 0;
    free(h->qq_soft[st]);
    h->qq_soft[st] = (uint8_t *)calloc((size_t)nq * nq + 1, 1);
  }
  for (int iq = 0; iq < nq - 1; iq++)
    for (int jq = iq + 1; jq < nq; jq++)
      for (int st = 0; st < ns; st++) {
        size_t k = (size_t)iq + nq * ((size_t)jq + (size_t)nq * st);
        if (!s->qq_set[k]) continue;
        h->qq_soft[st][h->qq[st].n] = s->qq_soft[k];
        double *pr = pl_push(&h->qq[st], iq, jq);
        pr[0] = s->qq_A[k];
        pr[1] = s->qq_B[k];
        pr[2] = s->qq_el[k];
        pr[3] = s->qq_sc[k];
      }
  for (int ja = 0; ja < nsol; ja++) {
    if (s->iqatom[ja] != 0) continue;
    if (!s->qconn3[ja]) continue;
    for (int iq = 0; iq < nq; iq++)
      for (int st = 0; st < ns; st++) {
        size_t k = (size_t)ja + nsol * ((size_t)iq + (size_t)nq * st);
        if (!s->qp_set[k]) continue;
        double *pr = pl_push(&h->qqp[st], iq, ja);
        pr[0] = s->qp_A[k];
        pr[1] = s->qp_B[k];
        pr[2] = s->qp_el[k];
        pr[3] = s->qp_sc[k];
      }
  }
}

/* -------------------------------------------------------- pair parameters */

/* Solute-solute pair (i, j) as the list orientation has it.  Follows
 * pp_int_comp (simprep.f90:2952-3065) and precompute_set_values_pp
 * (simprep.f90:3321-3350): exclusions from listex/listexlong, 1-4 pairs from
 * list14/list14long (LJ code 3), otherwise ljcod; the pair enters the list
 * only when one of |A|, |B|, |qq| exceeds QREAL_EPS (before 1-4 scaling).
 * Returns 0 when the pair does not interact. */
static int pp_param(const qnb_system *s, int i, int j, double *A, double *B, double *el) {
  int vdw;
  int lo = i < j ? i : j, dk = i < j ? j - i : i - j;
  int is14 = 0;
  if (dk <= QNB_MAX_NBR_RANGE) {
    if (s->listex[lo * QNB_MAX_NBR_RANGE + dk - 1]) return 0;
    if (s->list14[lo * QNB_MAX_NBR_RANGE + dk - 1]) is14 = 1;
  } else {
    for (int m = 0; m < s->nexlong; m++) {
      int a = s->listexlong[2 * m], b = s->listexlong[2 * m + 1];
      if ((a == i && b == j) || (a == j && b == i)) return 0;
    }
    for (int m = 0; m < s->n14long; m++) {
      int a = s->list14long[2 * m], b = s->list14long[2 * m + 1];
      if ((a == i && b == j) || (a == j && b == i)) {
        is14 = 1;
        break;
      }
    }
  }
  int ti = s->iac[i], tj = s->iac[j];
  vdw = is14 ? 3 : s->ljcod[ti + tj * s->ntypes];
  int k = vdw - 1;
  if (s->ivdw_rule == QNB_VDW_GEOMETRIC) {
    *A = s->avdw[ti * 3 + k] * s->avdw[tj * 3 + k];
    *B = s->bvdw[ti * 3 + k] * s->bvdw[tj * 3 + k];
  } else {
    double tA = s->avdw[ti * 3 + k] + s->avdw[tj * 3 + k];
    tA = tA * tA;
    tA = tA * tA * tA;
    double tB = s->bvdw[ti * 3 + k] * s->bvdw[tj * 3 + k];
    *A = (tA * tA) * tB;
    *B = 2.0 * tA * tB;
  }
  *el = s->crg[i] * s->crg[j];
  int set = fabs(*A - 0.0) > QREAL_EPS || fabs(*B - 0.0) > QREAL_EPS || fabs(*el - 0.0) > QREAL_EPS;
  if (vdw == 3) *el = *el * s->el14_scale;
  return set;
}

/* ---------------------------------------------------------------- LRF */

/* cgp_centers (nonbondene.f90:43-78) */
static void cgp_centers(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  for (int g = 0; g < s->ncgp; g++) {
    double *L = h->lrf + (size_t)NLRF * g;
    memset(L, 0, sizeof(double) * NLRF);
    for (int k = s->cgp_first[g]; k <= s->cgp_last[g]; k++) {
      int a = s->cgpatom[k];
      L[0] = L[0] + x[3 * a];
      L[1] = L[1] + x[3 * a + 1];
      L[2] = L[2] + x[3 * a + 2];
    }
    double n = (double)(s->cgp_last[g] - s->cgp_first[g] + 1);
    L[0] = L[0] / n;
    L[1] = L[1] / n;
    L[2] = L[2] / n;
  }
}

/* lrf_update (nonbondene.f90:628-725): expansion of group g1's non-Q charges
 * about the centre of g2.  Table layout per group: cent[0..2], phi0[3],
 * phi1[4..6], phi2(k)[7+3k .. 9+3k], phi3(m)[16+3m .. 18+3m]. */
static void lrf_update(orc_handle *h, const double *x, int g1, int g2) {
  const qnb_system *s = &h->s;
  double *L = h->lrf + (size_t)NLRF * g2;
  double *phi2 = L + 7, *phi3 = L + 16;
  for (int ia = s->cgp_first[g1]; ia <= s->cgp_last[g1]; ia++) {
    int i = s->cgpatom[ia];
    if (s->iqatom[i] != 0) continue;
    vec3 dr = {x[3 * i] - L[0] - 0.0, x[3 * i + 1] - L[1] - 0.0, x[3 * i + 2] - L[2] - 0.0};
    double r2 = sq(dr);
    double f0 = s->crg[i] / (r2 * sqrt(r2));
    double f1 = 3.0 * f0 / r2;
    double f2 = -f1 / r2;
    L[3] = L[3] + f0 * r2;
    L[4] = L[4] - dr.x * f0;
    L[5] = L[5] - dr.y * f0;
    L[6] = L[6] - dr.z * f0;
    double t1[3] = {dr.x * f1, dr.y * f1, dr.z * f1};
    double d3[3] = {dr.x, dr.y, dr.z};
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) phi2[3 * k + c] = phi2[3 * k + c] + d3[c] * t1[k];
    phi2[0] = phi2[0] - f0;
    phi2[4] = phi2[4] - f0;
    phi2[8] = phi2[8] - f0;
    /* phi3(m), m = j + 3(n-1): component c += ((((dr_n*dr_j)*dr_c)*5)*f2) */
    for (int n = 0; n < 3; n++)
      for (int j = 0; j < 3; j++) {
        int m = j + 3 * n;
        double tk = d3[n] * d3[j];
        for (int c = 0; c < 3; c++) {
          double tl = tk * d3[c];
          tl = tl * 5.0;
          tl = tl * f2;
          phi3[3 * m + c] = phi3[3 * m + c] + tl;
        }
      }
    double t[3] = {dr.x * r2, dr.y * r2, dr.z * r2};
#define P3(m, c) phi3[3 * ((m)-1) + (c)]
    P3(1, 0) = P3(1, 0) - f2 * (3.0 * t[0]);
    P3(1, 1) = P3(1, 1) - f2 * (t[1]);
    P3(1, 2) = P3(1, 2) - f2 * (t[2]);
    P3(2, 0) = P3(2, 0) - f2 * (t[1]);
    P3(2, 1) = P3(2, 1) - f2 * (t[0]);
    P3(3, 0) = P3(3, 0) - f2 * (t[2]);
    P3(3, 2) = P3(3, 2) - f2 * (t[0]);
    P3(4, 0) = P3(4, 0) - f2 * (t[1]);
    P3(4, 1) = P3(4, 1) - f2 * (t[0]);
    P3(5, 0) = P3(5, 0) - f2 * (t[0]);
    P3(5, 1) = P3(5, 1) - f2 * (3.0 * t[1]);
    P3(5, 2) = P3(5, 2) - f2 * (t[2]);
    P3(6, 1) = P3(6, 1) - f2 * (t[2]);
    P3(6, 2) = P3(6, 2) - f2 * (t[1]);
    P3(7, 0) = P3(7, 0) - f2 * (t[2]);
    P3(7, 2) = P3(7, 2) - f2 * (t[0]);
    P3(8, 1) = P3(8, 1) - f2 * (t[2]);
    P3(8, 2) = P3(8, 2) - f2 * (t[1]);
    P3(9, 0) = P3(9, 0) - f2 * (t[0]);
    P3(9, 1) = P3(9, 1) - f2 * (t[1]);
    P3(9, 2) = P3(9, 2) - f2 * (3.0 * t[2]);
#undef P3
  }
}

/* ------------------------------------------------------------- list builders */

static inline int parity_skip(int ig, int jg) {
  /* count each group pair once (nonbondene.f90:1589-1591); 0-based indices
   * have the same parity sum as the reference's 1-based ones */
  return (ig > jg && ((ig + jg) % 2 == 0)) || (ig < jg && ((ig + jg) % 2 == 1));
}

/* Group test of the iuse_switch_atom = 0 builders: the atom pairs of groups
 * ig and jg are scanned ig-major until one lies within rc2 (nbpplis2
 * nonbondene.f90:1031-1047, nbpplis2_lrf :1451-1468).  Returns 1 when one
 * does; otherwise 0 with *r2last the squared distance of the last pair
 * scanned (the two groups' last atoms), which the _lrf form compares with
 * RcLRF^2 (:1499). */
static int any_pair_inside(const qnb_system *s, const double *x, int ig, int jg, double rc2, double *r2last) {
  double r2 = 1.0;
  for (int ia = s->cgp_first[ig]; ia <= s->cgp_last[ig]; ia++)
    for (int ja = s->cgp_first[jg]; ja <= s->cgp_last[jg]; ja++) {
      r2 = dist4(v3(x, s->cgpatom[ia]), v3(x, s->cgpatom[ja]));
      if (r2 <= rc2) return 1;
    }
  *r2last = r2;
  return 0;
}

/* the same with a single point on the other side: the water's first atom
 * (nbpwlis2 :2197-2208, nbpwlis2_lrf :2566-2577; LRF test :2601) or xpcent
 * (nbqplis2 :3477-3488) */
static int any_atom_inside(const qnb_system *s, const double *x, int ig, vec3 p, double rc2, double *r2last) {
  double r2 = 1.0;
  for (int ia = s->cgp_first[ig]; ia <= s->cgp_last[ig]; ia++) {
    r2 = dist4(v3(x, s->cgpatom[ia]), p);
    if (r2 <= rc2) return 1;
  }
  *r2last = r2;
  return 0;
}

/* nbpplist_lrf (nonbondene.f90:1777-1908) and, for use_LRF = 0, nbpplist
 * (nonbondene.f90:1520-1639); with any_atom_cutoff, nbpplis2_lrf (:1369-1488)
 * and nbpplis2 (:950-1089). */
static void nbpplist(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  const double rc2 = h->p.Rcpp * h->p.Rcpp, rl2 = h->p.RcLRF * h->p.RcLRF;
  h->pp.n = 0;
  for (int ig = 0; ig < s->ncgp_solute; ig++) {
    int is = s->cgp_iswitch[ig];
    if (s->excl[is]) continue;
    for (int jg = 0; jg < s->ncgp_solute; jg++) {
      if (parity_skip(ig, jg)) continue;
      int js = s->cgp_iswitch[jg];
      if (s->excl[js]) continue;
      double r2 = 0.0;
      int inside;
      if (s->any_atom_cutoff) inside = any_pair_inside(s, x, ig, jg, rc2, &r2);
      else {
        r2 = dist4(v3(x, is), v3(x, js));
        inside = r2 <= rc2;
      }
      if (inside) {
        for (int ia = s->cgp_first[ig]; ia <= s->cgp_last[ig]; ia++) {
          int i = s->cgpatom[ia];
          if (s->iqatom[i] != 0) continue;
          for (int ja = s->cgp_first[jg]; ja <= s->cgp_last[jg]; ja++) {
            int j = s->cgpatom[ja];
            if (s->iqatom[j] != 0) continue;
            if (ig == jg && i >= j) continue;
            double A, B, el;
            if (!pp_param(s, i, j, &A, &B, &el)) continue;
            double *pr = pl_push(&h->pp, i, j);
            pr[0] = A;
            pr[1] = B;
            pr[2] = el;
          }
        }
      } else if (h->p.use_LRF && r2 <= rl2) {
        lrf_update(h, x, ig, jg);
        lrf_update(h, x, jg, ig);
      }
    }
  }
}

/* nbpwlist_lrf (nonbondene.f90:2842-2955) and nbpwlist (2622-2725); with
 * any_atom_cutoff, nbpwlis2_lrf (:2493-2620) and nbpwlis2 (:2137-2230). */
static void nbpwlist(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  const double rc2 = h->p.Rcpw * h->p.Rcpw, rl2 = h->p.RcLRF * h->p.RcLRF;
  const int sa = s->solv_atom, ns = s->nat_solute;
  const int quirk = !h->p.use_LRF && (h->p.ref_quirks & 1) && !s->any_atom_cutoff;
  h->pw.n = 0;
  for (int ig = 0; ig < s->ncgp_solute; ig++) {
    int is = s->cgp_iswitch[ig];
    if (s->excl[is]) continue;
    int ref = is;
    for (int jw = 0; jw < s->nwat; jw++) {
      int ja = ns + sa * jw;
      int jg = h->iwhich_cgp[ja];
      if (s->excl[ja]) continue;
      double r2 = 0.0;
      int inside;
      if (s->any_atom_cutoff) inside = any_atom_inside(s, x, ig, v3(x, ja), rc2, &r2);
      else {
        r2 = dist4(v3(x, ref), v3(x, ja));
        inside = r2 <= rc2;
      }
      if (inside) {
        for (int ia = s->cgp_first[ig]; ia <= s->cgp_last[ig]; ia++) {
          int i = s->cgpatom[ia];
          if (s->iqatom[i] != 0) continue;
          for (int j = 0; j < sa; j++) {
            size_t k = (size_t)i + (size_t)ns * j;
            double *pr = pl_push(&h->pw, i, ja + j);
            pr[0] = s->pw_A[k];
            pr[1] = s->pw_B[k];
            pr[2] = s->pw_el[k];
          }
        }
        /* the reference's loop variable `ia` ends at cgp(ig)%last + 1 and is
         * then used as the distance test atom (nonbondene.f90:2686-2691) */
        if (quirk) ref = s->cgp_last[ig] + 1;
      } else if (h->p.use_LRF && r2 <= rl2) {
        lrf_update(h, x, ig, jg);
        lrf_update(h, x, jg, ig);
      }
    }
  }
}

/* nbwwlist_lrf (nonbondene.f90:4274-4395) and nbwwlist (4079-4172). */
static void nbwwlist(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  const double rc2 = h->p.Rcww * h->p.Rcww, rl2 = h->p.RcLRF * h->p.RcLRF;
  const int sa = s->solv_atom, ns = s->nat_solute;
  h->ww.n = 0;
  for (int iw = 0; iw < s->nwat; iw++) {
    int is = ns + sa * iw;
    if (s->excl[is]) continue;
    int ig = h->iwhich_cgp[is];
    for (int jw = 0; jw < s->nwat; jw++) {
      int ja = ns + sa * jw;
      if (s->excl[ja]) continue;
      int jg = h->iwhich_cgp[ja];
      if (parity_skip(iw, jw) || iw == jw) continue;
      double r2 = dist4(v3(x, is), v3(x, ja));
      if (r2 <= rc2) {
        double *pr = pl_push(&h->ww, iw, jw); /* stored as a molecule pair */
        pr[0] = pr[1] = pr[2] = 0.0;
      } else if (h->p.use_LRF && r2 <= rl2) {
        lrf_update(h, x, ig, jg);
        lrf_update(h, x, jg, ig);
      }
    }
  }
}

/* nbqplist (nonbondene.f90:3647-3742): per-state parameters of every
 * (Q-atom, solute atom) pair whose group's switch atom is within Rcq of
 * xpcent (with any_atom_cutoff, any atom of the group: nbqplis2
 * :3408-3470); built once when Rcq > rexcl_o. */
static void nbqplist(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  const double rq2 = h->p.Rcq * h->p.Rcq;
  if (h->qp_done && rq2 > h->p.rexcl_o * h->p.rexcl_o) return;
  if (s->nqat == 0) return;
  const int nq = s->nqat, nst = s->nstates, nsol = s->nat_solute;
  vec3 xp = {h->p.xpcent[0], h->p.xpcent[1], h->p.xpcent[2]};
  h->qp.n = 0;
  for (int ig = 0; ig < s->ncgp_solute; ig++) {
    int ia = s->cgp_iswitch[ig];
    if (s->excl[ia]) continue;
    if (s->any_atom_cutoff) {
      double r2l;
      if (!any_atom_inside(s, x, ig, xp, rq2, &r2l)) continue;
    } else if (dist4(v3(x, ia), xp) > rq2)
      continue;
    for (int k = s->cgp_first[ig]; k <= s->cgp_last[ig]; k++) {
      int i = s->cgpatom[k];
      if (i < nsol && s->qconn3[i]) continue;
      for (int iq = 0; iq < nq; iq++) {
        double *pr = pl_push(&h->qp, iq, i);
        for (int st = 0; st < nst; st++) {
          size_t q = (size_t)i + nsol * ((size_t)iq + (size_t)nq * st);
          pr[4 * st + 0] = s->qp_A[q];
          pr[4 * st + 1] = s->qp_B[q];
          pr[4 * st + 2] = s->qp_el[q];
          pr[4 * st + 3] = s->qp_sc[q];
        }
      }
    }
  }
  h->qp_done = 1;
}

/* nbqwlist (nonbondene.f90:3862-3944): every non-excluded water whose first
 * atom is within Rcq of xpcent, with every Q-atom; stored here as (iq, water)
 * entries, each standing for solv_atom atom pairs in the reference's order. */
static void nbqwlist(orc_handle *h, const double *x) {
  const qnb_system *s = &h->s;
  const double rq2 = h->p.Rcq * h->p.Rcq;
  if (h->qw_done && rq2 > h->p.rexcl_o * h->p.rexcl_o) return;
  if (s->nqat == 0) return;
  vec3 xp = {h->p.xpcent[0], h->p.xpcent[1], h->p.xpcent[2]};
  h->qw.n = 0;
  for (int iw = 0; iw < s->nwat; iw++) {
    int ia = s->nat_solute + s->solv_atom * iw;
    if (s->excl[ia]) continue;
    double r2 = dist4(v3(x, ia), xp);
    if (r2 <= rq2)
      for (int iq = 0; iq < s->nqat; iq++) pl_push(&h->qw, iq, iw);
  }
  h->qw_done = 1;
}

/* make_pair_lists, spherical branch (nonbondene.f90:796-825). */
void orc_make_pair_lists(orc_handle *h, const double *x) {
  if (h->p.use_LRF) cgp_centers(h, x);
  nbpplist(h, x);
  nbpwlist(h, x);
  nbqplist(h, x);
  nbwwlist(h, x);
  nbqwlist(h, x);
}

/* ------------------------------------------------------------ pair kernels */

/* nbe (nonbonded.f90:45-75) with q_dist2 (math.f90:226-238); returns dv and
 * the vector u = x_j - x_i, energies added to el/vdw. */
static inline double nbe(const double *x, int i, int j, double A, double B, double q, vec3 *u, double *el,
                         double *vdw) {
  vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
  double r2 = 1.0 / sq(t);
  double r = sqrt(r2);
  double r6 = r2 * r2 * r2;
  double r12 = r6 * r6;
  double Vel = q * r, Va = A * r12, Vb = B * r6;
  double dv = r2 * (-Vel - 12.0 * Va + 6.0 * Vb);
  *u = t;
  *el = *el + Vel;
  *vdw = *vdw + Va - Vb;
  return dv;
}

/* nbe_spc (nonbonded.f90:247-267) */
static inline double nbe_spc(const double *x, int i, int j, double q, vec3 *u, double *el) {
  vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
  double r2 = 1.0 / sq(t);
  double r = sqrt(r2);
  double Vel = q * r;
  *u = t;
  *el = *el + Vel;
  return r2 * (-Vel);
}

static inline void dupd(double *d, int i, int j, vec3 u, double dv) {
  d[3 * i] = d[3 * i] - u.x * dv;
  d[3 * i + 1] = d[3 * i + 1] - u.y * dv;
  d[3 * i + 2] = d[3 * i + 2] - u.z * dv;
  d[3 * j] = d[3 * j] + u.x * dv;
  d[3 * j + 1] = d[3 * j + 1] + u.y * dv;
  d[3 * j + 2] = d[3 * j + 2] + u.z * dv;
}

/* nonbond_pp / nonbond_pw (nonbondene.f90:4694-4760, 4856-4916) */
static void nonbond_list(const plist *l, const double *x, double *d, double *el, double *vdw) {
  for (int64_t p = 0; p < l->n; p++) {
    vec3 u;
    const double *pr = l->par + 3 * p;
    double dv = nbe(x, l->i[p], l->j[p], pr[0], pr[1], pr[2], &u, el, vdw);
    dupd(d, l->i[p], l->j[p], u, dv);
  }
}

/* nonbond_ww_spc (nonbondene.f90:5886-5975) and nonbond_ww (5509-5573) over
 * the molecule-pair list: 9 atom pairs (la, ka) per entry in the order the
 * reference list stores them (nonbondene.f90:4366-4375). */
static void nonbond_ww(orc_handle *h, const double *x, double *d, double *el, double *vdw) {
  const qnb_system *s = &h->s;
  const int sa = s->solv_atom, ns = s->nat_solute;
  for (int64_t p = 0; p < h->ww.n; p++) {
    int iw = h->ww.i[p], jw = h->ww.j[p];
    for (int la = 0; la < sa; la++)
      for (int ka = 0; ka < sa; ka++) {
        int i = ns + sa * iw + la, j = ns + sa * jw + ka;
        int k = la + ka * sa;
        vec3 u;
        double dv;
        if (s->solvent_spc && (la != 0 || ka != 0))
          dv = nbe_spc(x, i, j, s->ww_el[k], &u, el);
        else
          dv = nbe(x, i, j, s->ww_A[k], s->ww_B[k], s->ww_el[k], &u, el, vdw);
        dupd(d, i, j, u, dv);
      }
  }
}

/* nonbond_solvent_internal (nonbondene.f90:6082-6154): the listed pairs
 * inside every solvent molecule, excluded molecules included, with q_dist2
 * and the routine's own grouping of dv, into the solvent-solvent energies */
static void nonbond_solvent_internal(orc_handle *h, const double *x, double *d, double *el, double *vdw) {
  const qnb_system *s = &h->s;
  for (int iw = 0; iw < s->nwat; iw++) {
    const int is = s->nat_solute + s->solv_atom * iw;
    for (int ip = 0; ip < s->n_solv_int; ip++) {
      const int ia = is + s->solv_int_ij[2 * ip], ib = is + s->solv_int_ij[2 * ip + 1];
      const double *pr = s->solv_int_par + 3 * ip;
      vec3 t = {x[3 * ib] - x[3 * ia], x[3 * ib + 1] - x[3 * ia + 1], x[3 * ib + 2] - x[3 * ia + 2]};
      double r2 = 1.0 / sq(t);
      double r = sqrt(r2);
      double r6 = r2 * r2 * r2;
      double r12 = r6 * r6;
      double Vel = pr[2] * r, Va = pr[0] * r12, Vb = pr[1] * r6;
      double dv = r2 * (-Vel) + r2 * (-12.0 * Va + 6.0 * Vb);
      *el = *el + Vel;
      *vdw = *vdw + Va - Vb;
      dupd(d, ia, ib, t, dv);
    }
  }
}

/* nbe_qx (nonbonded.f90:200-222) with q_dist3's r2, r, r6 */
static inline double nbe_qx(double r2, double r, double r6in, const double *pr, double lam, double *Vel,
                            double *Va, double *Vb) {
  double r6_hc = 1.0 / r6in;
  double r6 = r6_hc + pr[3];
  r6 = 1.0 / r6;
  double r12 = r6 * r6;
  *Vel = pr[2] * r;
  *Va = pr[0] * r12;
  *Vb = pr[1] * r6;
  return r2 * (-*Vel - (12.0 * *Va - 6.0 * *Vb) * r6 * r6_hc) * lam;
}

/* nonbond_qp (nonbondene.f90:5157-5230): shared distance, d updated per state */
static void nonbond_qp(orc_handle *h, const double *x, double *d, qnb_qenergies *EQ) {
  const qnb_system *s = &h->s;
  for (int64_t p = 0; p < h->qp.n; p++) {
    int iq = h->qp.i[p], i = s->iqseq[iq], j = h->qp.j[p];
    vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    double r2 = 1.0 / sq(t), r = sqrt(r2), r6 = r2 * r2 * r2;
    for (int st = 0; st < s->nstates; st++) {
      double Vel, Va, Vb;
      double dv = nbe_qx(r2, r, r6, h->qp.par + (size_t)p * h->qp.npar + 4 * st, h->p.lambda[st], &Vel, &Va, &Vb);
      dupd(d, i, j, t, dv);
      EQ[st].qp_el = EQ[st].qp_el + Vel;
      EQ[st].qp_vdw = EQ[st].qp_vdw + Va - Vb;
    }
  }
}

/* nonbond_qw_spc (nonbondene.f90:5658-5767) / nonbond_qw (5339-5418) */
static void nonbond_qw(orc_handle *h, const double *x, double *d, qnb_qenergies *EQ) {
  const qnb_system *s = &h->s;
  const int sa = s->solv_atom, nq = s->nqat, nst = s->nstates;
  for (int64_t p = 0; p < h->qw.n; p++) {
    int iq = h->qw.i[p], iw = h->qw.j[p];
    int i = s->iqseq[iq], j0 = s->nat_solute + sa * iw;
    if (s->solvent_spc) {
      vec3 t = {x[3 * j0] - x[3 * i], x[3 * j0 + 1] - x[3 * i + 1], x[3 * j0 + 2] - x[3 * i + 2]};
      double r2 = 1.0 / sq(t), r = sqrt(r2), r6 = r2 * r2 * r2;
      double dv = 0.0;
      for (int st = 0; st < nst; st++) {
        size_t k = (size_t)iq + nq * (0 + (size_t)sa * st);
        double pr[4] = {s->qw_A[k], s->qw_B[k], s->qw_el[k], s->qw_sc[k]};
        double Vel, Va, Vb;
        dv = dv + nbe_qx(r2, r, r6, pr, h->p.lambda[st], &Vel, &Va, &Vb);
        EQ[st].qw_el = EQ[st].qw_el + Vel;
        EQ[st].qw_vdw = EQ[st].qw_vdw + Va - Vb;
      }
      dupd(d, i, j0, t, dv);
      for (int ip = 1; ip < sa; ip++) {
        int ja = j0 + ip;
        vec3 u = {x[3 * ja] - x[3 * i], x[3 * ja + 1] - x[3 * i + 1], x[3 * ja + 2] - x[3 * i + 2]};
        double rr2 = 1.0 / sq(u), rr = sqrt(rr2);
        dv = 0.0;
        for (int st = 0; st < nst; st++) {
          size_t k = (size_t)iq + nq * ((size_t)ip + (size_t)sa * st);
          double Vel = s->qw_el[k] * rr; /* nbe_qspc (nonbonded.f90:226-243) */
          EQ[st].qw_el = EQ[st].qw_el + Vel;
          dv = dv + (-rr2 * Vel * h->p.lambda[st]);
        }
        dupd(d, i, ja, u, dv);
      }
    } else {
      for (int jw = 0; jw < sa; jw++) {
        int j = j0 + jw;
        vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
        double r2 = 1.0 / sq(t), r = sqrt(r2), r6 = r2 * r2 * r2;
        for (int st = 0; st < nst; st++) {
          size_t k = (size_t)iq + nq * ((size_t)jw + (size_t)sa * st);
          double pr[4] = {s->qw_A[k], s->qw_B[k], s->qw_el[k], s->qw_sc[k]};
          double Vel, Va, Vb;
          double dv = nbe_qx(r2, r, r6, pr, h->p.lambda[st], &Vel, &Va, &Vb);
          dupd(d, i, j, t, dv);
          EQ[st].qw_el = EQ[st].qw_el + Vel;
          EQ[st].qw_vdw = EQ[st].qw_vdw + Va - Vb;
        }
      }
    }
  }
}

/* nonbond_qq (nonbondene.f90:5013-5082) with nbe_qq (nonbonded.f90:111-150) */
static void nonbond_qq(orc_handle *h, const double *x, double *d, qnb_qenergies *EQ) {
  const qnb_system *s = &h->s;
  for (int st = 0; st < s->nstates; st++) {
    const plist *l = &h->qq[st];
    double lam = h->p.lambda[st];
    for (int64_t p = 0; p < l->n; p++) {
      int i = s->iqseq[l->i[p]], j = s->iqseq[l->j[p]];
      const double *pr = l->par + 4 * p;
      vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      double r2 = 1.0 / sq(t), r = sqrt(r2), r6d = r2 * r2 * r2;
      double r6_hc = 1.0 / r6d;
      double r6 = r6_hc + pr[3];
      r6 = 1.0 / r6;
      double r12 = r6 * r6;
      double Vel = pr[2] * r, Va, Vb, dv;
      if (h->qq_soft[st][p]) {
        Vb = 0.0;
        Va = pr[0] * exp(-pr[1] / r);
        dv = r2 * (-Vel - pr[1] * Va / r) * lam;
      } else {
        Va = pr[0] * r12;
        Vb = pr[1] * r6;
        dv = r2 * (-Vel - (12.0 * Va - 6.0 * Vb) * r6 * r6_hc) * lam;
      }
      dupd(d, i, j, t, dv);
      EQ[st].qq_el = EQ[st].qq_el + Vel;
      EQ[st].qq_vdw = EQ[st].qq_vdw + Va - Vb;
    }
  }
}

/* nonbond_qqp (nonbondene.f90:5085-5153) */
static void nonbond_qqp(orc_handle *h, const double *x, double *d, qnb_qenergies *EQ) {
  const qnb_system *s = &h->s;
  for (int st = 0; st < s->nstates; st++) {
    const plist *l = &h->qqp[st];
    for (int64_t p = 0; p < l->n; p++) {
      int i = s->iqseq[l->i[p]], j = l->j[p];
      vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      double r2 = 1.0 / sq(t), r = sqrt(r2), r6 = r2 * r2 * r2;
      double Vel, Va, Vb;
      double dv = nbe_qx(r2, r, r6, l->par + 4 * p, h->p.lambda[st], &Vel, &Va, &Vb);
      dupd(d, i, j, t, dv);
      EQ[st].qp_el = EQ[st].qp_el + Vel;
      EQ[st].qp_vdw = EQ[st].qp_vdw + Va - Vb;
    }
  }
}

/* lrf_taylor (nonbondene.f90:507-571) */
static void lrf_taylor(orc_handle *h, const double *x, double *d, double *E) {
  const qnb_system *s = &h->s;
  for (int i = 0; i < s->natom; i++) {
    if (s->excl[i] || s->iqatom[i] != 0) continue;
    const double *L = h->lrf + (size_t)NLRF * h->iwhich_cgp[i];
    vec3 dr = {L[0] - x[3 * i], L[1] - x[3 * i + 1], L[2] - x[3 * i + 2]};
    vec3 phi1 = {L[4], L[5], L[6]};
    vec3 p2[3], p3[9];
    for (int k = 0; k < 3; k++) p2[k] = (vec3){L[7 + 3 * k], L[8 + 3 * k], L[9 + 3 * k]};
    for (int m = 0; m < 9; m++) p3[m] = (vec3){L[16 + 3 * m], L[17 + 3 * m], L[18 + 3 * m]};
    vec3 tmp = {dot(dr, p2[0]), dot(dr, p2[1]), dot(dr, p2[2])};
    double Vij = L[3] + dot(dr, phi1) + 0.5 * dot(dr, tmp);
    *E = *E + 0.5 * s->crg[i] * Vij;
    vec3 df;
    tmp = (vec3){dot(dr, p3[0]), dot(dr, p3[1]), dot(dr, p3[2])};
    df.x = phi1.x + dot(dr, p2[0]) + 0.5 * dot(dr, tmp);
    tmp = (vec3){dot(dr, p3[3]), dot(dr, p3[4]), dot(dr, p3[5])};
    df.y = phi1.y + dot(dr, p2[1]) + 0.5 * dot(dr, tmp);
    tmp = (vec3){dot(dr, p3[6]), dot(dr, p3[7]), dot(dr, p3[8])};
    df.z = phi1.z + dot(dr, p2[2]) + 0.5 * dot(dr, tmp);
    d[3 * i] = d[3 * i] - df.x * s->crg[i];
    d[3 * i + 1] = d[3 * i + 1] - df.y * s->crg[i];
    d[3 * i + 2] = d[3 * i + 2] - df.z * s->crg[i];
  }
}

/* restrain_solvent (nonbondene.f90:6466-6544) */
static void restrain_solvent(orc_handle *h, const double *x, double *d, double Tfree, double *E) {
  const qnb_system *s = &h->s;
  const qnb_params *p = &h->p;
  double shift = p->fk_wsphere != 0.0 ? sqrt(p->Boltz * Tfree / p->fk_wsphere) : 0.0;
  vec3 xw = {p->xwcent[0], p->xwcent[1], p->xwcent[2]};
  const int sa = s->solv_atom;
  for (int iw = s->ncgp_solute; iw < s->ncgp; iw++) {
    int jsolv = iw - s->ncgp_solute;
    for (int isolv = 0; isolv < sa; isolv++) {
      int i;
      if (sa == 3) {
        if (isolv > 0) break;
        i = s->cgp_iswitch[iw];
      } else {
        if (!s->heavy[s->nat_solute + isolv]) continue;
        i = s->nat_solute + sa * jsolv; /* as the reference: base atom, offset at the update */
      }
      if (s->excl[s->cgp_iswitch[iw]]) break;
      vec3 dv3 = {x[3 * i] - xw.x, x[3 * i + 1] - xw.y, x[3 * i + 2] - xw.z}; /* q_dist5(xwcent, x(i)) */
      double b = sqrt(sq(dv3));
      double db = b - (p->rwat - shift);
      double erst, dv;
      if (db > 0.0) {
        erst = 0.5 * p->fk_wsphere * (db * db) - p->Dwmz;
        dv = p->fk_wsphere * db / b;
      } else if (b > 0.0) {
        double fexp = exp(p->awmz * db);
        erst = p->Dwmz * (fexp * fexp - 2.0 * fexp);
        dv = -2.0 * p->Dwmz * p->awmz * (fexp - fexp * fexp) / b;
      } else {
        dv = 0.0;
        erst = 0.0;
      }
      int ii = (sa == 3) ? i : i + isolv;
      d[3 * ii] = d[3 * ii] + dv3.x * dv;
      d[3 * ii + 1] = d[3 * ii + 1] + dv3.y * dv;
      d[3 * ii + 2] = d[3 * ii + 2] + dv3.z * dv;
      *E = *E + erst;
    }
  }
}

/* dipole direction and radial direction of a solvent molecule whose first
 * atom is i (watpol, nonbondene.f90:6572-6604): a 3-site solvent takes
 * H1 + H2 - 2 O and the radial vector of O; any other takes the
 * charge-weighted sum of its positions (the charges of the first solvent
 * molecule, chg_solv) and the radial vector of its centroid. */
static void wat_vectors(const orc_handle *h, const double *x, int i, vec3 *rmu, double *rm, vec3 *rcu, double *rc) {
  const qnb_params *p = &h->p;
  const int sa = h->s.solv_atom;
  vec3 m = {0.0, 0.0, 0.0}, c;
  if (sa == 3) {
    for (int k = 1; k < 3; k++) {
      m.x = m.x + x[3 * (i + k)];
      m.y = m.y + x[3 * (i + k) + 1];
      m.z = m.z + x[3 * (i + k) + 2];
    }
    m.x = m.x - x[3 * i] * 2.0;
    m.y = m.y - x[3 * i + 1] * 2.0;
    m.z = m.z - x[3 * i + 2] * 2.0;
    c.x = x[3 * i] - p->xwcent[0];
    c.y = x[3 * i + 1] - p->xwcent[1];
    c.z = x[3 * i + 2] - p->xwcent[2];
  } else {
    vec3 g = {0.0, 0.0, 0.0};
    const double *chg = h->s.crg + h->s.nat_solute;
    for (int k = 0; k < sa; k++) {
      g.x = g.x + x[3 * (i + k)];
      g.y = g.y + x[3 * (i + k) + 1];
      g.z = g.z + x[3 * (i + k) + 2];
      m.x = m.x + x[3 * (i + k)] * chg[k];
      m.y = m.y + x[3 * (i + k) + 1] * chg[k];
      m.z = m.z + x[3 * (i + k) + 2] * chg[k];
    }
    g.x = g.x / (double)sa;
    g.y = g.y / (double)sa;
    g.z = g.z / (double)sa;
    c.x = g.x - p->xwcent[0];
    c.y = g.y - p->xwcent[1];
    c.z = g.z - p->xwcent[2];
  }
  *rm = sqrt(sq(m));
  rmu->x = m.x / *rm;
  rmu->y = m.y / *rm;
  rmu->z = m.z / *rm;
  *rc = sqrt(sq(c));
  rcu->x = c.x / *rc;
  rcu->y = c.y / *rc;
  rcu->z = c.z / *rc;
}

/* watpol (nonbondene.f90:6548-6746), md = .true.; the force on a molecule's
 * first atom is (f1 + f2) f0 and on each other atom f3 f0, for any solvent */
static void watpol(orc_handle *h, const double *x, double *d, int istep, double *E) {
  const qnb_system *s = &h->s;
  const qnb_params *p = &h->p;
  const int nw = s->nwat, nsh = p->nwpolr_shell;
  const double pi = 4.0 * atan(1.0);
  for (int k = 0; k < nsh; k++) h->n_insh[k] = 0;
  for (int iw = 0; iw < nw; iw++) {
    h->theta[iw] = 0.0;
    int i = s->nat_solute + s->solv_atom * iw;
    if (s->excl[i]) continue;
    vec3 rmu, rcu;
    double rm, rc;
    wat_vectors(h, x, i, &rmu, &rm, &rcu, &rc);
    double scp = dot(rmu, rcu);
    if (scp > 1.0) scp = 1.0;
    if (scp < -1.0) scp = -1.0;
    h->theta[iw] = acos(scp);
    h->tdum[iw] = h->theta[iw];
    if (rc > p->wshell_rout[nsh - 1] - p->wshell_dr[nsh - 1]) {
      int is;
      for (is = nsh - 1; is >= 1; is--)
        if (rc <= p->wshell_rout[is]) break;
      if (is < 1) is = 0;
      h->list_sh[(size_t)is * nw + h->n_insh[is]] = iw;
      h->n_insh[is]++;
    }
  }
  /* selection sort by theta (nonbondene.f90:6629-6648) */
  int jmin = 0;
  for (int is = 0; is < nsh; is++) {
    int n = h->n_insh[is];
    for (int il = 0; il < n; il++) {
      double tmin = 2.0 * pi;
      for (int jl = 0; jl < n; jl++) {
        int jw = h->list_sh[(size_t)is * nw + jl];
        if (h->tdum[jw] < tmin) {
          jmin = jw;
          tmin = h->theta[jw];
        }
      }
      h->nsort[(size_t)is * nw + il] = jmin;
      h->tdum[jmin] = 99999.0;
    }
  }
  /* theta_corr feedback (nonbondene.f90:6653-6666) */
  if (istep > 0 && istep % (int)lround(p->itdis) == 0) {
    for (int is = 0; is < nsh; is++) {
      h->avtheta[is] = h->avtheta[is] / p->itdis;
      h->avn_insh[is] = h->avn_insh[is] / p->itdis;
      h->theta_corr[is] = h->theta_corr[is] + h->avtheta[is] - acos(p->wshell_cstb[is]);
      h->avtheta[is] = 0.0;
      h->avn_insh[is] = 0.0;
    }
  }
  for (int is = 0; is < nsh; is++) {
    int n = h->n_insh[is];
    if (n == 0) continue;
    double avtdum = 0.0;
    for (int il = 0; il < n; il++) {
      int iw = h->nsort[(size_t)is * nw + il];
      double arg = 1.0 + (1.0 - 2.0 * (double)(il + 1)) / (double)n;
      double th0 = acos(arg);
      th0 = th0 - 3.0 * sin(th0) * p->wshell_cstb[is] / 2.0;
      if (th0 < 0.0) th0 = 0.0;
      if (th0 > pi) th0 = pi;
      avtdum = avtdum + h->theta[iw];
      double dth = h->theta[iw] - th0 + h->theta_corr[is];
      *E = *E + 0.5 * p->fkwpol * (dth * dth);
      double dv = p->fkwpol * (h->theta[iw] - th0 + h->theta_corr[is]);
      int i = s->nat_solute + s->solv_atom * iw;
      vec3 rmu, rcu;
      double rm, rc;
      wat_vectors(h, x, i, &rmu, &rm, &rcu, &rc);
      double scp = dot(rmu, rcu);
      if (scp > 1.0) scp = 1.0;
      if (scp < -1.0) scp = -1.0;
      double f0 = sin(acos(scp));
      if (fabs(f0) < QREAL_EPS) f0 = QREAL_EPS;
      f0 = -1.0 / f0;
      f0 = dv * f0;
      vec3 a = {rcu.x - rmu.x * scp, rcu.y - rmu.y * scp, rcu.z - rmu.z * scp};
      vec3 f1 = {(a.x * (-2.0)) / rm, (a.y * (-2.0)) / rm, (a.z * (-2.0)) / rm};
      vec3 f3 = {a.x / rm, a.y / rm, a.z / rm};
      vec3 f2 = {(rmu.x - rcu.x * scp) / rc, (rmu.y - rcu.y * scp) / rc, (rmu.z - rcu.z * scp) / rc};
      d[3 * i] = d[3 * i] + (f1.x + f2.x) * f0;
      d[3 * i + 1] = d[3 * i + 1] + (f1.y + f2.y) * f0;
      d[3 * i + 2] = d[3 * i + 2] + (f1.z + f2.z) * f0;
      for (int k = 1; k < s->solv_atom; k++) {
        d[3 * (i + k)] = d[3 * (i + k)] + f3.x * f0;
        d[3 * (i + k) + 1] = d[3 * (i + k) + 1] + f3.y * f0;
        d[3 * (i + k) + 2] = d[3 * (i + k) + 2] + f3.z * f0;
      }
    }
    h->avtheta[is] = h->avtheta[is] + avtdum / (double)n;
    h->avn_insh[is] = h->avn_insh[is] + (double)n;
  }
}

/* ---------------------------------------------------------------- drivers */

int orc_eval_terms(orc_handle *h, uint32_t mask, const double *x, double *d, double Tfree, int istep,
                   qnb_energies *E, qnb_qenergies *EQ) {
  const qnb_system *s = &h->s;
  /* order of pot_energy_nonbonds (potene.f90:345-377) then pot_energy
   * (potene.f90:161-177) */
  if (s->natom > s->nat_solute) {
    if (s->solvent_spc) {
      if (mask & QNB_T_WW) nonbond_ww(h, x, d, &E->ww_el, &E->ww_vdw);
      if (mask & QNB_T_QW) nonbond_qw(h, x, d, EQ);
    } else {
      if (mask & QNB_T_QW) nonbond_qw(h, x, d, EQ);
      if (mask & QNB_T_WW) nonbond_ww(h, x, d, &E->ww_el, &E->ww_vdw);
    }
    if ((mask & QNB_T_SINT) && s->n_solv_int > 0) nonbond_solvent_internal(h, x, d, &E->ww_el, &E->ww_vdw);
    if (mask & QNB_T_PW) nonbond_list(&h->pw, x, d, &E->pw_el, &E->pw_vdw);
  }
  if (mask & QNB_T_PP) nonbond_list(&h->pp, x, d, &E->pp_el, &E->pp_vdw);
  if (mask & QNB_T_QP) nonbond_qp(h, x, d, EQ);
  if ((mask & QNB_T_LRF) && h->p.use_LRF) lrf_taylor(h, x, d, &E->lrf);
  if (s->nwat > 0) {
    if (mask & QNB_T_RAD) restrain_solvent(h, x, d, Tfree, &E->solvent_radial);
    if ((mask & QNB_T_POL) && h->p.wpol_restr) watpol(h, x, d, istep, &E->water_pol);
  }
  if (mask & QNB_T_QQ) nonbond_qq(h, x, d, EQ);
  if (mask & QNB_T_QQP) nonbond_qqp(h, x, d, EQ);
  return 0;
}

int64_t orc_get_pairs(orc_handle *h, int which, int32_t *ab, int64_t cap) {
  const plist *l = NULL;
  switch (which) {
    case QNB_LIST_PP: l = &h->pp; break;
    case QNB_LIST_PW: l = &h->pw; break;
    case QNB_LIST_WW: l = &h->ww; break;
    case QNB_LIST_QP: l = &h->qp; break;
    case QNB_LIST_QW: l = &h->qw; break;
    case QNB_LIST_QQ: l = &h->qq[0]; break;
    case QNB_LIST_QQP: l = &h->qqp[0]; break;
    default: return -1;
  }
  if (ab && l->n <= cap)
    for (int64_t p = 0; p < l->n; p++) {
      ab[2 * p] = l->i[p];
      ab[2 * p + 1] = l->j[p];
    }
  return l->n;
}

int64_t orc_get_list_params(orc_handle *h, int which, double *par, int64_t cap) {
  const plist *l = NULL;
  switch (which) {
    case QNB_LIST_PP: l = &h->pp; break;
    case QNB_LIST_PW: l = &h->pw; break;
    case QNB_LIST_QP: l = &h->qp; break;
    default: return -1;
  }
  if (par && l->n <= cap) memcpy(par, l->par, sizeof(double) * l->npar * l->n);
  return l->n;
}

void orc_get_lrf(orc_handle *h, double *lrf) { memcpy(lrf, h->lrf, sizeof(double) * NLRF * h->s.ncgp); }

void orc_set_lrf(orc_handle *h, const double *lrf) { memcpy(h->lrf, lrf, sizeof(double) * NLRF * h->s.ncgp); }

void orc_get_shell_state(orc_handle *h, double *st) {
  int n = h->p.nwpolr_shell;
  for (int k = 0; k < n; k++) {
    st[k] = h->theta_corr[k];
    st[n + k] = h->avtheta[k];
    st[2 * n + k] = h->avn_insh[k];
  }
}

void orc_set_shell_state(orc_handle *h, const double *st) {
  int n = h->p.nwpolr_shell;
  for (int k = 0; k < n; k++) {
    h->theta_corr[k] = st[k];
    h->avtheta[k] = st[n + k];
    h->avn_insh[k] = st[2 * n + k];
  }
}

void orc_get_theta(orc_handle *h, double *theta, int32_t *n_insh) {
  memcpy(theta, h->theta, sizeof(double) * h->s.nwat);
  for (int k = 0; k < h->p.nwpolr_shell; k++) n_insh[k] = h->n_insh[k];
}

/* calculate_exc (nonbondene.f90:6825-6920) with the lists exc_assign builds
 * (simprep.f90:2868-2947).  For each energy array a (mask over the solute
 * atoms, type 1 full / 2 electro / 3 vdW) and state s, the Q energies of
 * eq_in[s] (qq el, vdW, qp el, vdW, qw el, vdW, Q bond, angle, torsion,
 * improper, restraint: what pot_energy left) minus the interactions of the
 * Q-atoms with the masked atoms, pair by pair in list order: first the
 * masked Q-atoms (Q-Q pairs, nbe_qq), then the masked atoms bonded to a
 * Q-atom (qconn <= 3, nbe_qx), then the other masked atoms (nbe_qx).  The
 * lists are built atom-major, Q-atom-minor; a Q-Q pair enters when its
 * qq_precomp entry is set, a bonded-neighbour pair when its qp_precomp entry
 * is set, and a plain Q-solute pair when qp_precomp(iq, iq, s) is set, the
 * entry exc_assign tests (atom index = Q-atom number).  out[a][s][7]: qq el,
 * vdW, qp el, vdW, qx el, vdW, total. */
static void exc_sub(int type, double Vel, double Va, double Vb, double *el, double *vdw) {
  if (type == 1) {
    *el = *el - Vel;
    *vdw = *vdw - (Va - Vb);
  } else if (type == 2) {
    *el = *el - Vel;
  } else if (type == 3) {
    *vdw = *vdw - (Va - Vb);
  }
}

int orc_calculate_exc(orc_handle *h, const double *x, int narr, const uint8_t *masks, const int32_t *types,
                      const double *eq_in, double *out) {
  const qnb_system *s = &h->s;
  const int nsol = s->nat_solute, nq = s->nqat, nst = s->nstates;
  for (int a = 0; a < narr; a++) {
    const uint8_t *mask = masks + (size_t)a * nsol;
    for (int st = 0; st < nst; st++) {
      const double *e = eq_in + 11 * st;
      double qqel = e[0], qqvdw = e[1], qpel = e[2], qpvdw = e[3];
      for (int ig = 0; ig < nsol; ig++) { /* Q-atoms of the group */
        if (!mask[ig] || s->iqatom[ig] == 0) continue;
        const int jq = s->iqatom[ig] - 1;
        for (int iq = 0; iq < nq; iq++) {
          const size_t k = (size_t)iq + (size_t)nq * ((size_t)jq + (size_t)nq * st);
          if (!s->qq_set[k]) continue;
          const int i = s->iqseq[iq], j = s->iqseq[jq];
          vec3 t = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
          double r2 = 1.0 / sq(t), r = sqrt(r2), r6d = r2 * r2 * r2;
          double r6_hc = 1.0 / r6d;
          double r6 = r6_hc + s->qq_sc[k];
          r6 = 1.0 / r6;
          double r12 = r6 * r6;
          double Vel = s->qq_el[k] * r, Va, Vb;
          if (s->qq_soft[k]) {
            Vb = 0.0;
            Va = s->qq_A[k] * exp(-s->qq_B[k] / r);
          } else {
            Va = s->qq_A[k] * r12;
            Vb = s->qq_B[k] * r6;
          }
          exc_sub(types[a], Vel, Va, Vb, &qqel, &qqvdw);
        }
      }
      for (int pass = 0; pass < 2; pass++) { /* bonded neighbours, then the others */
        for (int ig = 0; ig < nsol; ig++) {
          if (!mask[ig] || s->iqatom[ig] != 0) continue;
          if ((pass == 0) != (s->qconn3[ig] != 0)) continue;
          for (int iq = 0; iq < nq; iq++) {
            const size_t k = (size_t)ig + (size_t)nsol * ((size_t)iq + (size_t)nq * st);
            const size_t kt = pass == 0 ? k : (size_t)iq + (size_t)nsol * ((size_t)iq + (size_t)nq * st);
            if (!s->qp_set[kt]) continue;
            const int i = s->iqseq[iq];
            vec3 t = {x[3 * ig] - x[3 * i], x[3 * ig + 1] - x[3 * i + 1], x[3 * ig + 2] - x[3 * i + 2]};
            double r2 = 1.0 / sq(t), r = sqrt(r2), r6 = r2 * r2 * r2;
            const double pr[4] = {s->qp_A[k], s->qp_B[k], s->qp_el[k], s->qp_sc[k]};
            double Vel, Va, Vb;
            (void)nbe_qx(r2, r, r6, pr, h->p.lambda[st], &Vel, &Va, &Vb);
            exc_sub(types[a], Vel, Va, Vb, &qpel, &qpvdw);
          }
        }
      }
      double *o = out + 7 * ((size_t)a * nst + st);
      o[0] = qqel;
      o[1] = qqvdw;
      o[2] = qpel;
      o[3] = qpvdw;
      o[4] = qqel + qpel + e[4];
      o[5] = qqvdw + qpvdw + e[5];
      o[6] = e[6] + e[7] + e[8] + e[9] + o[4] + o[5] + e[10];
    }
  }
  return 0;
}
