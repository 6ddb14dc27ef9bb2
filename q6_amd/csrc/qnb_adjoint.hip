// qnb_adjoint.hip -- adjoint of the per-state Q-atom energies over stored
// frames: for every frame f, state s and Q-atom iq the kernel returns
//   dE_s/d(charge of iq in s), dE_s/d(LJ A and B factors of iq in s) for
//   normal and 1-4 pairs, and iq's share of E_s (el, vdw),
// where E_s = EQ(s)%qq + %qp + %qw (src/potene.f90:224-229) recomputed from the
// frame.  Each Q table entry is a product of a Q-side parameter and a partner
// factor (precompute_set_values_qp/qq/qw, src/simprep.f90:3420-3589); the
// caller passes the derivative of every entry with respect to its Q-side
// parameter (the tangent tables), so the kernel only sums tangent x radial
// function over the pairs.  The Q lists are recomputed per frame with the
// criteria of nbqplist/nbqwlist (src/nonbondene.f90:3647-3742, 3862-3944):
// nothing from the MD's lists is stored.  Pair math in R (float for the
// FP32 backward, double for validation), every sum in FP64.
#include <hip/hip_runtime.h>

#include "qnb_host.h"

namespace qnb {

template <typename R, int NS>
__device__ __forceinline__ void adj_pair(double (&g)[NS][QNB_ADJ_NV], int s, R r2, R r, R r6in, R A, R B, R el,
                                         R sc, R tq, R tA, R tB, bool is14, bool lj, bool energy) {
  R Vel = el * r;
  R rq = tq * r;
  double e_vdw = 0.0, gA = 0.0, gB = 0.0;
  if (lj) {
    R r6_hc = recip(r6in);
    R r6 = recip(r6_hc + sc);
    R r12 = r6 * r6;
    R Va = A * r12, Vb = B * r6;
    e_vdw = (double)Va - (double)Vb;
    gA = (double)(tA * r12);
    gB = -(double)(tB * r6);
  }
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (k == s) {
      g[k][ADJ_DQ] += (double)rq;
      g[k][is14 ? ADJ_DA14 : ADJ_DA] += gA;
      g[k][is14 ? ADJ_DB14 : ADJ_DB] += gB;
      if (energy) {
        g[k][ADJ_EEL] += (double)Vel;
        g[k][ADJ_EVDW] += e_vdw;
      }
    }
}

// one wavefront per (frame, Q-atom); grid.y = instance
template <typename R, int NS>
__global__ __launch_bounds__(256) void k_adjoint(const DevSys* __restrict__ systems, const AdjDev* __restrict__ adjs) {
  const DevSys& S = systems[blockIdx.y];
  const AdjDev& A = adjs[blockIdx.y];
  const int nq = S.nqat, ns = S.nstates, sa = S.sa, nsol = S.nat_solute;
  if (nq <= 0) return;
  const int wv = blockIdx.x * (blockDim.x / QNB_WAVE) + threadIdx.x / QNB_WAVE;
  const int lane = threadIdx.x % QNB_WAVE;
  const int f = wv / nq, iq = wv % nq;
  if (f >= A.nframes) return;
  const double* __restrict__ x = A.frames + (size_t)f * 3 * S.natom;
  const int i = S.iqseq[iq];
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  double g[NS][QNB_ADJ_NV];
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int c = 0; c < QNB_ADJ_NV; c++) g[s][c] = 0.0;

  // Q-water: every water whose oxygen lies within Rcq of the Q-list centre
  for (int w = lane; w < S.nwat; w += QNB_WAVE) {
    const int o = nsol + sa * w;
    if (S.excl[o] || dist4p(x, o, S.xp) > S.rcq2) continue;
    for (int k = 0; k < sa; k++) {
      const int j = o + k;
      const R ux = (R)(x[3 * j] - xi0), uy = (R)(x[3 * j + 1] - xi1), uz = (R)(x[3 * j + 2] - xi2);
      R r2, r;
      rinv(ux * ux + uy * uy + uz * uz, r2, r);
      const R r6in = r2 * r2 * r2;
      const bool lj = k == 0 || !S.spc;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        if (s >= ns) break;
        const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * s);
        adj_pair<R, NS>(g, s, r2, r, r6in, (R)S.qwA[t], (R)S.qwB[t], (R)S.qwEl[t], (R)S.qwSc[t], (R)A.qw_dq[t],
                        (R)A.qw_dA[t], (R)A.qw_dB[t], false, lj, true);
      }
    }
  }
  // Q-solute: atoms of solute groups inside Rcq (switch atom, or any atom
  // with iuse_switch_atom = 0), not 1-2/1-3 bonded to any Q-atom
  for (int j = lane; j < nsol; j += QNB_WAVE) {
    if (S.qconn3[j] || S.atom_group[j] < 0) continue;
    if (!qp_group_in(S, x, S.atom_group[j])) continue;
    const R ux = (R)(x[3 * j] - xi0), uy = (R)(x[3 * j + 1] - xi1), uz = (R)(x[3 * j + 2] - xi2);
    R r2, r;
    rinv(ux * ux + uy * uy + uz * uz, r2, r);
    const R r6in = r2 * r2 * r2;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (s >= ns) break;
      const size_t t = (size_t)j + (size_t)nsol * ((size_t)iq + (size_t)nq * s);
      adj_pair<R, NS>(g, s, r2, r, r6in, (R)S.qpA[t], (R)S.qpB[t], (R)S.qpEl[t], (R)S.qpSc[t], (R)A.qp_dq[t],
                      (R)A.qp_dA[t], (R)A.qp_dB[t], A.qp_is14[t] != 0, true, true);
    }
  }
  // Q-Q pairs with iq as either member; the energy is counted by the lower one
  for (int p = lane; p < S.nqq; p += QNB_WAVE) {
    const int ia = S.qq_iq[p], ja = S.qq_jq[p];
    if (ia != iq && ja != iq) continue;
    const int st = S.qq_st[p];
    const int a = S.iqseq[ia], b = S.iqseq[ja];
    const R ux = (R)(x[3 * b] - x[3 * a]), uy = (R)(x[3 * b + 1] - x[3 * a + 1]), uz = (R)(x[3 * b + 2] - x[3 * a + 2]);
    R r2, r;
    rinv(ux * ux + uy * uy + uz * uz, r2, r);
    const double* pr = S.qq_par + 4 * p;
    const size_t t = (size_t)ia + (size_t)nq * ((size_t)ja + (size_t)nq * st);
    const bool mine = ia == iq;
    const R tq = (R)(mine ? A.qq_dq_i[t] : A.qq_dq_j[t]);
    if (S.qq_soft[p]) {
      // exponential soft pair: its A, B are not LJ factors (no LJ tangent)
      R Vel = (R)pr[2] * r;
      R Va = (R)pr[0] * exp(-(R)pr[1] / r);
#pragma unroll
      for (int k = 0; k < NS; k++)
        if (k == st) {
          g[k][ADJ_DQ] += (double)(tq * r);
          if (mine) {
            g[k][ADJ_EEL] += (double)Vel;
            g[k][ADJ_EVDW] += (double)Va;
          }
        }
    } else {
      const R tA = (R)(mine ? A.qq_dA_i[t] : A.qq_dA_j[t]);
      const R tB = (R)(mine ? A.qq_dB_i[t] : A.qq_dB_j[t]);
      adj_pair<R, NS>(g, st, r2, r, r2 * r2 * r2, (R)pr[0], (R)pr[1], (R)pr[2], (R)pr[3], tq, tA, tB,
                      A.qq_is14[t] != 0, true, mine);
    }
  }
  // Q-bonded-neighbour pairs of iq (entries of the qp tables)
  for (int p = lane; p < S.nqqp; p += QNB_WAVE) {
    if (S.qqp_iq[p] != iq) continue;
    const int j = S.qqp_j[p], st = S.qqp_st[p];
    const R ux = (R)(x[3 * j] - xi0), uy = (R)(x[3 * j + 1] - xi1), uz = (R)(x[3 * j + 2] - xi2);
    R r2, r;
    rinv(ux * ux + uy * uy + uz * uz, r2, r);
    const double* pr = S.qqp_par + 4 * p;
    const size_t t = (size_t)j + (size_t)nsol * ((size_t)iq + (size_t)nq * st);
    adj_pair<R, NS>(g, st, r2, r, r2 * r2 * r2, (R)pr[0], (R)pr[1], (R)pr[2], (R)pr[3], (R)A.qp_dq[t],
                    (R)A.qp_dA[t], (R)A.qp_dB[t], A.qp_is14[t] != 0, true, true);
  }
  double* out = A.out + (size_t)f * ns * nq * QNB_ADJ_NV;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if (s >= ns) break;
#pragma unroll
    for (int c = 0; c < QNB_ADJ_NV; c++) {
      const double v = wave_sum(g[s][c]);
      if (lane == 0) out[((size_t)s * nq + iq) * QNB_ADJ_NV + c] = v;
    }
  }
}

// Copies a caller's tangent tables to the device (sizes from the system).
int adjoint_upload(DevAlloc& M, const DevSys& S, const qnb_qtangent* t, AdjDev& A) {
  if (!t) return fail(QNB_EINVAL, "null tangent tables");
  const size_t nq = S.nqat, ns = S.nstates, nsol = S.nat_solute, sa = S.sa;
  const size_t np = nsol * nq * ns, nw = S.nwat > 0 ? nq * sa * ns : 0, nqq = nq * nq * ns;
  int rc = 0;
#define UPT(field, n)                                                                     \
  do {                                                                                    \
    if ((n) > 0 && !t->field) return fail(QNB_EINVAL, "tangent table " #field " missing"); \
    if ((rc = M.upload(&A.field, t->field, (n)))) return rc;                              \
  } while (0)
  UPT(qp_dq, np);
  UPT(qp_dA, np);
  UPT(qp_dB, np);
  UPT(qp_is14, np);
  UPT(qw_dq, nw);
  UPT(qw_dA, nw);
  UPT(qw_dB, nw);
  UPT(qq_dq_i, nqq);
  UPT(qq_dA_i, nqq);
  UPT(qq_dB_i, nqq);
  UPT(qq_dq_j, nqq);
  UPT(qq_dA_j, nqq);
  UPT(qq_dB_j, nqq);
  UPT(qq_is14, nqq);
#undef UPT
  return 0;
}

template <typename R>
static void launch_adj_r(hipStream_t st, int ns_max, int waves, int B, const DevSys* dsys, const AdjDev* dadj) {
  const int wpb = 4;
  dim3 grid((waves + wpb - 1) / wpb, B), block(QNB_WAVE * wpb);
  if (ns_max <= 1) hipLaunchKernelGGL((k_adjoint<R, 1>), grid, block, 0, st, dsys, dadj);
  else if (ns_max <= 2) hipLaunchKernelGGL((k_adjoint<R, 2>), grid, block, 0, st, dsys, dadj);
  else if (ns_max <= 4) hipLaunchKernelGGL((k_adjoint<R, 4>), grid, block, 0, st, dsys, dadj);
  else hipLaunchKernelGGL((k_adjoint<R, 8>), grid, block, 0, st, dsys, dadj);
}

// waves = max over instances of nframes * nqat
void launch_adjoint(hipStream_t st, int prec, int ns_max, long waves, int B, const DevSys* dsys, const AdjDev* dadj) {
  if (waves <= 0 || B <= 0) return;
  if (prec == QNB_PREC_MIXED) launch_adj_r<float>(st, ns_max, (int)waves, B, dsys, dadj);
  else launch_adj_r<double>(st, ns_max, (int)waves, B, dsys, dadj);
}

}  // namespace qnb
