// qnb_exc.hip -- group contributions (calculate_exc, src/nonbondene.f90:
// 6825-6920): energy arrays that repeat the per-state Q-atom energies with the
// interactions between the Q-atoms and a group of solute atoms removed.
//
// The pair lists are those exc_assign builds once (src/simprep.f90:2868-2947),
// with no cut-off: for each array and state, atom by atom over the group in
// index order and Q-atom by Q-atom,
//   * a group atom that is itself Q-atom jq: the pair (iq, jq) when its
//     qq_precomp entry is set (nbe_qq, soft pairs by the exponential form);
//   * a group atom bonded or angled to a Q-atom (qconn <= 3): (iq, atom) when
//     its qp_precomp entry is set (nbe_qx);
//   * any other group atom: (iq, atom) when qp_precomp(iq, iq, s) is set --
//     the entry exc_assign tests there, the atom index being the Q-atom
//     number -- with the pair's own parameters (nbe_qx).
// calculate_exc walks the Q-Q pairs, then the bonded neighbours, then the
// others, subtracting each pair's energies from the state's totals (el and
// vdW for a full array, one of them for an electrostatic or a vdW array) and
// forms qx and the total as pot_energy does.  One thread per (array, state)
// repeats that walk in FP64 in the same order (exc_walk, qnb_dev.h); the batch
// runs it after the energy record of each sampled step (k_exc_record,
// qnb_batch.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "qnb_host.h"

#pragma clang fp contract(off)

namespace qnb {

int exc_build(const qnb_system* s, int narr, const uint8_t* mask, const int32_t* type, ExcHost& X) {
  const int nsol = s->nat_solute, nq = s->nqat, ns = s->nstates;
  if (narr < 0 || (narr > 0 && (!mask || !type))) return fail(QNB_EINVAL, "groups: masks and types needed");
  if (narr > 0 && nq == 0) return fail(QNB_EINVAL, "groups: the system has no Q-atoms");
  for (int a = 0; a < narr; a++)
    if (type[a] < QNB_GC_FULL || type[a] > QNB_GC_VDW) return fail(QNB_EINVAL, "group type %d", type[a]);
  X.narr = narr;
  X.ns = ns;
  X.type.assign(type, type + narr);
  X.off.assign(1, 0);
  X.ent.clear();
  for (int a = 0; a < narr; a++) {
    const uint8_t* m = mask + (size_t)a * nsol;
    for (int st = 0; st < ns; st++) {
      for (int ig = 0; ig < nsol; ig++) {  // Q-atoms of the group
        if (!m[ig] || s->iqatom[ig] == 0) continue;
        const int jq = s->iqatom[ig] - 1;
        for (int iq = 0; iq < nq; iq++) {
          const size_t k = (size_t)iq + (size_t)nq * ((size_t)jq + (size_t)nq * st);
          if (!s->qq_set[k]) continue;
          X.ent.push_back(ExcEnt{0, s->iqseq[iq], s->iqseq[jq], s->qq_soft[k] ? 1 : 0, s->qq_A[k], s->qq_B[k],
                                 s->qq_el[k], s->qq_sc[k]});
        }
      }
      for (int pass = 0; pass < 2; pass++)  // bonded neighbours, then the others
        for (int ig = 0; ig < nsol; ig++) {
          if (!m[ig] || s->iqatom[ig] != 0 || (pass == 0) != (s->qconn3[ig] != 0)) continue;
          for (int iq = 0; iq < nq; iq++) {
            const size_t k = (size_t)ig + (size_t)nsol * ((size_t)iq + (size_t)nq * st);
            const size_t kt = pass == 0 ? k : (size_t)iq + (size_t)nsol * ((size_t)iq + (size_t)nq * st);
            if (!s->qp_set[kt]) continue;
            X.ent.push_back(ExcEnt{1, s->iqseq[iq], ig, 0, s->qp_A[k], s->qp_B[k], s->qp_el[k], s->qp_sc[k]});
          }
        }
      X.off.push_back((int)X.ent.size());
    }
  }
  return 0;
}

int exc_upload(const ExcHost& X, ExcDev& D, DevAlloc& M) {
  D.narr = X.narr;
  D.ns = X.ns;
  D.off = nullptr;
  D.ent = nullptr;
  D.type = nullptr;
  if (X.narr == 0) return 0;
  if (int rc = M.upload(&D.off, X.off.data(), X.off.size())) return rc;
  if (int rc = M.upload(&D.type, X.type.data(), X.type.size())) return rc;
  if (!X.ent.empty())
    if (int rc = M.upload(&D.ent, X.ent.data(), X.ent.size())) return rc;
  return 0;
}

// single system: eq_in[s][11], out[a][s][7]
__global__ void k_exc_eval(ExcDev D, const double* __restrict__ x, const double* __restrict__ eq_in, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.narr * D.ns) return;
  const int a = t / D.ns, st = t % D.ns;
  exc_walk(D, x, a, st, eq_in + 11 * st, out + (size_t)QNB_GC_NV * t);
}

void launch_exc_eval(hipStream_t st, const ExcDev& D, const double* x, const double* eq_in, double* out) {
  const int n = D.narr * D.ns;
  if (n > 0) hipLaunchKernelGGL(k_exc_eval, dim3((n + 63) / 64), dim3(64), 0, st, D, x, eq_in, out);
}

}  // namespace qnb
