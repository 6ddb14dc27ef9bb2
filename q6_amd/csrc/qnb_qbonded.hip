// qnb_qbonded.hip -- Q-atom bonded terms per state on the device.
//
// One block per instance computes what pot_energy's per-state loop
// (src/potene.f90:183-190) computes: for each state, the Q bonds (qbond,
// src/bondene.f90:506-551), then every angle (+ Urey-Bradley), torsion and
// improper (qangle :387-446, qurey_bradley :450-502, qtorsion :624-685,
// qimproper :555-620).  All states run at once: a thread per (bond, state),
// then a thread per (term, state).  A coupled term reads the Morse state of
// its bond (EMorseD, dMorse_i) that the sequential loop would hold at that
// point: the bond's evaluation in the latest state up to its own where the
// bond is present, else the value kept from the previous call, which lives in
// device memory between calls as the reference's module arrays do.  Gradient
// contributions go to fixed slots per state; every atom the terms touch sums
// its own slots in a fixed order, so results do not depend on thread timing,
// and adds the total to d once at the end.  The term counts are small (tens
// per FEP file), so one block of 256 threads per instance is the whole launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "qnb_host.h"

namespace qnb {

#define QB_THREADS 256

namespace {

constexpr double kQrealEps = 1.0e-10;  // QREAL_EPS, src/sizes.f90:50

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 ld3(const double* x, int i) {
  const auto g = gp(x);
  return {g[3 * i], g[3 * i + 1], g[3 * i + 2]};
}
__device__ __forceinline__ void st3(double* p, V3 v) {
  stg(p, v.x);
  stg(p + 1, v.y);
  stg(p + 2, v.z);
}

// dtheta/dx of the angle a-b-c (angle_calc, src/bonded.f90:78-118): the
// derivative uses -1/sin(theta) with |sin| floored at QREAL_EPS
struct Ang {
  double t;
  V3 a, b, c;
};
__device__ Ang angle_geo(V3 pa, V3 pb, V3 pc) {
  const V3 ab = pa - pb, bc = pc - pb;
  const double ab2 = 1.0 / dot(ab, ab), bc2 = 1.0 / dot(bc, bc);
  const double abd = sqrt(ab2), bcd = sqrt(bc2);
  double sc = dot(ab, bc) * abd * bcd;
  sc = fmin(1.0, fmax(-1.0, sc));
  Ang r;
  r.t = acos(sc);
  double inv = sin(r.t);
  if (fabs(inv) < kQrealEps) inv = kQrealEps;
  inv = -1.0 / inv;
  r.a = (bc * (abd * bcd) - ab * (ab2 * sc)) * inv;
  r.c = (ab * (abd * bcd) - bc * (bc2 * sc)) * inv;
  r.b = r.a + r.c;
  return r;
}

// dihedral a-b-c-d and its derivative vectors (torsion_calc / improper_calc,
// src/bonded.f90:151-320, the same formula)
struct Dih {
  double t;
  V3 a, b, c, d;
};
__device__ Dih dihedral_geo(V3 pa, V3 pb, V3 pc, V3 pd) {
  const V3 ab = pa - pb, bc = pc - pb, cd = pd - pc;
  const V3 n1 = cross(ab, bc), n2 = cross(-bc, cd);
  const double a2 = dot(n1, n1), b2 = dot(n2, n2);
  const double aa = sqrt(a2), bb = sqrt(b2);
  double sc = dot(n1, n2) / (aa * bb);
  sc = fmin(1.0, fmax(-1.0, sc));
  Dih r;
  r.t = acos(sc);
  if (dot(bc, cross(n1, n2)) < 0.0) r.t = -r.t;
  double inv = sin(r.t);
  if (fabs(inv) < kQrealEps) inv = kQrealEps;
  inv = -1.0 / inv;
  const double ab_ = 1.0 / (aa * bb);
  const V3 v1 = (n2 * ab_ - (n1 * (1.0 / a2)) * sc) * inv;
  const V3 v2 = (n1 * ab_ - (n2 * (1.0 / b2)) * sc) * inv;
  const V3 ca = ab - bc, db = -bc - cd;
  r.a = cross(bc, v1);
  r.b = cross(ca, v1) + cross(cd, v2);
  r.c = cross(db, v2) - cross(ab, v1);
  r.d = cross(bc, v2);
  return r;
}

// The Morse state (EMorseD, dMorse_i) a term of state s reads for bond b: in
// the reference's sequential loop, the value qbond left at the bond's last
// evaluation, i.e. in the latest state s' <= s of this call where the bond is
// present, else the value kept from the previous call
__device__ __forceinline__ const double* morse_for(const QbDev& Q, int bond, int s) {
  for (int s2 = s; s2 >= 0; s2--)
    if (Q.pres[bond * Q.ns + s2]) return Q.snap + 4 * (bond * Q.ns + s2);
  return Q.morse + 4 * bond;
}

// gamma of a coupled term: EMorseD(bond), or 1 - EMorseD with flag 1
__device__ __forceinline__ double coupling_gamma(const double* m, int flag) {
  const double g = m[0];
  return flag == 1 ? 1.0 - g : g;
}

// the coupled bond's share: d(i) += dMorse_i * e, d(j) += dMorse_j * e with
// dMorse_j = -dMorse_i (src/bondene.f90:537-541)
__device__ __forceinline__ void coupling_slots(const double* m, double e, double* slot) {
  const V3 dm = {m[1], m[2], m[3]};
  st3(slot, dm * e);
  st3(slot + 3, -dm * e);
}

__device__ void qb_bond(const QbDev& Q, double* sl, const double* x, int b, int s, double lam, double& eb) {
  double* slot = sl + 3 * Q.slot0[b];
  if (!Q.pres[b * Q.ns + s]) {
    st3(slot, {0.0, 0.0, 0.0});
    st3(slot + 3, {0.0, 0.0, 0.0});
    return;
  }
  const double* p = Q.par + 4 * (b * Q.ns + s);
  const double Dmz = p[0], amz = p[1], r0 = p[2], fk = p[3];
  const int i = Q.atoms[2 * b], j = Q.atoms[2 * b + 1];
  const V3 av = ld3(x, j) - ld3(x, i);
  const double dist = sqrt(dot(av, av));
  const double db = dist - r0;
  const double fexp = exp(-amz * db);
  eb += Dmz * (fexp * fexp - 2.0 * fexp) + 0.5 * fk * db * db;
  const double dv = (2.0 * Dmz * amz * (fexp - fexp * fexp) + fk * db) * lam / dist;
  st3(slot, -av * dv);
  st3(slot + 3, av * dv);
  double* m = Q.snap + 4 * (b * Q.ns + s);
  if (db > 0.0) {
    m[0] = -(fexp * fexp - 2.0 * fexp);
    const double c = 2.0 * amz * (fexp - fexp * fexp) * lam / dist;
    m[1] = c * av.x;
    m[2] = c * av.y;
    m[3] = c * av.z;
  } else {
    m[0] = 1.0;
    m[1] = m[2] = m[3] = 0.0;
  }
}

// angle term a (+ its Urey-Bradley term): energies into ea
__device__ void qb_angle(const QbDev& Q, double* sl, const double* x, int a, int s, double lam, double& ea) {
  const int t = Q.nb + a;
  double* slot = sl + 3 * Q.slot0[t];
  const int* c = Q.cpl + 4 * a;
  const int i = Q.atoms[2 * Q.nb + 3 * a], j = Q.atoms[2 * Q.nb + 3 * a + 1], k = Q.atoms[2 * Q.nb + 3 * a + 2];
  const bool on = Q.pres[t * Q.ns + s];
  const double* p = Q.par + 4 * (t * Q.ns + s);
  const V3 xi = ld3(x, i), xk = ld3(x, k);
  V3 fa = {0.0, 0.0, 0.0}, fb = fa, fc = fa;
  double ae = 0.0;
  const double* m0 = c[0] >= 0 ? morse_for(Q, c[0], s) : nullptr;
  if (on) {
    const double g = m0 ? coupling_gamma(m0, c[1]) : 1.0;
    const Ang A = angle_geo(xi, ld3(x, j), xk);
    const double da = A.t - p[1];
    ae = 0.5 * p[0] * da * da;
    ea += ae * g;
    const double dv = g * p[0] * da * lam;
    fa = A.a * dv;
    fb = -(A.b * dv);
    fc = A.c * dv;
  }
  st3(slot, fa);
  st3(slot + 3, fb);
  st3(slot + 6, fc);
  int q = 3;
  if (c[0] >= 0) {
    if (on) coupling_slots(m0, ae, slot + 3 * q);
    else {
      st3(slot + 3 * q, {0.0, 0.0, 0.0});
      st3(slot + 3 * q + 3, {0.0, 0.0, 0.0});
    }
    q += 2;
  }
  if (!Q.urey) return;
  // qurey_bradley: skipped when the term is absent or ureyfk is zero
  const bool uon = on && p[2] != 0.0;
  V3 ui = {0.0, 0.0, 0.0}, uk = ui;
  double eu = 0.0;
  const double* m2 = c[2] >= 0 ? morse_for(Q, c[2], s) : nullptr;
  if (uon) {
    const double g = m2 ? coupling_gamma(m2, c[3]) : 1.0;
    const V3 av = xk - xi;
    const double dist = sqrt(dot(av, av));
    const double ru = dist - p[3];
    eu = p[2] * ru * ru;
    ea += eu * g;
    const double du = g * 2.0 * (p[2] * ru / dist) * lam;
    ui = -(av * du);
    uk = av * du;
  }
  st3(slot + 3 * q, ui);
  st3(slot + 3 * q + 3, uk);
  q += 2;
  if (c[2] >= 0) {
    if (uon) coupling_slots(m2, eu, slot + 3 * q);
    else {
      st3(slot + 3 * q, {0.0, 0.0, 0.0});
      st3(slot + 3 * q + 3, {0.0, 0.0, 0.0});
    }
  }
}

// torsion (improper = false) or improper term u of its kind
__device__ void qb_dihedral(const QbDev& Q, double* sl, const double* x, int u, bool improper, int s, double lam,
                            double& e) {
  const int t = Q.nb + Q.na + (improper ? Q.nt : 0) + u;
  double* slot = sl + 3 * Q.slot0[t];
  const int* c = Q.cpl + 4 * (t - Q.nb);
  const int* at = Q.atoms + 2 * Q.nb + 3 * Q.na + 4 * ((improper ? Q.nt : 0) + u);
  const bool on = Q.pres[t * Q.ns + s];
  V3 f[4] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  double pe = 0.0;
  const double* m0 = c[0] >= 0 ? morse_for(Q, c[0], s) : nullptr;
  if (on) {
    const double* p = Q.par + 4 * (t * Q.ns + s);
    const double g = m0 ? coupling_gamma(m0, c[1]) : 1.0;
    const Dih D = dihedral_geo(ld3(x, at[0]), ld3(x, at[1]), ld3(x, at[2]), ld3(x, at[3]));
    double dv;
    if (!improper) {
      const double arg = p[1] * D.t - p[2];
      pe = p[0] * (1.0 + cos(arg));
      dv = -p[1] * p[0] * sin(arg) * g * lam;
    } else {
      constexpr double twopi = 6.283185307179586;
      double arg = D.t - p[1];
      arg = arg - twopi * round(arg / twopi);  // Fortran nint: halves away from zero
      const double k = p[0] * arg;
      pe = 0.5 * k * arg;
      dv = k * g * lam;
    }
    e += pe * g;
    f[0] = D.a * dv;
    f[1] = D.b * dv;
    f[2] = D.c * dv;
    f[3] = D.d * dv;
  }
  for (int r = 0; r < 4; r++) st3(slot + 3 * r, f[r]);
  if (c[0] >= 0) {
    if (on) coupling_slots(m0, pe, slot + 12);
    else {
      st3(slot + 12, {0.0, 0.0, 0.0});
      st3(slot + 15, {0.0, 0.0, 0.0});
    }
  }
}

}  // namespace

// blockIdx.x = instance.  Every state at once: first the Q bonds of all
// states, each present (bond, state) writing its Morse snapshot; after a
// barrier every (term, state), each coupled term reading the snapshot
// morse_for() picks, which is what the sequential per-state loop gives it;
// after a second barrier the kept Morse state takes each bond's last present
// state, the per-state energies are summed in term order and every touched
// atom adds its slots of all states to d.  Each (term, state) writes its own
// energy and slots, so the result does not depend on timing.
__global__ __launch_bounds__(QB_THREADS) void k_qbonded(const DevSys* __restrict__ systems,
                                                       const QbDev* __restrict__ qbs) {
  const QbDev& Q = qbs[blockIdx.x];
  const int nterm = Q.nb + Q.na + Q.nt + Q.ni;
  if (nterm == 0) return;
  const DevSys& S = systems[blockIdx.x];
  const double* x = S.x;
  const int ns = Q.ns;
  for (int k = threadIdx.x; k < Q.nb * ns; k += blockDim.x) {
    const int b = k / ns, s = k - b * ns;
    double e = 0.0;
    qb_bond(Q, Q.slots + 3 * (size_t)Q.nslot * s, x, b, s, Q.lambda[s], e);
    stg(Q.eterm + k, e);
  }
  __syncthreads();
  const int nrest = Q.na + Q.nt + Q.ni;
  for (int k = threadIdx.x; k < nrest * ns; k += blockDim.x) {
    const int t = k / ns, s = k - t * ns;
    double* sl = Q.slots + 3 * (size_t)Q.nslot * s;
    double e = 0.0;
    if (t < Q.na) qb_angle(Q, sl, x, t, s, Q.lambda[s], e);
    else if (t < Q.na + Q.nt) qb_dihedral(Q, sl, x, t - Q.na, false, s, Q.lambda[s], e);
    else qb_dihedral(Q, sl, x, t - Q.na - Q.nt, true, s, Q.lambda[s], e);
    stg(Q.eterm + (size_t)Q.nb * ns + k, e);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < Q.nb; b += blockDim.x)
    for (int s = ns - 1; s >= 0; s--)
      if (Q.pres[b * ns + s]) {
        const double* m = Q.snap + 4 * (b * ns + s);
        for (int c = 0; c < 4; c++) stg(Q.morse + 4 * b + c, m[c]);
        break;
      }
  // energies: wave w sums kind k of state s for (s, k) = w, w + NW, ...
  const int wave = threadIdx.x / QNB_WAVE, lane = threadIdx.x % QNB_WAVE;
  const int lo[5] = {0, Q.nb, Q.nb + Q.na, Q.nb + Q.na + Q.nt, nterm};
  for (int sk = wave; sk < 4 * ns; sk += QB_THREADS / QNB_WAVE) {
    const int s = sk / 4, kind = sk - 4 * s;
    double v = 0.0;
    for (int t = lo[kind] + lane; t < lo[kind + 1]; t += QNB_WAVE) v += ldg(Q.eterm + (size_t)t * ns + s);
    v = wave_sum(v);
    if (lane == 0) stg(Q.e + sk, v);
  }
  for (int a = threadIdx.x; a < Q.nat; a += blockDim.x) {
    V3 sum = {0.0, 0.0, 0.0};
    for (int s = 0; s < ns; s++) {
      const double* sl = Q.slots + 3 * (size_t)Q.nslot * s;
      for (int m = Q.coff[a]; m < Q.coff[a + 1]; m++) sum = sum + ld3(sl, Q.cslot[m]);
    }
    const int i = Q.catom[a];
    st3(S.d + 3 * i, ld3(S.d, i) + sum);
  }
}

void launch_qbonded(hipStream_t st, int B, const DevSys* dsys, const QbDev* dqb) {
  hipLaunchKernelGGL(k_qbonded, dim3(B), dim3(QB_THREADS), 0, st, dsys, dqb);
}

// ------------------------------------------------------------------ host

namespace {

// coupling of term `term` in a table [n][3]: the first match (qangle's loop
// exits, src/bondene.f90:404-413) or the last (qurey_bradley, qtorsion and
// qimproper keep scanning)
void find_coupling(const int32_t* tab, int n, int term, bool first, int& bond, int& flag) {
  bond = -1;
  flag = 0;
  for (int m = 0; m < n; m++)
    if (tab[3 * m] == term) {
      bond = tab[3 * m + 1];
      flag = tab[3 * m + 2] == 1 ? 1 : 0;  // only 1 selects 1 - EMorseD (src/bondene.f90:415)
      if (first) return;
    }
}

}  // namespace

int qbonded_build(const qnb_qbonded* q, int nstates, int natom, const double* lambda, DevAlloc& M,
                  std::vector<void*>& owned, QbDev& out) {
  out = QbDev{};
  out.ns = nstates;
  for (int s = 0; s < QNB_MAXST; s++) out.lambda[s] = s < nstates ? lambda[s] : 0.0;
  if (!q || q->nbond + q->nangle + q->ntorsion + q->nimproper == 0) return 0;
  const int nb = q->nbond, na = q->nangle, nt = q->ntorsion, ni = q->nimproper;
  if (nb < 0 || na < 0 || nt < 0 || ni < 0) return fail(QNB_EINVAL, "negative Q bonded term count");
  if (nb + na + nt + ni > QNB_QB_MAX_TERMS || nb > QNB_QB_MAX_BONDS)
    return fail(QNB_EUNSUPPORTED, "Q bonded set too large: %d terms, %d bonds (max %d, %d)", nb + na + nt + ni, nb,
                QNB_QB_MAX_TERMS, QNB_QB_MAX_BONDS);
  const int ncp[3] = {q->nangle_coupl, q->ntorsion_coupl, q->nimproper_coupl};
  const int32_t* cpt[3] = {q->angle_coupl, q->torsion_coupl, q->improper_coupl};
  const int nterm_k[3] = {na, nt, ni};
  for (int k = 0; k < 3; k++) {
    if (ncp[k] < 0 || (ncp[k] > 0 && !cpt[k])) return fail(QNB_EINVAL, "Q bonded coupling table missing");
    for (int m = 0; m < ncp[k]; m++) {
      const int t = cpt[k][3 * m], b = cpt[k][3 * m + 1];
      if (t < 0 || t >= nterm_k[k] || b < 0 || b >= nb)
        return fail(QNB_EINVAL, "Q bonded coupling %d of kind %d out of range (term %d, bond %d)", m, k, t, b);
    }
  }
  // per-kind tables
  struct Kind {
    int n, nat, nlib, npar;
    const int32_t *idx, *cod;
    const double* lib;
    const char* name;
  };
  const Kind kinds[4] = {{nb, 2, q->nbond_lib, 4, q->bond_ij, q->bond_cod, q->bond_lib, "bond"},
                         {na, 3, q->nangle_lib, 4, q->angle_ijk, q->angle_cod, q->angle_lib, "angle"},
                         {nt, 4, q->ntorsion_lib, 3, q->torsion_ijkl, q->torsion_cod, q->torsion_lib, "torsion"},
                         {ni, 4, q->nimproper_lib, 2, q->improper_ijkl, q->improper_cod, q->improper_lib, "improper"}};
  const int nterm = nb + na + nt + ni;
  std::vector<int> atoms;
  std::vector<double> par((size_t)nterm * nstates * 4, 0.0);
  std::vector<uint8_t> pres((size_t)nterm * nstates, 0);
  int t = 0;
  for (const Kind& K : kinds) {
    if (K.n > 0 && (!K.idx || !K.cod || (K.nlib > 0 && !K.lib)))
      return fail(QNB_EINVAL, "Q %s arrays missing", K.name);
    for (int u = 0; u < K.n; u++, t++) {
      for (int r = 0; r < K.nat; r++) {
        const int a = K.idx[K.nat * u + r];
        if (a < 0 || a >= natom) return fail(QNB_EINVAL, "Q %s %d: atom %d out of range", K.name, u, a);
        atoms.push_back(a);
      }
      for (int s = 0; s < nstates; s++) {
        const int c = K.cod[u * nstates + s];
        if (c < 0 || c > K.nlib) return fail(QNB_EINVAL, "Q %s %d: code %d in state %d out of range", K.name, u, c, s);
        if (c == 0) continue;
        pres[(size_t)t * nstates + s] = 1;
        for (int r = 0; r < K.npar; r++) par[((size_t)t * nstates + s) * 4 + r] = K.lib[(c - 1) * K.npar + r];
      }
    }
  }
  // couplings and slots
  std::vector<int> cpl(4 * (size_t)(na + nt + ni), -1), slot0(nterm), slot_atom;
  auto add_slot = [&](int atom) { slot_atom.push_back(atom); };
  t = 0;
  for (int b = 0; b < nb; b++, t++) {
    slot0[t] = (int)slot_atom.size();
    add_slot(atoms[2 * b]);
    add_slot(atoms[2 * b + 1]);
  }
  for (int a = 0; a < na; a++, t++) {
    int* c = &cpl[4 * (size_t)a];
    find_coupling(q->angle_coupl, q->nangle_coupl, a, true, c[0], c[1]);
    find_coupling(q->angle_coupl, q->nangle_coupl, a, false, c[2], c[3]);
    slot0[t] = (int)slot_atom.size();
    for (int r = 0; r < 3; r++) add_slot(atoms[2 * nb + 3 * a + r]);
    if (c[0] >= 0) {
      add_slot(atoms[2 * c[0]]);
      add_slot(atoms[2 * c[0] + 1]);
    }
    if (q->urey_bradley) {
      add_slot(atoms[2 * nb + 3 * a]);
      add_slot(atoms[2 * nb + 3 * a + 2]);
      if (c[2] >= 0) {
        add_slot(atoms[2 * c[2]]);
        add_slot(atoms[2 * c[2] + 1]);
      }
    }
  }
  for (int k = 1; k < 3; k++)
    for (int u = 0; u < nterm_k[k]; u++, t++) {
      int* c = &cpl[4 * (size_t)(t - nb)];
      find_coupling(cpt[k], ncp[k], u, false, c[0], c[1]);
      slot0[t] = (int)slot_atom.size();
      const int base = 2 * nb + 3 * na + 4 * (k == 2 ? nt : 0) + 4 * u;
      for (int r = 0; r < 4; r++) add_slot(atoms[base + r]);
      if (c[0] >= 0) {
        add_slot(atoms[2 * c[0]]);
        add_slot(atoms[2 * c[0] + 1]);
      }
    }
  // distinct atoms and their slots, in atom order then slot order
  std::vector<int> catom(slot_atom.begin(), slot_atom.end());
  std::sort(catom.begin(), catom.end());
  catom.erase(std::unique(catom.begin(), catom.end()), catom.end());
  const int nat = (int)catom.size(), nslot = (int)slot_atom.size();
  std::vector<int> coff(nat + 1, 0), cslot(nslot);
  for (int sl = 0; sl < nslot; sl++) {
    const int a = (int)(std::lower_bound(catom.begin(), catom.end(), slot_atom[sl]) - catom.begin());
    coff[a + 1]++;
  }
  for (int a = 0; a < nat; a++) coff[a + 1] += coff[a];
  {
    std::vector<int> fill(coff.begin(), coff.end() - 1);
    for (int sl = 0; sl < nslot; sl++) {
      const int a = (int)(std::lower_bound(catom.begin(), catom.end(), slot_atom[sl]) - catom.begin());
      cslot[fill[a]++] = sl;
    }
  }
  // device copies
  out.nb = nb;
  out.na = na;
  out.nt = nt;
  out.ni = ni;
  out.urey = q->urey_bradley ? 1 : 0;
  out.nslot = nslot;
  out.nat = nat;
  int rc = 0;
  auto up = [&](auto** dst, const auto& v) {
    using T = std::remove_const_t<std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>>;
    const T* p = nullptr;
    if (!rc) rc = M.upload(&p, v.data(), v.size());
    if (!rc) owned.push_back((void*)p);
    *dst = p;
  };
  up(&out.atoms, atoms);
  up(&out.par, par);
  up(&out.pres, pres);
  up(&out.cpl, cpl);
  up(&out.slot0, slot0);
  up(&out.coff, coff);
  up(&out.cslot, cslot);
  up(&out.catom, catom);
  auto scratch = [&](double** dst, size_t n) {
    if (!rc) rc = M.alloc(dst, n);
    if (!rc) {
      owned.push_back(*dst);
      if (hipMemset(*dst, 0, n * sizeof(double)) != hipSuccess) rc = fail(QNB_EDEVICE, "hipMemset failed");
    }
  };
  scratch(&out.morse, 4 * (size_t)nb);
  scratch(&out.slots, 3 * (size_t)nslot * nstates);
  scratch(&out.snap, 4 * (size_t)nb * nstates);
  scratch(&out.eterm, (size_t)nterm * nstates);
  scratch(&out.e, 4 * (size_t)nstates);
  return rc;
}

void qbonded_free(DevAlloc& M, std::vector<void*>& owned) {
  for (void* p : owned) M.release(p);
  owned.clear();
}

}  // namespace qnb
