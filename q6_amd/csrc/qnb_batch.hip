// qnb_batch.hip -- device-resident MD of a batch of instances (lambda windows
// or independent systems) on top of the nonbonded engine: step kernels
// (solute bonded terms, leap-frog + Berendsen, SHAKE of rigid waters,
// temperature, energy sampling) and the batch C ABI.  One launch of each
// kernel covers every instance (grid.y or grid.x = instance), and a
// list-update cycle of nbcycle steps is captured once as a HIP graph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "qnb_host.h"

namespace qnb {
template <typename R, bool ANY>
__global__ void k_lists_lrf(const DevSys*);
__global__ void k_qlists(const DevSys*);
__global__ void k_pack_xq(const DevSys*);
__global__ void k_pw_first(const DevSys*);
template <typename R, int NS, bool E, int CAP>
__global__ void k_wforces(const DevSys*, uint32_t, int, int, int);
template <typename R, int NS, bool E, int CAP>
__global__ void k_sforces(const DevSys*, uint32_t, int, int, int);
template <typename R, int NS>
__global__ void k_qforces(const DevSys*, uint32_t, int, double*, int);
template <typename R, int NS>
__global__ void k_qq(const DevSys*, uint32_t);
void launch_watpol(hipStream_t st, int nblk, int B, const DevSys* dsys, const int* istep);
__global__ void k_reduce_energies(const DevSys*, int, int);
__global__ void k_atom_terms(const DevSys*, uint32_t, const double*);

// per-instance MD record
struct MdDev {
  double dt, dt_over_tau, T0, Boltz;
  int nbcycle, ene_every, max_frames;
  int nbonds, nangles;
  const int* bond_ij;
  const double* bond_r0;
  const double* bond_k;
  const int* angle_ijk;
  const double* angle_t0;
  const double* angle_k;
  const double* winv;
  double roh2, rhh2;
  // temperature groups (temperature, src/md.f90:1379-1447): with separate
  // scaling, group 0 = solute atoms [0, nat_solute), group 1 = the solvent;
  // otherwise one group.  Degrees of freedom as init_constraints counts them
  // (src/simprep.f90:2439-2448).
  int ngroup;
  double ndegf, ndegfree;
  double gndegf[2], gndegfree[2];
  double* v;
  double* xx;
  double* hist;   // [max_frames][QNB_FRAME_LEN]
  double* st;     // [8]: sfact of group 0, Tfree, Temp, Ekin, tfree of groups 0 and 1, sfact of group 1
  int* ctr;       // [4]: frame, error flags
  double* tfree;  // &tfree_all[k]
  int* istep;     // &istep_all[k]
  double* xs;     // [xs_max][natom*3] coordinate snapshots (counter ctr[2])
  int xs_every, xs_max;
  double* bpart;  // [nbpart][2] bond and angle energy per k_bonded block
  int nbpart;
  // atom-centred bonded terms: for atom i, items boff[i]..boff[i+1]-1 of
  // bitem (2 * bond + role) and aoff[i]..aoff[i+1]-1 of aitem (3 * angle +
  // role); atoms [0, nbat) carry all of them
  const int* boff;
  const int* bitem;
  const int* aoff;
  const int* aitem;
  int nbat;
  // fused integrator: per-block kinetic energy sums
  double* tpart;  // [ceil((nwat + nat_solute) / 256)][4]: Ekin, free Ekin of groups 0 and 1
  // EVB off-diagonal elements sampled with the energies
  int noffd;
  int offd_k[QNB_MAX_OFFD], offd_l[QNB_MAX_OFFD];  // atom indices
  double offd_A[QNB_MAX_OFFD], offd_mu[QNB_MAX_OFFD];
  // per-state Q bonded energies of the last step ([nstates][4]), or null
  const double* qbe;
};

enum { ERR_NONFINITE = 1, ERR_SHAKE = 2, ERR_OVERFLOW = 4 };

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / QNB_WAVE, l = threadIdx.x % QNB_WAVE, nw = blockDim.x / QNB_WAVE;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < nw; k++) s += red[k];
  return s;
}

// harmonic bonds and angles of the solute (bond, angle: src/bondene.f90:41-246)
// One thread per atom sums the terms of every bond and angle it belongs to
// (each term is evaluated once per member atom), so the gradient needs no
// atomics and does not depend on scheduling.  A bond's energy is counted by
// its first atom, an angle's by its vertex; each block's energy sums go to
// bpart[block] and k_record adds them in block order.  bonded_atom is the
// per-atom part, used by k_bonded (the minimiser) and k_atom_step (MD).
__device__ __forceinline__ void bonded_atom(const DevSys& S, const MdDev& M, int i, double* f, double& eb,
                                            double& ea) {
  const auto x = gp(S.x);
  for (int it = ldg(M.boff + i), e = ldg(M.boff + i + 1); it < e; it++) {
    const int item = ldg(M.bitem + it), b = item >> 1, role = item & 1;
    const int bi = ldg(M.bond_ij + 2 * b), bj = ldg(M.bond_ij + 2 * b + 1);
    const double ax = x[3 * bj] - x[3 * bi], ay = x[3 * bj + 1] - x[3 * bi + 1], az = x[3 * bj + 2] - x[3 * bi + 2];
    const double r = sqrt(ax * ax + ay * ay + az * az);
    const double kb = ldg(M.bond_k + b), db = r - ldg(M.bond_r0 + b);
    const double dv = kb * db / r;
    const double sg = role == 0 ? -dv : dv;
    f[0] += ax * sg;
    f[1] += ay * sg;
    f[2] += az * sg;
    if (role == 0) eb += 0.5 * kb * db * db;
  }
  for (int it = ldg(M.aoff + i), e = ldg(M.aoff + i + 1); it < e; it++) {
    const int item = ldg(M.aitem + it), a = item / 3, role = item - 3 * a;
    const int ai = ldg(M.angle_ijk + 3 * a), aj = ldg(M.angle_ijk + 3 * a + 1), ak = ldg(M.angle_ijk + 3 * a + 2);
    double u[3], w[3];
    for (int c = 0; c < 3; c++) {
      u[c] = x[3 * ai + c] - x[3 * aj + c];
      w[c] = x[3 * ak + c] - x[3 * aj + c];
    }
    const double ru = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double rw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double ct = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (ru * rw);
    ct = fmin(1.0, fmax(-1.0, ct));
    const double th = acos(ct);
    // sin(theta) from the cross product, consistent with the numerator of
    // the derivative below even for nearly linear angles, where
    // sin(acos(ct)) loses all digits; at a linear angle the derivative has
    // no direction
    const double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
    const double st = sqrt(cx * cx + cy * cy + cz * cz) / (ru * rw);
    const double ka = ldg(M.angle_k + a), da = th - ldg(M.angle_t0 + a);
    if (role == 1) ea += 0.5 * ka * da * da;
    if (!(st > 1.0e-8)) continue;
    const double dv = ka * da;
    // dtheta/dx_i = -(w/(ru rw) - ct u/ru^2)/sin(theta), likewise for k; j balances
    for (int c = 0; c < 3; c++) {
      const double gi = -(w[c] / (ru * rw) - ct * u[c] / (ru * ru)) / st;
      const double gk = -(u[c] / (ru * rw) - ct * w[c] / (rw * rw)) / st;
      f[c] += dv * (role == 0 ? gi : (role == 2 ? gk : -(gi + gk)));
    }
  }
}

// solute bonds and angles alone (the minimiser's step)
__global__ __launch_bounds__(256) void k_bonded(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  __shared__ double red[4];
  double eb = 0.0, ea = 0.0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M.nbat) {
    double f[3] = {0.0, 0.0, 0.0};
    bonded_atom(S, M, i, f, eb, ea);
    const auto d = gpm(S.d);
    d[3 * i] += f[0];
    d[3 * i + 1] += f[1];
    d[3 * i + 2] += f[2];
  }
  const double sb = block_sum(eb, red);
  const double sa = block_sum(ea, red);
  if (threadIdx.x == 0 && blockIdx.x < M.nbpart) {
    M.bpart[2 * blockIdx.x] = sb;
    M.bpart[2 * blockIdx.x + 1] = sa;
  }
}

// The per-atom terms of an MD step in one pass: LRF Taylor application and
// the SCAAS radial restraint (as k_atom_terms with every term on) and the
// solute bonds and angles, added to the atom's gradient in one update.
__global__ __launch_bounds__(256) void k_atom_step(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds,
                                                   const double* __restrict__ tfree, double* __restrict__ dq,
                                                   int dq_stride) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  __shared__ double red[4];
  double eb = 0.0, ea = 0.0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S.natom) {
    double f[3] = {0.0, 0.0, 0.0};
    double el = 0.0, er = 0.0;
    if (S.use_lrf && !ldg(S.excl + i) && ldg(S.iqatom + i) == 0) lrf_taylor_atom(S, i, f, el);
    const int src = radial_source(S, i);
    if (src >= 0) {
      const double T = tfree[blockIdx.y];
      const double shift = S.fk_wsphere != 0.0 ? sqrt(S.Boltz * T / S.fk_wsphere) : 0.0;
      radial_restraint(S, src, shift, f, er);
    }
    if (i < M.nbat) bonded_atom(S, M, i, f, eb, ea);
    // the Q-atom side's rows of this step, cleared for the next one
    const int iq = i < S.nat_solute ? ldg(S.iqatom + i) : 0;
    if (iq > 0) {
      double* q = dq + (size_t)blockIdx.y * dq_stride + 3 * (iq - 1);
      for (int c = 0; c < 3; c++) {
        f[c] += q[c];
        q[c] = 0.0;
      }
    }
    store_atom_gradient(S, i, f);
    if (want_energies(S)) {
      stg(S.eatom + 2 * i, el);
      stg(S.eatom + 2 * i + 1, er);
    }
  }
  const double sb = block_sum(eb, red);
  const double sa = block_sum(ea, red);
  if (threadIdx.x == 0 && blockIdx.x < M.nbpart) {
    M.bpart[2 * blockIdx.x] = sb;
    M.bpart[2 * blockIdx.x + 1] = sa;
  }
}

// One MD step's integration of every atom, fused: leap-frog with the
// Berendsen factor of the previous step (newvel_ber, src/md.f90:1451-1466),
// SHAKE of each rigid water (src/bondene.f90:1069-1151), v = (x - x_old)/dt
// for every atom (src/md.f90:1787-1793) and per-block sums of the kinetic
// energy (k_thermo adds them).  Thread t < nwat takes water t (3 atoms), the
// next nat_solute threads one solute atom each.
__global__ __launch_bounds__(256) void k_integrate(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  const int nthr = S.nwat + S.nat_solute;
  const int nblk = (nthr + 255) / 256;
  if ((int)blockIdx.x >= nblk) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const double dt = M.dt;
  const auto xg = gpm(S.x);
  const auto vg = gpm(M.v);
  const auto dg = gp(S.d);
  // waters belong to the solvent group, solute atoms to group 0
  double ek = 0.0, ekf = 0.0;
  const bool solvent = t < S.nwat;
  const double s = (solvent && M.ngroup == 2) ? M.st[6] : M.st[0];
  if (t < S.nwat) {
    const int o = S.nat_solute + 3 * t;
    double xo[3][3], xn[3][3], wi[3];
    for (int k = 0; k < 3; k++) {
      wi[k] = ldg(M.winv + o + k);
      for (int c = 0; c < 3; c++) {
        const int q = 3 * (o + k) + c;
        const double v = (vg[q] - dg[q] * wi[k] * dt) * s;
        xo[k][c] = xg[q];
        xn[k][c] = xo[k][c] + v * dt;
      }
    }
    const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
    const double d2[3] = {M.roh2, M.roh2, M.rhh2};
    bool ready[3] = {false, false, false};
    int nits = 0;
    for (; nits < 1000; nits++) {
      for (int c = 0; c < 3; c++) {
        if (ready[c]) continue;
        const int i = ci[c], j = cj[c];
        const double vx = xn[i][0] - xn[j][0], vy = xn[i][1] - xn[j][1], vz = xn[i][2] - xn[j][2];
        const double diff = d2[c] - (vx * vx + vy * vy + vz * vz);
        if (fabs(diff) < 0.0001 * d2[c]) ready[c] = true;
        const double px = xo[i][0] - xo[j][0], py = xo[i][1] - xo[j][1], pz = xo[i][2] - xo[j][2];
        const double scp = vx * px + vy * py + vz * pz;
        const double corr = diff / (2.0 * scp * (wi[i] + wi[j]));
        xn[i][0] += px * corr * wi[i];
        xn[i][1] += py * corr * wi[i];
        xn[i][2] += pz * corr * wi[i];
        xn[j][0] -= px * corr * wi[j];
        xn[j][1] -= py * corr * wi[j];
        xn[j][2] -= pz * corr * wi[j];
      }
      if (ready[0] && ready[1] && ready[2]) break;
    }
    if (nits >= 1000) atomicOr(&M.ctr[1], ERR_SHAKE);
    for (int k = 0; k < 3; k++) {
      double v2 = 0.0;
      for (int c = 0; c < 3; c++) {
        const int q = 3 * (o + k) + c;
        const double v = (xn[k][c] - xo[k][c]) / dt;
        xg[q] = xn[k][c];
        vg[q] = v;
        v2 += v * v;
      }
      const double e = 0.5 * (1.0 / wi[k]) * v2;
      ek += e;
      if (!ldg(S.excl + o + k)) ekf += e;
    }
  } else if (t < nthr) {
    const int i = t - S.nwat;
    const double wi = ldg(M.winv + i);
    double v2 = 0.0;
    for (int c = 0; c < 3; c++) {
      const int q = 3 * i + c;
      const double v0 = (vg[q] - dg[q] * wi * dt) * s;
      const double xo = xg[q], xn = xo + v0 * dt;
      const double v = (xn - xo) / dt;
      xg[q] = xn;
      vg[q] = v;
      v2 += v * v;
    }
    const double e = 0.5 * (1.0 / wi) * v2;
    ek += e;
    if (!ldg(S.excl + i)) ekf += e;
  }
  __shared__ double red[4];
  const bool g1 = solvent && M.ngroup == 2;
  const double ek0 = block_sum(g1 ? 0.0 : ek, red);
  const double ekf0 = block_sum(g1 ? 0.0 : ekf, red);
  const double ek1 = block_sum(g1 ? ek : 0.0, red);
  const double ekf1 = block_sum(g1 ? ekf : 0.0, red);
  if (threadIdx.x == 0) {
    M.tpart[4 * blockIdx.x] = ek0;
    M.tpart[4 * blockIdx.x + 1] = ekf0;
    M.tpart[4 * blockIdx.x + 2] = ek1;
    M.tpart[4 * blockIdx.x + 3] = ekf1;
  }
}

// Temperatures and Berendsen factors from the kinetic energies of the
// temperature groups (temperature, src/md.f90:1412-1443): Temp and Tfree over
// all atoms, each group's free temperature, and sfact = sqrt(1 + dt/tau
// (T0/tfree - 1)); a group with tfree = 0 keeps its previous sfact inside the
// square root, as the reference's update does.
__device__ void thermo_update(const MdDev& M, const double* ek, const double* ekf) {
  const double eks = M.ngroup == 2 ? ek[0] + ek[1] : ek[0];
  const double ekfs = M.ngroup == 2 ? ekf[0] + ekf[1] : ekf[0];
  const double T = 2.0 * eks / M.Boltz / M.ndegf;
  const double Tf = 2.0 * ekfs / M.Boltz / M.ndegfree;
  M.st[1] = Tf;
  M.st[2] = T;
  M.st[3] = eks;
  *M.tfree = Tf;
  for (int g = 0; g < M.ngroup; g++) {
    const double tf = 2.0 * ekf[g] / M.Boltz / M.gndegfree[g];
    double sf = M.st[g == 0 ? 0 : 6];
    if (tf != 0.0) sf = M.T0 / tf - 1.0;
    M.st[g == 0 ? 0 : 6] = sqrt(1.0 + M.dt_over_tau * sf);
    M.st[4 + g] = tf;
  }
}

// After k_integrate, one thread per system: the kinetic energy from the
// block sums in block order, the temperature, Tfree and the Berendsen factor
// of the next step (temperature, src/md.f90:1379-1447); advances the step
// counter and records errors.
__global__ __launch_bounds__(64) void k_thermo(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.x];
  const MdDev& M = mds[blockIdx.x];
  if (threadIdx.x != 0) return;
  const int nblk = (S.nwat + S.nat_solute + 255) / 256;
  double ek[2] = {0.0, 0.0}, ekf[2] = {0.0, 0.0};
  for (int k = 0; k < nblk; k++) {
    ek[0] += M.tpart[4 * k];
    ekf[0] += M.tpart[4 * k + 1];
    ek[1] += M.tpart[4 * k + 2];
    ekf[1] += M.tpart[4 * k + 3];
  }
  thermo_update(M, ek, ekf);
  const double T = M.st[2];
  if (!isfinite(T) || !isfinite(S.ered[E_PP_EL] + S.ered[E_WW_EL])) atomicOr(&M.ctr[1], ERR_NONFINITE);
  if (*S.need > S.cap) atomicOr(&M.ctr[1], ERR_OVERFLOW);
  *M.istep = *M.istep + 1;
}

// shake (src/bondene.f90:1069-1151) on each rigid water (O-H1, O-H2, H1-H2),
// then v = (x - xx)/dt for every atom (src/md.f90:1787-1793)
__global__ __launch_bounds__(256) void k_shake(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double* x = S.x;
  const double* xx = M.xx;
  if (t < S.nwat) {
    const int o = S.nat_solute + 3 * t;
    const int ci[3] = {o, o, o + 1}, cj[3] = {o + 1, o + 2, o + 2};
    const double d2[3] = {M.roh2, M.roh2, M.rhh2};
    bool ready[3] = {false, false, false};
    int nits = 0;
    for (; nits < 1000; nits++) {
      for (int c = 0; c < 3; c++) {
        if (ready[c]) continue;
        const int i = ci[c], j = cj[c];
        double vx = x[3 * i] - x[3 * j], vy = x[3 * i + 1] - x[3 * j + 1], vz = x[3 * i + 2] - x[3 * j + 2];
        const double diff = d2[c] - (vx * vx + vy * vy + vz * vz);
        if (fabs(diff) < 0.0001 * d2[c]) ready[c] = true;
        const double px = xx[3 * i] - xx[3 * j], py = xx[3 * i + 1] - xx[3 * j + 1], pz = xx[3 * i + 2] - xx[3 * j + 2];
        const double scp = vx * px + vy * py + vz * pz;
        const double wi = M.winv[i], wj = M.winv[j];
        const double corr = diff / (2.0 * scp * (wi + wj));
        x[3 * i] += px * corr * wi;
        x[3 * i + 1] += py * corr * wi;
        x[3 * i + 2] += pz * corr * wi;
        x[3 * j] -= px * corr * wj;
        x[3 * j + 1] -= py * corr * wj;
        x[3 * j + 2] -= pz * corr * wj;
      }
      if (ready[0] && ready[1] && ready[2]) break;
    }
    if (nits >= 1000) atomicOr(&M.ctr[1], ERR_SHAKE);
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) M.v[3 * (o + k) + c] = (x[3 * (o + k) + c] - xx[3 * (o + k) + c]) / M.dt;
  } else if (t < S.nwat + S.nat_solute) {
    const int i = t - S.nwat;
    for (int c = 0; c < 3; c++) M.v[3 * i + c] = (x[3 * i + c] - xx[3 * i + c]) / M.dt;
  }
}

// steepest-descent move used to relax a start structure before MD: every atom
// moves along -d by k0 * |d|, capped at max_disp (A); the previous positions
// go to xx so k_shake restores the rigid waters afterwards
__global__ __launch_bounds__(256) void k_sd_move(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds,
                                                 double k0, double max_disp) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.natom) return;
  const double gx = S.d[3 * i], gy = S.d[3 * i + 1], gz = S.d[3 * i + 2];
  const double g = sqrt(gx * gx + gy * gy + gz * gz);
  M.xx[3 * i] = S.x[3 * i];
  M.xx[3 * i + 1] = S.x[3 * i + 1];
  M.xx[3 * i + 2] = S.x[3 * i + 2];
  if (!isfinite(g)) {  // leave the atom in place and report
    atomicOr(&M.ctr[1], ERR_NONFINITE);
    return;
  }
  double sc = k0;
  if (g * k0 > max_disp) sc = max_disp / g;
  S.x[3 * i] -= gx * sc;
  S.x[3 * i + 1] -= gy * sc;
  S.x[3 * i + 2] -= gz * sc;
}

// temperature (src/md.f90:1379-1447) of the starting velocities, one
// temperature group: Temp, Tfree and the Berendsen factor of the first step
// (during MD, k_integrate computes them)
__global__ __launch_bounds__(256) void k_temperature(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.x];
  const MdDev& M = mds[blockIdx.x];
  __shared__ double red[4];
  double e0 = 0.0, ef0 = 0.0, e1 = 0.0, ef1 = 0.0;
  for (int i = threadIdx.x; i < S.natom; i += blockDim.x) {
    const double m = 1.0 / M.winv[i];
    const double v2 = M.v[3 * i] * M.v[3 * i] + M.v[3 * i + 1] * M.v[3 * i + 1] + M.v[3 * i + 2] * M.v[3 * i + 2];
    const double e = 0.5 * m * v2;
    const bool g1 = M.ngroup == 2 && i >= S.nat_solute;
    (g1 ? e1 : e0) += e;
    if (!S.excl[i]) (g1 ? ef1 : ef0) += e;
  }
  double ek[2], ekf[2];
  ek[0] = block_sum(e0, red);
  ekf[0] = block_sum(ef0, red);
  ek[1] = block_sum(e1, red);
  ekf[1] = block_sum(ef1, red);
  if (threadIdx.x == 0) {
    thermo_update(M, ek, ekf);
    const double T = M.st[2];
    if (!isfinite(T) || !isfinite(S.ered[E_PP_EL] + S.ered[E_WW_EL])) atomicOr(&M.ctr[1], ERR_NONFINITE);
    if (*S.need > S.cap) atomicOr(&M.ctr[1], ERR_OVERFLOW);
  }
}

// energy sampling every ene_every steps (put_ene cadence, src/md.f90:1829-1834)
// coordinate snapshot of the sampled step into slot ctr[2] (k_record, next in
// the stream, advances the counter)
__global__ void k_snap(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  if (M.xs_every <= 0 || *M.istep % M.xs_every != 0) return;
  const int f = M.ctr[2];
  if (f >= M.xs_max) return;
  const int n = 3 * S.natom;
  double* dst = M.xs + (size_t)f * n;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) dst[k] = S.x[k];
}

__global__ void k_record(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds) {
  const DevSys& S = systems[blockIdx.x];
  const MdDev& M = mds[blockIdx.x];
  const int step = *M.istep;
  if (threadIdx.x == 0 && M.xs_every > 0 && step % M.xs_every == 0 && M.ctr[2] < M.xs_max) M.ctr[2] += 1;
  if (M.ene_every <= 0 || step % M.ene_every != 0) return;
  const int f = M.ctr[0];
  if (f >= M.max_frames) return;
  double* h = M.hist + (size_t)f * QNB_FRAME_LEN;
  const int k = threadIdx.x;
  if (k < QNB_NE) h[k] = S.ered[k];
  __syncthreads();
  if (k == 0) {
    double eb = 0.0, ea = 0.0;
    for (int p = 0; p < M.nbpart; p++) {
      eb += M.bpart[2 * p];
      ea += M.bpart[2 * p + 1];
    }
    h[QNB_F_BOND] = eb;
    h[QNB_F_ANGLE] = ea;
    h[QNB_F_EKIN] = M.st[3];
    h[QNB_F_TEMP] = M.st[2];
    h[QNB_F_TFREE] = M.st[1];
    h[QNB_F_STEP] = (double)step;
    for (int z = QNB_F_STEP + 1; z < QNB_FRAME_LEN; z++) h[z] = 0.0;
    // offdiag (nonbondene.f90:6238-6260): H = A exp(-mu r_kl)
    for (int m = 0; m < M.noffd; m++) {
      const double r = sqrt(dist4(S.x, M.offd_k[m], M.offd_l[m]));
      h[QNB_F_OFFD + 2 * m] = M.offd_A[m] * exp(-M.offd_mu[m] * r);
      h[QNB_F_OFFD + 2 * m + 1] = r;
    }
    if (M.qbe)
      for (int z = 0; z < 4 * S.nstates; z++) h[QNB_F_QB0 + z] = M.qbe[z];
    M.ctr[0] = f + 1;
  }
}

}  // namespace qnb

using namespace qnb;

struct qnb_batch {
  int n = 0;
  std::vector<DevSys> hs;
  std::vector<MdDev> hm;
  DevSys* dsys = nullptr;
  MdDev* dmd = nullptr;
  double* tfree = nullptr;
  int* istep = nullptr;
  DevAlloc mem;
  std::vector<std::unique_ptr<HostTopo>> topos;
  std::vector<int> topo_of;  // instance -> index into topos
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // second branch of the MD step's force pass (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t cycle_exec = nullptr;
  hipGraph_t cycle_graph = nullptr;
  int maxg = 0, maxat = 0, maxsw = 0, maxw = 0, maxs = 0, ns_max = 1, prec = 0;
  int nbblk = 1;  // k_bonded blocks per instance
  bool anyq = false, anypol = false, static_q = true;
  int qchunks = 1, nq_max = 0;
  double* dq = nullptr;  // [n][nq_max][3]: the Q-atom side's gradient rows of the MD step
  int wpb_lists = 4, wpb_forces = 4;  // waves per block, picked by qnb_batch_autotune
  int imap_water = 1;                 // block placement of the water-receiver kernel (map_block), autotuned
  int imap_solute = 0;                // block placement of the solute-receiver kernel, autotuned
  bool small_groups = true;           // every solute group fits the water kernel's small staging
  std::vector<int> nbcycle;
  std::vector<int> ene_every;  // per instance energy sampling interval
  // Q bonded terms per instance
  std::vector<QbDev> hqb;
  std::vector<std::vector<void*>> qb_owned;
  QbDev* dqb = nullptr;
  bool anyqb = false;
  bool anysnap = false;  // some instance stores coordinate snapshots (k_snap launched)
  // group-contribution arrays per instance (k_exc_record after the energy record)
  std::vector<ExcDev> hexc;
  std::vector<std::unique_ptr<DevAlloc>> exc_mem;
  ExcDev* dexc = nullptr;
  int exc_threads = 0;  // largest narr * nstates over the instances (0: not launched)
  long long steps_done = 0;
  ~qnb_batch() {
    if (cycle_exec) (void)hipGraphExecDestroy(cycle_exec);
    if (cycle_graph) (void)hipGraphDestroy(cycle_graph);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

__global__ void k_exc_record(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds,
                             const ExcDev* __restrict__ excs);

// launchers shared by the step, the minimiser, the timer and the autotuner
static void launch_lists(qnb_batch* b, int wpb) {
  dim3 g((b->maxg + wpb - 1) / wpb, b->n), t(64 * wpb);
  bool any = false, quirk = false;
  int maxsg = 0;
  for (const DevSys& S : b->hs) {
    any = any || S.any_atom;
    quirk = quirk || pw_quirk(S);
    maxsg = std::max(maxsg, S.ncgp_solute);
  }
  if (quirk) hipLaunchKernelGGL(k_pw_first, dim3((maxsg + 3) / 4, b->n), dim3(256), 0, b->stream, b->dsys);
  hipLaunchKernelGGL(k_pack_xq, dim3((b->maxat + 255) / 256, b->n), dim3(256), 0, b->stream, b->dsys);
  if (b->prec == QNB_PREC_MIXED) {
    if (any) hipLaunchKernelGGL((k_lists_lrf<float, true>), g, t, 0, b->stream, b->dsys);
    else hipLaunchKernelGGL((k_lists_lrf<float, false>), g, t, 0, b->stream, b->dsys);
  } else {
    if (any) hipLaunchKernelGGL((k_lists_lrf<double, true>), g, t, 0, b->stream, b->dsys);
    else hipLaunchKernelGGL((k_lists_lrf<double, false>), g, t, 0, b->stream, b->dsys);
  }
}

static const uint32_t kStepMask = QNB_T_NONBONDED | QNB_T_QQ | QNB_T_QQP | QNB_T_RAD | QNB_T_POL;

// waves per block of the force kernel, kept within 64 KiB of LDS per block
// (the per-wave partner slots plus the solute launch's type-pair tables)
template <typename R>
static int forces_wpb(int wpb) {
  const int cap = (int)((65536 - fc_static_bytes<R>()) / fc_wave_bytes<R>());
  return std::max(1, std::min(std::min(wpb, FC_WPB), cap));
}

// Blocks are dealt to the 8 XCDs round-robin in linear block order
// (blockIdx.y * gridDim.x + blockIdx.x); with gridDim.x a multiple of 8 the
// XCD of a block depends on blockIdx.x alone, so the blocks of every instance
// that read the same rows of the shared tables (Q tables, topology) run on
// the same XCD and find them in its L2.  The padding blocks exit at once.
static unsigned xcd_cols(int blocks) { return (unsigned)((blocks + 7) / 8 * 8); }

// water receivers (k_wforces) with the block placement imap (map_block)
template <typename R, int NS, bool E, int CAP>
static void launch_water(qnb_batch* b, int wpb, int imap, hipStream_t st) {
  const int nb = (b->maxw + wpb - 1) / wpb;
  const dim3 grid = imap == 0 ? dim3(xcd_cols(nb), b->n) : dim3((unsigned)(8 * (((long)nb * b->n + 7) / 8)), 1);
  hipLaunchKernelGGL((k_wforces<R, NS, E, CAP>), grid, dim3(64 * wpb), (wf_wave_bytes<R, CAP>() * wpb), st, b->dsys,
                     kStepMask, imap, nb, b->n);
}

template <typename R, int NS, bool E>
static void launch_forces_e(qnb_batch* b, int wpb, int parts, hipStream_t wst) {
  if (b->maxw > 0 && (parts & 1)) {
    if (b->small_groups) launch_water<R, NS, E, WF_CAP_SMALL>(b, wpb, b->imap_water, wst);
    else launch_water<R, NS, E, FC_CAP>(b, wpb, b->imap_water, wst);
  }
  if (b->maxs > 0 && (parts & 2)) {
    const int nb = (b->maxs + wpb - 1) / wpb, imap = b->imap_solute;
    const dim3 grid = imap == 0 ? dim3(xcd_cols(nb), b->n) : dim3((unsigned)(8 * (((long)nb * b->n + 7) / 8)), 1);
    if (b->small_groups)
      hipLaunchKernelGGL((k_sforces<R, NS, E, WF_CAP_SMALL>), grid, dim3(64 * wpb),
                         (sf_wave_bytes<R, WF_CAP_SMALL>() * wpb), b->stream, b->dsys, kStepMask, imap, nb, b->n);
    else
      hipLaunchKernelGGL((k_sforces<R, NS, E, FC_CAP>), grid, dim3(64 * wpb), (sf_wave_bytes<R, FC_CAP>() * wpb),
                         b->stream, b->dsys, kStepMask, imap, nb, b->n);
  }
}

// energy: the step samples energies in at least one instance (or the host
// cannot tell, e.g. inside a graph whose period is not a multiple of every
// instance's sampling interval); parts: 1 water receivers, 2 solute receivers
template <typename R, int NS>
static void launch_forces_t(qnb_batch* b, int wpb, bool energy = true, bool atom_terms = true, int parts = 3,
                            hipStream_t wst = nullptr) {
  wpb = forces_wpb<R>(wpb);
  if (!wst) wst = b->stream;
  if (energy) launch_forces_e<R, NS, true>(b, wpb, parts, wst);
  else launch_forces_e<R, NS, false>(b, wpb, parts, wst);
  if (atom_terms)
    hipLaunchKernelGGL(k_atom_terms, dim3((b->maxat + 255) / 256, b->n), dim3(256), 0, b->stream, b->dsys, kStepMask,
                       (const double*)b->tfree);
}

static void launch_forces(qnb_batch* b, int wpb, bool energy = true, bool atom_terms = true, int parts = 3) {
  const bool mixed = b->prec == QNB_PREC_MIXED;
#define QNB_LF(NS)                                                 \
  (mixed ? launch_forces_t<float, NS>(b, wpb, energy, atom_terms, parts) \
         : launch_forces_t<double, NS>(b, wpb, energy, atom_terms, parts))
  switch (b->ns_max) {
    case 1: QNB_LF(1); break;
    case 2: QNB_LF(2); break;
    case 3:
    case 4: QNB_LF(4); break;
    default: QNB_LF(8); break;
  }
#undef QNB_LF
}

// the Q-atom side of the Q-solute and Q-water pairs alone (kernel timing)
static void launch_qforces(qnb_batch* b) {
  const bool mixed = b->prec == QNB_PREC_MIXED;
  const int wpb = 4;
  const dim3 grid(qforces_blocks(b->nq_max, b->qchunks, wpb), b->n), block(64 * wpb);
#define QNB_LQ(NS)                                                                                     \
  if (mixed)                                                                                                  \
    hipLaunchKernelGGL((k_qforces<float, NS>), grid, block, 0, b->stream, b->dsys, kStepMask, b->qchunks,     \
                       (double*)nullptr, 0);                                                                  \
  else hipLaunchKernelGGL((k_qforces<double, NS>), grid, block, 0, b->stream, b->dsys, kStepMask, b->qchunks, \
                          (double*)nullptr, 0)
  switch (b->ns_max) {
    case 1: QNB_LQ(1); break;
    case 2: QNB_LQ(2); break;
    case 3:
    case 4: QNB_LQ(4); break;
    default: QNB_LQ(8); break;
  }
#undef QNB_LQ
}

// atom_terms = false: the caller runs the per-atom terms itself (the MD step
// fuses them with the bonded terms in k_atom_step)
// The MD step's force pass on two streams: the Q-atom side (first: its few
// blocks start beside the solute receivers), the water receivers and the
// polarisation restraint on the side stream, the solute receivers and the
// Q-Q pairs on the batch stream.  The two branches write
// disjoint gradient rows (waters and the Q-atom side's own rows dq / solute
// atoms) and disjoint energy slots; the batch stream waits for the side
// branch before the per-atom terms, which add dq to the Q-atoms' gradients.
template <typename R, int NS>
static void enqueue_forces_forked(qnb_batch* b, bool energy) {
  const int wpb = 4, B = b->n;
  (void)hipEventRecord(b->ev_fork, b->stream);
  (void)hipStreamWaitEvent(b->side, b->ev_fork, 0);
  if (b->anyq)
    hipLaunchKernelGGL((k_qforces<R, NS>), dim3(qforces_blocks(b->nq_max, b->qchunks, wpb), B), dim3(64 * wpb), 0,
                       b->side, b->dsys, kStepMask, b->qchunks, b->dq, 3 * b->nq_max);
  if (b->maxw > 0) launch_forces_t<R, NS>(b, b->wpb_forces, energy, false, 1, b->side);
  if (b->anypol)
    launch_watpol(b->side, (b->maxw + WP_THREADS - 1) / WP_THREADS, B, b->dsys, (const int*)b->istep);
  launch_forces_t<R, NS>(b, b->wpb_forces, energy, false, 2);
  if (b->anyq) hipLaunchKernelGGL((k_qq<R, NS>), dim3(B), dim3(QQ_THREADS), 0, b->stream, b->dsys, kStepMask);
  (void)hipEventRecord(b->ev_join, b->side);
  (void)hipStreamWaitEvent(b->stream, b->ev_join, 0);
}

template <typename R, int NS>
static void enqueue_forces(qnb_batch* b, bool energy = true, bool atom_terms = true) {
  const int wpb = 4, B = b->n;
  const uint32_t mask = kStepMask;
  launch_forces_t<R, NS>(b, b->wpb_forces, energy, atom_terms);
  if (b->anyq) {
    hipLaunchKernelGGL((k_qforces<R, NS>), dim3(qforces_blocks(b->nq_max, b->qchunks, wpb), B), dim3(64 * wpb), 0,
                       b->stream, b->dsys, mask, b->qchunks, (double*)nullptr, 0);
    hipLaunchKernelGGL((k_qq<R, NS>), dim3(B), dim3(QQ_THREADS), 0, b->stream, b->dsys, mask);
  }
  if (b->anypol) launch_watpol(b->stream, (b->maxw + WP_THREADS - 1) / WP_THREADS, B, b->dsys, (const int*)b->istep);
}

// leap-frog + SHAKE + new velocities, then the temperatures and Berendsen
// factors of the next step
static void launch_integrate(qnb_batch* b) {
  hipLaunchKernelGGL(k_integrate, dim3((b->maxsw + 255) / 256, b->n), dim3(256), 0, b->stream, b->dsys, b->dmd);
  hipLaunchKernelGGL(k_thermo, dim3(b->n), dim3(64), 0, b->stream, b->dsys, b->dmd);
}

template <typename R>
static void enqueue_step(qnb_batch* b, bool lists, bool energy) {
  const int B = b->n;
  if (lists) {
    launch_lists(b, b->wpb_lists);
    if (b->anyq && !b->static_q) hipLaunchKernelGGL(k_qlists, dim3(B), dim3(1024), 0, b->stream, b->dsys);
  }
  switch (b->ns_max) {
    case 1: enqueue_forces_forked<R, 1>(b, energy); break;
    case 2: enqueue_forces_forked<R, 2>(b, energy); break;
    case 3:
    case 4: enqueue_forces_forked<R, 4>(b, energy); break;
    default: enqueue_forces_forked<R, 8>(b, energy); break;
  }
  hipLaunchKernelGGL(k_atom_step, dim3((b->maxat + 255) / 256, B), dim3(256), 0, b->stream, b->dsys, b->dmd,
                     (const double*)b->tfree, b->dq, 3 * b->nq_max);
  if (b->anyqb) launch_qbonded(b->stream, B, b->dsys, b->dqb);
  // the reduction and the record do nothing on a step no instance samples
  // (energy is false only then); the record also counts stored snapshots
  if (energy)
    hipLaunchKernelGGL(k_reduce_energies, dim3(B), dim3(1024), 0, b->stream, b->dsys, (int)b->anyq, (int)b->anypol);
  if (b->anysnap)
    hipLaunchKernelGGL(k_snap, dim3((3 * b->maxat + 255) / 256, B), dim3(256), 0, b->stream, b->dsys, b->dmd);
  if (energy || b->anysnap) hipLaunchKernelGGL(k_record, dim3(B), dim3(64), 0, b->stream, b->dsys, b->dmd);
  if (energy && b->exc_threads > 0)
    hipLaunchKernelGGL(k_exc_record, dim3((b->exc_threads + 63) / 64, B), dim3(64), 0, b->stream, b->dsys, b->dmd,
                       (const ExcDev*)b->dexc);
  launch_integrate(b);
}

template <typename R>
static void enqueue_min_step(qnb_batch* b, bool lists, double k0, double max_disp) {
  const int B = b->n;
  if (lists) {
    launch_lists(b, b->wpb_lists);
    if (b->anyq && !b->static_q) hipLaunchKernelGGL(k_qlists, dim3(B), dim3(1024), 0, b->stream, b->dsys);
  }
  switch (b->ns_max) {
    case 1: enqueue_forces<R, 1>(b); break;
    case 2: enqueue_forces<R, 2>(b); break;
    case 3:
    case 4: enqueue_forces<R, 4>(b); break;
    default: enqueue_forces<R, 8>(b); break;
  }
  hipLaunchKernelGGL(k_bonded, dim3(b->nbblk, B), dim3(256), 0, b->stream, b->dsys, b->dmd);
  if (b->anyqb) launch_qbonded(b->stream, B, b->dsys, b->dqb);
  hipLaunchKernelGGL(k_sd_move, dim3((b->maxat + 255) / 256, B), dim3(256), 0, b->stream, b->dsys, b->dmd, k0, max_disp);
  hipLaunchKernelGGL(k_shake, dim3((b->maxsw + 255) / 256, B), dim3(256), 0, b->stream, b->dsys, b->dmd);
}

// does MD step `step` sample energies in any instance
static bool energy_step(const qnb_batch* b, long long step) {
  for (int e : b->ene_every)
    if (e <= 0 || step % e == 0) return true;
  return false;
}

static void enqueue(qnb_batch* b, bool lists, bool energy) {
  if (b->prec == QNB_PREC_MIXED) enqueue_step<float>(b, lists, energy);
  else enqueue_step<double>(b, lists, energy);
}

extern "C" int32_t qnb_batch_destroy(qnb_batch* b) {
  delete b;
  return 0;
}

extern "C" int32_t qnb_batch_create(int32_t n, const qnb_system* const* sys, const qnb_params* const* par,
                                    const qnb_md_params* const* md, const double* const* x0, const double* const* v0,
                                    qnb_batch** out) {
  if (!out || n <= 0 || !sys || !par || !md || !x0 || !v0) return fail(QNB_EINVAL, "bad batch arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(QNB_EDEVICE, "no HIP device");
  std::unique_ptr<qnb_batch> b(new qnb_batch());
  b->n = n;
  b->hs.resize(n);
  b->hm.resize(n);
  b->prec = par[0]->precision;
  std::map<const qnb_system*, int> shared;  // system pointer -> instance holding its topology
  int rc = 0;
  for (int k = 0; k < n; k++) {
    if ((rc = validate_system(sys[k], par[k]))) return rc;
    if (sys[k]->nwat > 0 && sys[k]->solv_atom != 3)
      return fail(QNB_EUNSUPPORTED, "batched MD integrates rigid 3-site water only (instance %d: solv_atom=%d)", k,
                  sys[k]->solv_atom);
    if (par[k]->precision != b->prec) return fail(QNB_EINVAL, "all instances need the same precision");
    DevSys& S = b->hs[k];
    memset(&S, 0, sizeof S);
    auto it = shared.find(sys[k]);
    if (it == shared.end()) {
      fill_params(S, sys[k], par[k]);
      b->topos.emplace_back(new HostTopo());
      if ((rc = topo_build(S, sys[k], *b->topos.back(), b->mem))) return rc;
      shared[sys[k]] = k;
      b->topo_of.push_back((int)b->topos.size() - 1);
    } else {
      S = b->hs[it->second];  // topology pointers shared; instance fields below
      fill_params(S, sys[k], par[k]);
      b->topo_of.push_back(b->topo_of[it->second]);
    }
    S.cap = 0;
    if ((rc = instance_alloc(S, b->mem, par[k], sys[k]->ncgp + 2))) return rc;
    HIPCHK(hipMemcpy(S.x, x0[k], sizeof(double) * 3 * S.natom, hipMemcpyHostToDevice));
    b->maxg = std::max(b->maxg, S.ncgp);
    b->maxw = std::max(b->maxw, S.nwat);
    b->maxs = std::max(b->maxs, S.ncgp_solute);
    b->small_groups = b->small_groups &&
                      max_solute_group(sys[k]->cgp_first, sys[k]->cgp_last, sys[k]->ncgp_solute) <= WF_CAP_SMALL;
    b->maxat = std::max(b->maxat, S.natom);
    b->maxsw = std::max(b->maxsw, S.nwat + S.nat_solute);
    b->ns_max = std::max(b->ns_max, S.nstates);
    b->anyq = b->anyq || S.nqat > 0;
    b->anypol = b->anypol || (S.nwat > 0 && S.wpol && S.nshell > 0);
    b->static_q = b->static_q && (S.rcq2 > S.rexcl2);
    // about QF_PART / 64 partners per lane of k_qforces
    b->qchunks = std::max(b->qchunks, std::min((S.nwat + S.nat_solute + QF_PART - 1) / QF_PART, 64));
    b->nbcycle.push_back(md[k]->nbcycle);
    if (md[k]->nbcycle != md[0]->nbcycle) return fail(QNB_EINVAL, "all instances need the same nbcycle");
  }
  for (int k = 0; k < n; k++) b->nq_max = std::max(b->nq_max, b->hs[k].nqat);
  HIPCHK(hipStreamCreate(&b->stream));
  HIPCHK(hipStreamCreate(&b->side));
  HIPCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  if ((rc = b->mem.alloc(&b->tfree, n)) || (rc = b->mem.alloc(&b->istep, n))) return rc;
  HIPCHK(hipMemset(b->istep, 0, sizeof(int) * n));
  if ((rc = b->mem.alloc(&b->dq, (size_t)n * 3 * std::max(b->nq_max, 1)))) return rc;
  HIPCHK(hipMemset(b->dq, 0, sizeof(double) * n * 3 * std::max(b->nq_max, 1)));
  // MD records
  for (int k = 0; k < n; k++) {
    const qnb_md_params* m = md[k];
    MdDev& M = b->hm[k];
    memset(&M, 0, sizeof M);
    const DevSys& S = b->hs[k];
    M.dt = 0.020462 * m->dt_fs;
    M.dt_over_tau = M.dt / (0.020462 * m->tau_T_fs);  // dt / tau_T in internal units (src/md.f90:207, 231)
    M.T0 = m->T0;
    M.Boltz = par[k]->Boltz;
    M.nbcycle = m->nbcycle;
    M.ene_every = m->ene_every;
    M.max_frames = m->max_frames;
    M.nbonds = m->nbonds;
    M.nangles = m->nangles;
    M.roh2 = m->water_roh * m->water_roh;
    M.rhh2 = m->water_rhh * m->water_rhh;
    // degrees of freedom (init_constraints, src/simprep.f90:2194-2250,
    // 2439-2448): three constraints per rigid water; a constraint counts half
    // for each excluded atom it joins (exclshk), which for a water adds its
    // number of excluded atoms
    M.ngroup = (m->separate_scaling && S.nat_solute > 0 && S.nat_solute < S.natom) ? 2 : 1;
    {
      int nex[2] = {0, 0};
      double exsh[2] = {0.0, 0.0};
      for (int i = 0; i < S.natom; i++) {
        const int g = (M.ngroup == 2 && i >= S.nat_solute) ? 1 : 0;
        if (sys[k]->excl[i]) {
          nex[g]++;
          if (i >= S.nat_solute) exsh[g] += 1.0;
        }
      }
      const int nat_g[2] = {M.ngroup == 2 ? S.nat_solute : S.natom, S.natom - S.nat_solute};
      const int shk_g[2] = {M.ngroup == 2 ? 0 : 3 * S.nwat, 3 * S.nwat};
      for (int g = 0; g < M.ngroup; g++) {
        M.gndegf[g] = 3.0 * nat_g[g] - shk_g[g];
        M.gndegfree[g] = (double)(long long)(M.gndegf[g] - 3.0 * nex[g] + exsh[g]);
      }
      M.ndegf = 3.0 * S.natom - 3.0 * S.nwat;
      M.ndegfree = (double)(long long)(M.ndegf - 3.0 * (nex[0] + nex[1]) + (exsh[0] + exsh[1]));
    }
    if ((rc = b->mem.upload(&M.bond_ij, m->bond_ij, 2 * (size_t)m->nbonds)) ||
        (rc = b->mem.upload(&M.bond_r0, m->bond_r0, m->nbonds)) || (rc = b->mem.upload(&M.bond_k, m->bond_k, m->nbonds)) ||
        (rc = b->mem.upload(&M.angle_ijk, m->angle_ijk, 3 * (size_t)m->nangles)) ||
        (rc = b->mem.upload(&M.angle_t0, m->angle_t0, m->nangles)) ||
        (rc = b->mem.upload(&M.angle_k, m->angle_k, m->nangles)) || (rc = b->mem.upload(&M.winv, m->winv, S.natom)) ||
        (rc = b->mem.alloc(&M.v, 3 * (size_t)S.natom)) || (rc = b->mem.alloc(&M.xx, 3 * (size_t)S.natom)) ||
        (rc = b->mem.alloc(&M.hist, (size_t)std::max(m->max_frames, 1) * QNB_FRAME_LEN)) ||
        (rc = b->mem.alloc(&M.st, 8)) || (rc = b->mem.alloc(&M.ctr, 4)))
      return rc;
    // atom-centred incidence lists of the bonds and angles (k_bonded)
    if (m->nbonds < 0 || m->nangles < 0 || (m->nbonds > 0 && !m->bond_ij) || (m->nangles > 0 && !m->angle_ijk))
      return fail(QNB_EINVAL, "instance %d: bad bond/angle arrays", k);
    {
      std::vector<int> boff(S.natom + 1, 0), aoff(S.natom + 1, 0);
      for (int q = 0; q < 2 * m->nbonds; q++) {
        const int at = m->bond_ij[q];
        if (at < 0 || at >= S.natom) return fail(QNB_EINVAL, "instance %d: bond atom %d out of range", k, at);
        boff[at + 1]++;
      }
      for (int q = 0; q < 3 * m->nangles; q++) {
        const int at = m->angle_ijk[q];
        if (at < 0 || at >= S.natom) return fail(QNB_EINVAL, "instance %d: angle atom %d out of range", k, at);
        aoff[at + 1]++;
      }
      int nbat = 0;
      for (int i = 0; i < S.natom; i++) {
        if (boff[i + 1] > 0 || aoff[i + 1] > 0) nbat = i + 1;
        boff[i + 1] += boff[i];
        aoff[i + 1] += aoff[i];
      }
      std::vector<int> bitem(std::max(boff[S.natom], 1)), aitem(std::max(aoff[S.natom], 1));
      std::vector<int> bpos(boff.begin(), boff.end() - 1), apos(aoff.begin(), aoff.end() - 1);
      for (int q = 0; q < 2 * m->nbonds; q++) bitem[bpos[m->bond_ij[q]]++] = q;  // q = 2 * bond + role
      for (int q = 0; q < 3 * m->nangles; q++) aitem[apos[m->angle_ijk[q]]++] = q;  // q = 3 * angle + role
      M.nbat = nbat;
      if ((rc = b->mem.upload(&M.boff, boff.data(), boff.size())) ||
          (rc = b->mem.upload(&M.bitem, bitem.data(), bitem.size())) ||
          (rc = b->mem.upload(&M.aoff, aoff.data(), aoff.size())) ||
          (rc = b->mem.upload(&M.aitem, aitem.data(), aitem.size())))
        return rc;
      b->nbblk = std::max(b->nbblk, (nbat + 255) / 256);
      const int nib = (S.nwat + S.nat_solute + 255) / 256;
      if ((rc = b->mem.alloc(&M.tpart, 4 * (size_t)std::max(nib, 1)))) return rc;
    }
    HIPCHK(hipMemcpy(M.v, v0[k], sizeof(double) * 3 * S.natom, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(M.st, 0, 8 * sizeof(double)));
    HIPCHK(hipMemset(M.ctr, 0, 4 * sizeof(int)));
    M.tfree = b->tfree + k;
    M.istep = b->istep + k;
    // the nonbonded kernels accumulate energies on the sampled steps only
    b->hs[k].step_ptr = M.istep;
    b->hs[k].ene_every = m->ene_every;
    b->ene_every.push_back(m->ene_every);
  }
  for (int k = 0; k < n; k++) {
    MdDev& M = b->hm[k];
    M.nbpart = b->nbblk;
    if ((rc = b->mem.alloc(&M.bpart, 2 * (size_t)b->nbblk))) return rc;
    HIPCHK(hipMemset(M.bpart, 0, sizeof(double) * 2 * b->nbblk));
  }
  if ((rc = b->mem.alloc(&b->dsys, n)) || (rc = b->mem.alloc(&b->dmd, n)) || (rc = b->mem.alloc(&b->dqb, n)))
    return rc;
  b->hqb.assign(n, QbDev{});
  b->qb_owned.assign(n, {});
  for (int k = 0; k < n; k++) {
    b->hqb[k].ns = b->hs[k].nstates;
    for (int st = 0; st < QNB_MAXST; st++) b->hqb[k].lambda[st] = b->hs[k].lambda[st];
  }
  HIPCHK(hipMemcpy(b->dqb, b->hqb.data(), sizeof(QbDev) * n, hipMemcpyHostToDevice));
  // the MD records change after this point only in qnb_batch_store_frames
  // and qnb_batch_set_offdiag (after a stream sync); every kernel that takes
  // b->dmd runs after this copy
  HIPCHK(hipMemcpy(b->dmd, b->hm.data(), sizeof(MdDev) * n, hipMemcpyHostToDevice));
  // first list build, regrown until every row fits, then rows sized to 1.5x
  // the largest neighbour count seen (re-checked at every update; an overflow
  // during a run is reported by qnb_batch_sync)
  for (int k = 0; k < n; k++) {
    DevSys& S = b->hs[k];
    b->mem.release(S.nbr);
    S.cap = std::min(((S.ncgp + 63) / 64) * 64, 1024);
    if ((rc = b->mem.alloc(&S.nbr, (size_t)S.cap * S.ncgp))) return rc;
  }
  for (int attempt = 0;; attempt++) {
    HIPCHK(hipMemcpy(b->dsys, b->hs.data(), sizeof(DevSys) * n, hipMemcpyHostToDevice));
    for (int k = 0; k < n; k++) HIPCHK(hipMemset(b->hs[k].need, 0, sizeof(int)));
    const int wpb = 4;
    launch_lists(b.get(), wpb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    bool again = false;
    for (int k = 0; k < n; k++) {
      DevSys& S = b->hs[k];
      int need = 0;
      HIPCHK(hipMemcpy(&need, S.need, sizeof(int), hipMemcpyDeviceToHost));
      if (need > S.cap) {
        b->mem.release(S.nbr);
        S.cap = ((need + 63) / 64) * 64;
        if ((rc = b->mem.alloc(&S.nbr, (size_t)S.cap * S.ncgp))) return rc;
        again = true;
      }
    }
    if (!again) break;
    if (attempt > 3) return fail(QNB_EOVERFLOW, "neighbour list capacity");
  }
  for (int k = 0; k < n; k++) {
    DevSys& S = b->hs[k];
    std::vector<int> nn(S.ncgp);
    HIPCHK(hipMemcpy(nn.data(), S.nnbr, sizeof(int) * S.ncgp, hipMemcpyDeviceToHost));
    int mx = 0;
    for (int v : nn) mx = std::max(mx, v);
    const int cap = ((mx * 3 / 2 + 64) + 63) / 64 * 64;
    if (cap < S.cap) {
      // shrink: rebuild happens at the first step of the run
      b->mem.release(S.nbr);
      S.cap = cap;
      if ((rc = b->mem.alloc(&S.nbr, (size_t)S.cap * S.ncgp))) return rc;
    }
    HIPCHK(hipMemset(S.need, 0, sizeof(int)));
  }
  HIPCHK(hipMemcpy(b->dsys, b->hs.data(), sizeof(DevSys) * n, hipMemcpyHostToDevice));
  {
    const int wpb = 4;
    launch_lists(b.get(), wpb);
  }
  if (b->anyq) hipLaunchKernelGGL(k_qlists, dim3(n), dim3(1024), 0, b->stream, b->dsys);
  // initial temperature (md_run calls temperature() before the loop)
  hipLaunchKernelGGL(k_temperature, dim3(n), dim3(256), 0, b->stream, b->dsys, b->dmd);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  *out = b.release();
  return 0;
}

// captures one list cycle (nbcycle steps, starting with the list update) as a
// graph, once; it is replayed until the launch list changes
static int capture_cycle(qnb_batch* b) {
  if (b->cycle_exec) return 0;
  const int nb = b->nbcycle[0];
  // a cycle starts at a multiple of nb: its sampled steps sit at fixed
  // offsets when every sampling interval divides nb
  bool fixed = true;
  for (int e : b->ene_every) fixed = fixed && e > 0 && nb % e == 0;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < nb; s++) enqueue(b, s == 0, fixed ? energy_step(b, s) : true);
  HIPCHK(hipStreamEndCapture(b->stream, &b->cycle_graph));
  HIPCHK(hipGraphInstantiate(&b->cycle_exec, b->cycle_graph, nullptr, nullptr, 0));
  return 0;
}

extern "C" int32_t qnb_batch_run(qnb_batch* b, int32_t nsteps, int32_t use_graph) {
  if (!b || nsteps < 0) return fail(QNB_EINVAL, "null batch or negative step count");
  const int nb = b->nbcycle[0];
  int rc;
  if (nsteps == 0) return use_graph ? capture_cycle(b) : 0;
  int done = 0;
  // steps up to the next list update, eagerly
  while (done < nsteps && (b->steps_done % nb) != 0) {
    enqueue(b, false, energy_step(b, b->steps_done));
    b->steps_done++;
    done++;
  }
  if (use_graph && nsteps - done >= nb) {
    if ((rc = capture_cycle(b))) return rc;
    while (nsteps - done >= nb) {
      HIPCHK(hipGraphLaunch(b->cycle_exec, b->stream));
      b->steps_done += nb;
      done += nb;
    }
  }
  while (done < nsteps) {
    enqueue(b, (b->steps_done % nb) == 0, energy_step(b, b->steps_done));
    b->steps_done++;
    done++;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int32_t qnb_batch_integrate(qnb_batch* b, const double* const* d) {
  if (!b || !d) return fail(QNB_EINVAL, "null batch or gradients");
  for (int k = 0; k < b->n; k++)
    if (!d[k]) return fail(QNB_EINVAL, "instance %d: null gradient", k);
  for (int k = 0; k < b->n; k++)
    HIPCHK(hipMemcpyAsync(b->hs[k].d, d[k], sizeof(double) * 3 * b->hs[k].natom, hipMemcpyHostToDevice, b->stream));
  launch_integrate(b);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  b->steps_done++;
  return 0;
}

extern "C" int32_t qnb_batch_thermostat(qnb_batch* b, int32_t k, double* out) {
  if (!b || k < 0 || k >= b->n || !out) return fail(QNB_EINVAL, "bad instance or output");
  HIPCHK(hipStreamSynchronize(b->stream));
  double st[8];
  HIPCHK(hipMemcpy(st, b->hm[k].st, sizeof st, hipMemcpyDeviceToHost));
  const int ng = b->hm[k].ngroup;
  out[0] = st[2];
  out[1] = st[1];
  out[2] = st[3];
  out[3] = st[4];
  out[4] = ng == 2 ? st[5] : 0.0;
  out[5] = st[0];
  out[6] = ng == 2 ? st[6] : 0.0;
  out[7] = ng;
  return 0;
}

extern "C" int32_t qnb_batch_sync(qnb_batch* b) {
  if (!b) return fail(QNB_EINVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < b->n; k++) {
    int ctr[4];
    HIPCHK(hipMemcpy(ctr, b->hm[k].ctr, sizeof ctr, hipMemcpyDeviceToHost));
    if (ctr[1] & ERR_OVERFLOW) return fail(QNB_EOVERFLOW, "instance %d: neighbour list overflow", k);
    if (ctr[1] & ERR_NONFINITE) return fail(QNB_ENONFINITE, "instance %d: non-finite energy or gradient", k);
    if (ctr[1] & ERR_SHAKE) return fail(QNB_ENONFINITE, "instance %d: shake failure", k);
  }
  return 0;
}

extern "C" int32_t qnb_batch_get_state(qnb_batch* b, int32_t k, double* x, double* v) {
  if (!b || k < 0 || k >= b->n) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  if (x) HIPCHK(hipMemcpy(x, b->hs[k].x, sizeof(double) * 3 * b->hs[k].natom, hipMemcpyDeviceToHost));
  if (v) HIPCHK(hipMemcpy(v, b->hm[k].v, sizeof(double) * 3 * b->hs[k].natom, hipMemcpyDeviceToHost));
  return 0;
}

// batch, after k_record on a sampled step: instance blockIdx.y, its frame of
// this step (the record just written: frame ctr[0] - 1 with this step's
// number); the state energies come from that record
__global__ void k_exc_record(const DevSys* __restrict__ systems, const MdDev* __restrict__ mds,
                             const ExcDev* __restrict__ excs) {
  const DevSys& S = systems[blockIdx.y];
  const MdDev& M = mds[blockIdx.y];
  const ExcDev& D = excs[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.narr * D.ns || M.ene_every <= 0) return;
  const int step = *M.istep;
  if (step % M.ene_every != 0) return;
  const int f = M.ctr[0] - 1;
  if (f < 0 || f >= D.max_frames) return;
  const double* h = M.hist + (size_t)f * QNB_FRAME_LEN;
  if (h[QNB_F_STEP] != (double)step) return;
  const int a = t / D.ns, st = t % D.ns;
  const double* q = h + QNB_F_Q0 + 6 * st;
  const double* b = h + QNB_F_QB0 + 4 * st;
  const double e[11] = {q[0], q[1], q[2], q[3], q[4], q[5], b[0], b[1], b[2], b[3], 0.0};
  exc_walk(D, S.x, a, st, e, D.hist + ((size_t)f * D.narr * D.ns + t) * QNB_GC_NV);
}

extern "C" int32_t qnb_batch_energy_history(qnb_batch* b, int32_t k, double* out, int32_t cap, int32_t* nframes) {
  if (!b || k < 0 || k >= b->n || !nframes) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  int ctr[4];
  HIPCHK(hipMemcpy(ctr, b->hm[k].ctr, sizeof ctr, hipMemcpyDeviceToHost));
  *nframes = ctr[0];
  if (out && ctr[0] <= cap && ctr[0] > 0)
    HIPCHK(hipMemcpy(out, b->hm[k].hist, sizeof(double) * QNB_FRAME_LEN * ctr[0], hipMemcpyDeviceToHost));
  return 0;
}

static void drop_step_graph(qnb_batch* b);

extern "C" int32_t qnb_batch_set_groups(qnb_batch* b, int32_t k, const qnb_system* sys, int32_t narr,
                                        const uint8_t* mask, const int32_t* type) {
  if (!b || k < 0 || k >= b->n || !sys) return fail(QNB_EINVAL, "bad instance or system");
  const DevSys& S = b->hs[k];
  if (sys->natom != S.natom || sys->nat_solute != S.nat_solute || sys->nqat != S.nqat || sys->nstates != S.nstates)
    return fail(QNB_EINVAL, "groups: not the system instance %d was created from", k);
  ExcHost X;
  if (int rc = exc_build(sys, narr, mask, type, X)) return rc;
  auto mem = std::make_unique<DevAlloc>();
  ExcDev D{};
  if (int rc = exc_upload(X, D, *mem)) return rc;
  D.max_frames = b->hm[k].max_frames;
  D.hist = nullptr;
  if (narr > 0) {
    const size_t nv = (size_t)std::max(D.max_frames, 1) * narr * S.nstates * QNB_GC_NV;
    if (int rc = mem->alloc(&D.hist, nv)) return rc;
    HIPCHK(hipMemset(D.hist, 0, sizeof(double) * nv));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->hexc.empty()) {
    b->hexc.assign(b->n, ExcDev{});
    b->exc_mem.resize(b->n);
    if (int rc = b->mem.alloc(&b->dexc, (size_t)b->n)) return rc;
  }
  b->hexc[k] = D;
  b->exc_mem[k] = std::move(mem);
  HIPCHK(hipMemcpy(b->dexc, b->hexc.data(), sizeof(ExcDev) * b->n, hipMemcpyHostToDevice));
  b->exc_threads = 0;
  for (const ExcDev& E : b->hexc) b->exc_threads = std::max(b->exc_threads, E.narr * E.ns);
  drop_step_graph(b);  // the step graph holds the launch list
  return 0;
}

extern "C" int32_t qnb_batch_group_energies(qnb_batch* b, int32_t k, double* out, int32_t cap, int32_t* nframes) {
  if (!b || k < 0 || k >= b->n || !nframes) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  int ctr[4];
  HIPCHK(hipMemcpy(ctr, b->hm[k].ctr, sizeof ctr, hipMemcpyDeviceToHost));
  *nframes = ctr[0];
  if (b->hexc.empty() || b->hexc[k].narr == 0) return 0;
  const ExcDev& D = b->hexc[k];
  if (out && ctr[0] <= cap && ctr[0] > 0)
    HIPCHK(hipMemcpy(out, D.hist, sizeof(double) * QNB_GC_NV * D.narr * D.ns * ctr[0], hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int32_t qnb_batch_set_offdiag(qnb_batch* b, int32_t k, int32_t n, const int32_t* kl, const double* A,
                                         const double* mu) {
  if (!b || k < 0 || k >= b->n) return fail(QNB_EINVAL, "bad instance");
  if (n < 0 || n > QNB_MAX_OFFD || (n > 0 && (!kl || !A || !mu)))
    return fail(QNB_EINVAL, "off-diagonals: 0..%d elements with pairs, A and mu", QNB_MAX_OFFD);
  const DevSys& S = b->hs[k];
  const HostTopo* T = b->topos[b->topo_of[k]].get();
  for (int m = 0; m < 2 * n; m++)
    if (kl[m] < 0 || kl[m] >= S.nqat) return fail(QNB_EINVAL, "off-diagonal Q-atom %d out of range", kl[m]);
  HIPCHK(hipStreamSynchronize(b->stream));
  MdDev& M = b->hm[k];
  M.noffd = n;
  for (int m = 0; m < n; m++) {
    M.offd_k[m] = T->iqseq[kl[2 * m]];
    M.offd_l[m] = T->iqseq[kl[2 * m + 1]];
    M.offd_A[m] = A[m];
    M.offd_mu[m] = mu[m];
  }
  HIPCHK(hipMemcpy(b->dmd, b->hm.data(), sizeof(MdDev) * b->n, hipMemcpyHostToDevice));
  return 0;
}

// the captured step graph embeds the launch list and shapes: a change to
// either drops it, and the next run recaptures
static void drop_step_graph(qnb_batch* b) {
  if (b->cycle_exec) {
    (void)hipGraphExecDestroy(b->cycle_exec);
    b->cycle_exec = nullptr;
  }
  if (b->cycle_graph) {
    (void)hipGraphDestroy(b->cycle_graph);
    b->cycle_graph = nullptr;
  }
}

extern "C" int32_t qnb_batch_set_qbonded(qnb_batch* b, int32_t k, const qnb_qbonded* qb) {
  if (!b || k < 0 || k >= b->n) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  const DevSys& S = b->hs[k];
  qbonded_free(b->mem, b->qb_owned[k]);
  int rc = qbonded_build(qb, S.nstates, S.natom, S.lambda, b->mem, b->qb_owned[k], b->hqb[k]);
  if (rc) {
    qbonded_free(b->mem, b->qb_owned[k]);
    b->hqb[k] = QbDev{};
    b->hqb[k].ns = S.nstates;
  }
  const QbDev& Q = b->hqb[k];
  b->hm[k].qbe = Q.nb + Q.na + Q.nt + Q.ni > 0 ? Q.e : nullptr;
  b->anyqb = false;
  for (const QbDev& q : b->hqb) b->anyqb = b->anyqb || q.nb + q.na + q.nt + q.ni > 0;
  HIPCHK(hipMemcpy(b->dqb, b->hqb.data(), sizeof(QbDev) * b->n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->dmd, b->hm.data(), sizeof(MdDev) * b->n, hipMemcpyHostToDevice));
  // the step graph holds the launch list: recapture with (or without) k_qbonded
  drop_step_graph(b);
  return rc;
}

extern "C" int32_t qnb_batch_store_frames(qnb_batch* b, int32_t every, int32_t max_frames) {
  if (!b || every < 0 || max_frames < 0) return fail(QNB_EINVAL, "bad frame-store arguments");
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < b->n; k++) {
    MdDev& M = b->hm[k];
    if (M.xs) {
      b->mem.release(M.xs);
      M.xs = nullptr;
    }
    M.xs_every = 0;
    M.xs_max = 0;
    if (every > 0 && max_frames > 0) {
      int rc = b->mem.alloc(&M.xs, (size_t)max_frames * 3 * b->hs[k].natom);
      if (rc) return rc;
      M.xs_every = every;
      M.xs_max = max_frames;
    }
    int zero = 0;
    HIPCHK(hipMemcpy(M.ctr + 2, &zero, sizeof zero, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(b->dmd, b->hm.data(), sizeof(MdDev) * b->n, hipMemcpyHostToDevice));
  // k_snap is launched only while snapshots are stored
  const bool snap = every > 0 && max_frames > 0;
  if (snap != b->anysnap) {
    b->anysnap = snap;
    drop_step_graph(b);
  }
  return 0;
}

static int frames_stored(qnb_batch* b, int k, int* nf) {
  int ctr[4];
  HIPCHK(hipMemcpy(ctr, b->hm[k].ctr, sizeof ctr, hipMemcpyDeviceToHost));
  *nf = b->hm[k].xs ? ctr[2] : 0;
  return 0;
}

extern "C" int32_t qnb_batch_frames(qnb_batch* b, int32_t k, double* out, int32_t cap, int32_t* nframes) {
  if (!b || k < 0 || k >= b->n || !nframes) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  int nf = 0, rc = frames_stored(b, k, &nf);
  if (rc) return rc;
  *nframes = nf;
  if (out && nf > 0 && nf <= cap)
    HIPCHK(hipMemcpy(out, b->hm[k].xs, sizeof(double) * 3 * b->hs[k].natom * nf, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int32_t qnb_batch_adjoint(qnb_batch* b, const qnb_qtangent* const* tan, double* out, int64_t cap,
                                     int64_t* n) {
  if (!b || !tan || !n) return fail(QNB_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<int> nf(b->n);
  std::vector<size_t> off(b->n + 1, 0);
  long waves = 0;
  for (int k = 0; k < b->n; k++) {
    int rc = frames_stored(b, k, &nf[k]);
    if (rc) return rc;
    const DevSys& S = b->hs[k];
    off[k + 1] = off[k] + (size_t)nf[k] * S.nstates * S.nqat * QNB_ADJ_NV;
    waves = std::max(waves, (long)nf[k] * S.nqat);
  }
  *n = (int64_t)off[b->n];
  if (!out || *n == 0 || *n > cap) return 0;
  DevAlloc tmp;
  std::vector<AdjDev> A(b->n);
  double* dout = nullptr;
  int rc = tmp.alloc(&dout, off[b->n]);
  if (rc) return rc;
  for (int k = 0; k < b->n; k++) {
    int same = -1;
    for (int m = 0; m < k && same < 0; m++)
      if (tan[m] == tan[k]) same = m;
    if (same >= 0) {
      if (b->topo_of[same] != b->topo_of[k]) return fail(QNB_EINVAL, "instances %d and %d share tangents but not a system", same, k);
      A[k] = A[same];
    } else if (b->hs[k].nqat > 0 && (rc = adjoint_upload(tmp, b->hs[k], tan[k], A[k]))) {
      return rc;
    }
    A[k].frames = b->hm[k].xs;
    A[k].nframes = nf[k];
    A[k].out = dout + off[k];
  }
  AdjDev* dadj = nullptr;
  if ((rc = tmp.upload((const AdjDev**)&dadj, A.data(), A.size()))) return rc;
  launch_adjoint(b->stream, b->prec, b->ns_max, waves, b->n, b->dsys, dadj);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * off[b->n], hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

extern "C" int32_t qnb_batch_pair_counts(qnb_batch* b, int32_t k, int64_t* out) {
  if (!b || k < 0 || k >= b->n || !out) return fail(QNB_EINVAL, "bad instance");
  HIPCHK(hipStreamSynchronize(b->stream));
  const DevSys& S = b->hs[k];
  const HostTopo* T = b->topos[b->topo_of[k]].get();
  std::vector<int> nn(S.ncgp), nbr((size_t)S.cap * S.ncgp);
  HIPCHK(hipMemcpy(nn.data(), S.nnbr, sizeof(int) * S.ncgp, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nbr.data(), S.nbr, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
  int nql[2];
  HIPCHK(hipMemcpy(nql, S.nq_lists, sizeof nql, hipMemcpyDeviceToHost));
  // atom pairs of a group pair, filtered as the list builders filter them
  auto pp_pairs = [&](int ig, int jg) {
    int64_t c = 0;
    for (int ia = T->gfirst[ig]; ia <= T->glast[ig]; ia++) {
      const int i = T->cgpatom[ia];
      if (T->iqatom[i]) continue;
      for (int ja = T->gfirst[jg]; ja <= T->glast[jg]; ja++) {
        const int j = T->cgpatom[ja];
        if (T->iqatom[j] || (ig == jg && i >= j)) continue;
        c += host_pp_param(*T, i, j);
      }
    }
    return c;
  };
  auto nonq = [&](int g) {
    int64_t c = 0;
    for (int ia = T->gfirst[g]; ia <= T->glast[g]; ia++) c += T->iqatom[T->cgpatom[ia]] == 0;
    return c;
  };
  int64_t pp = 0, pw = 0, ww = 0;
  const int ns = S.ncgp_solute;
  for (int a = 0; a < S.ncgp; a++) {
    if (a < ns && T->excl[T->gswitch[a]]) continue;
    for (int m = 0; m < std::min(nn[a], S.cap); m++) {
      const int g = nbr[(size_t)a * S.cap + m];
      const bool own = (a < g && ((a + g) & 1) == 0) || (a > g && ((a + g) & 1) == 1);
      if (a < ns && g < ns) {
        if (own) pp += pp_pairs(a, g);
      } else if (a < ns) {
        pw += nonq(a) * S.sa;
      } else if (g >= ns && own) {
        ww += (int64_t)S.sa * S.sa;
      }
    }
    if (a < ns) pp += pp_pairs(a, a);
  }
  out[0] = pp;
  out[1] = pw;
  out[2] = ww;
  out[3] = (int64_t)nql[1] * S.nqat;
  out[4] = (int64_t)nql[0] * S.nqat * S.sa;
  out[5] = S.nqq;
  out[6] = S.nqqp;
  return 0;
}

extern "C" int32_t qnb_batch_time_kernel(qnb_batch* b, int32_t which, int32_t reps, double* ms_per_launch) {
  if (!b || reps <= 0 || !ms_per_launch || which < 0 || which > 4) return fail(QNB_EINVAL, "bad arguments");
  if (which == 4 && !b->anyq) return fail(QNB_EINVAL, "no instance has Q-atoms");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto launch = [&]() {
    switch (which) {
      case 0: launch_lists(b, b->wpb_lists); break;
      case 1: launch_forces(b, b->wpb_forces, false); break;
      case 2: launch_forces(b, b->wpb_forces, false, false, 1); break;
      case 3: launch_forces(b, b->wpb_forces, false, false, 2); break;
      default: launch_qforces(b); break;
    }
  };
  launch();  // warm-up
  HIPCHK(hipEventRecord(e0, b->stream));
  for (int r = 0; r < reps; r++) launch();
  HIPCHK(hipEventRecord(e1, b->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_launch = (double)ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

extern "C" int32_t qnb_batch_free_energy(qnb_batch* b, const qnb_fep_params* p, qnb_fep_out* out) {
  if (!b || !out) return fail(QNB_EINVAL, "bad arguments");
  const int ns = b->hs[0].nstates;
  for (const DevSys& S : b->hs)
    if (S.nstates != ns) return fail(QNB_EINVAL, "instances differ in their number of states");
  FepPar P;
  if (int rc = fep_params(p, ns, &P)) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<FepWin> hw(b->n);
  for (int k = 0; k < b->n; k++) {
    int nf = 0;
    HIPCHK(hipMemcpy(&nf, b->hm[k].ctr, sizeof(int), hipMemcpyDeviceToHost));
    nf = std::min(nf, b->hm[k].max_frames);
    if (nf <= P.nskip) return fail(QNB_EINVAL, "instance %d has %d frames, not more than the %d skipped", k, nf, P.nskip);
    FepWin& W = hw[k];
    W.hist = b->hm[k].hist;
    W.V = nullptr;
    W.rkl = nullptr;
    W.nf = nf;
    W.noffd = b->hm[k].noffd;
    for (int s = 0; s < QNB_MAX_STATES; s++) W.lam[s] = b->hs[k].lambda[s];
  }
  return fep_run(b->stream, hw, P, out);
}

extern "C" int32_t qnb_batch_minimize(qnb_batch* b, int32_t nsteps, double max_disp) {
  if (!b || nsteps < 0 || !(max_disp > 0.0)) return fail(QNB_EINVAL, "bad arguments");
  const double k0 = 1.0e-3;  // A per kcal/mol/A
  for (int it = 0; it < nsteps; it++) {
    const bool lists = (it % b->nbcycle[0]) == 0;
    if (b->prec == QNB_PREC_MIXED) enqueue_min_step<float>(b, lists, k0, max_disp);
    else enqueue_min_step<double>(b, lists, k0, max_disp);
  }
  // lists for the relaxed positions before MD continues
  launch_lists(b, b->wpb_lists);
  if (b->anyq && !b->static_q) hipLaunchKernelGGL(k_qlists, dim3(b->n), dim3(1024), 0, b->stream, b->dsys);
  HIPCHK(hipGetLastError());
  return 0;
}

// Autotuner: times the list/LRF sweep and the force kernel on the live batch
// for 1, 2, 4 and 8 receiver groups (wavefronts) per block and keeps the
// fastest shape of each.  Both kernels recompute their outputs from the
// current coordinates, so the trials leave the state unchanged.
extern "C" int32_t qnb_batch_autotune(qnb_batch* b, int32_t reps, int32_t* wpb_out) {
  if (!b || reps <= 0) return fail(QNB_EINVAL, "bad arguments");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int cand[4] = {1, 2, 4, 8};
  for (int which = 0; which < 2; which++) {
    float best = 1e30f;
    int pick = 4, pick_map = b->imap_water | (b->imap_solute << 1);
    // the force kernels also try both block placements (map_block) of the
    // water receivers (bit 0) and of the solute receivers (bit 1)
    for (int imap = 0; imap < (which == 0 ? 1 : 4); imap++) {
      if (which == 1) {
        b->imap_water = imap & 1;
        b->imap_solute = imap >> 1;
      }
      for (int c : cand) {
        if (which == 0) launch_lists(b, c);
        else launch_forces(b, c, false);
        HIPCHK(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; r++) {
          if (which == 0) launch_lists(b, c);
          else launch_forces(b, c, false);
        }
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          pick = c;
          pick_map = imap;
        }
      }
    }
    if (which == 0) b->wpb_lists = pick;
    else {
      b->wpb_forces = pick;
      b->imap_water = pick_map & 1;
      b->imap_solute = pick_map >> 1;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the step graph embeds launch shapes: drop it so the next run recaptures
  if (b->cycle_exec) {
    (void)hipGraphExecDestroy(b->cycle_exec);
    b->cycle_exec = nullptr;
  }
  if (b->cycle_graph) {
    (void)hipGraphDestroy(b->cycle_graph);
    b->cycle_graph = nullptr;
  }
  if (wpb_out) {
    wpb_out[0] = b->wpb_lists;
    wpb_out[1] = b->wpb_forces;
  }
  return 0;
}
