// qnb_fep.hip -- Qfep6's free-energy analysis on the device, over the energy
// frames a batch has recorded (one window per instance, in instance order):
//
//  * Part 1, free-energy perturbation (src/qfep.f90:587-648): per window i the
//    forward average dgf_i = -kT ln <exp(-(V.lam_{i+1} - V.lam_i)/kT)>_i and the
//    reverse dgr_i towards window i-1, then dG_i = dG_{i-1} + (dgf_{i-1} -
//    dgr_i)/2;
//  * Part 2, reaction free energy per window and energy-gap bin (:700-795):
//    the EVB ground state E_g of two states (:492-500), the gap sum_s c_s V_s
//    binned between the smallest and largest gap of all windows, per bin the
//    averages of E_g - eps and V_s - eps (eps = V.lam_i) and the exponential
//    averages about them, dGg = dG_i - kT ln <exp(-(E_g - eps - av)/kT)> + av;
//  * Part 3 (:797-827): per bin the point-weighted averages over windows.
//
// State energies are formed from each frame as pot_energy forms EQ(s)%total
// (src/potene.f90:224-231: Q bonded terms, then qx = qq + qp + qw, el and
// vdW), plus the alpha shifts (qfep.f90:447-452).  Every sum runs over the
// frames in order, as the reference's loops do (one thread per sum), with no
// contraction of multiply-add pairs, so the results differ from Qfep6's only
// by the last bits of exp and log.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "qnb_host.h"

#pragma clang fp contract(off)

namespace qnb {

// V_s of frame f (alpha applied), the ground state, c1 and the gap
__device__ void fep_frame(const FepWin& W, const FepPar& P, int f, double* v, double& vg, double& c1, double& gap,
                          double& r) {
  if (W.V) {
    for (int s = 0; s < P.ns; s++) v[s] = W.V[(size_t)f * P.ns + s] + (s == 0 ? 0.0 : P.alpha[s]);
    r = W.noffd > 0 ? W.rkl[f] : 0.0;
  } else {
    const double* h = W.hist + (size_t)f * QNB_FRAME_LEN;
    for (int s = 0; s < P.ns; s++) {
      const double* q = h + QNB_F_Q0 + 6 * s;
      const double* b = h + QNB_F_QB0 + 4 * s;
      const double qxel = q[0] + q[2] + q[4];
      const double qxvdw = q[1] + q[3] + q[5];
      const double total = b[0] + b[1] + b[2] + b[3] + qxel + qxvdw + 0.0;  // + EQ(s)%restraint (none)
      v[s] = total + (s == 0 ? 0.0 : P.alpha[s]);
    }
    r = W.noffd > 0 ? h[QNB_F_OFFD + 1] : 0.0;
  }
  vg = 0.0;
  c1 = 0.0;
  if (P.ns == 2) {
    const double dr = r - P.r0;
    const double H = P.A * exp(-P.mu * dr) * exp(-P.eta * (dr * dr));
    vg = 0.5 * (v[0] + v[1]) - 0.5 * sqrt((v[0] - v[1]) * (v[0] - v[1]) + 4.0 * (H * H));
    if (W.noffd > 0) {
      const double t = (vg - v[0]) / H;
      c1 = 1.0 / (1.0 + t * t);
    }
  }
  gap = 0.0;
  for (int s = 0; s < P.ns; s++) gap = gap + v[s] * P.coeff[s];
}

__device__ double veff_of(const double* lam, const double* v, int ns) {
  double e = 0.0;
  for (int s = 0; s < ns; s++) e = e + lam[s] * v[s];
  return e;
}

// Part 1 sums and the gap range of window blockIdx.x: thread 0 the forward
// average, thread 1 the reverse one, all threads the gap minimum and maximum
__global__ __launch_bounds__(64) void k_fep_window(const FepWin* __restrict__ wins, int n, FepPar P, double* dgf,
                                                   double* dgr, double* wmin, double* wmax) {
  const int i = blockIdx.x;
  const FepWin& W = wins[i];
  double v[QNB_MAX_STATES], vg, c1, gap, r;
  const int t = threadIdx.x;
  if (t < 2) {
    const int o = t == 0 ? i + 1 : i - 1;
    double out = NAN;
    if (o >= 0 && o < n) {
      double sum = 0.0;
      for (int f = P.nskip; f < W.nf; f++) {
        fep_frame(W, P, f, v, vg, c1, gap, r);
        const double veff1 = veff_of(W.lam, v, P.ns);
        const double veff2 = veff_of(wins[o].lam, v, P.ns);
        const double dv = veff2 - veff1;
        sum = sum + exp(-dv / P.kT);
      }
      sum = sum / (double)(W.nf - P.nskip);
      out = -P.kT * log(sum);
    }
    (t == 0 ? dgf : dgr)[i] = out;
  }
  double lo = INFINITY, hi = -INFINITY;
  for (int f = P.nskip + t; f < W.nf; f += blockDim.x) {
    fep_frame(W, P, f, v, vg, c1, gap, r);
    lo = fmin(lo, gap);
    hi = fmax(hi, gap);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off));
    hi = fmax(hi, __shfl_xor(hi, off));
  }
  if (t == 0) {
    wmin[i] = lo;
    wmax[i] = hi;
  }
}

// dG chain and the overall gap range (sequential over windows)
__global__ void k_fep_chain(int n, const double* dgf, const double* dgr, const double* wmin, const double* wmax,
                            double* dG, double* range) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  dG[0] = 0.0;
  for (int i = 1; i < n; i++) dG[i] = dG[i - 1] + 0.5 * (dgf[i - 1] - dgr[i]);
  double lo = wmin[0], hi = wmax[0];
  for (int i = 1; i < n; i++) {
    lo = fmin(lo, wmin[i]);
    hi = fmax(hi, wmax[i]);
  }
  range[0] = lo;
  range[1] = hi;
}

// Part 2 of window blockIdx.x: thread t owns bins t, t + blockDim.x, ...
// and takes the window's frames in order for each of its bins.  Per (window,
// bin) rec[] holds dGg, c1, c2, r, then dGv of each state; npts the points.
#define FEP_REC (4 + QNB_MAX_STATES)
__global__ __launch_bounds__(256) void k_fep_bins(const FepWin* __restrict__ wins, FepPar P, const double* dG,
                                                  const double* range, double* rec, int* npts) {
  const int i = blockIdx.x;
  const FepWin& W = wins[i];
  const int ns = P.ns;
  const double gapmin = range[0];
  const double xint = (range[1] - range[0]) / (double)P.nbins;
  double v[QNB_MAX_STATES], vg, c1, gap, r;
  for (int k = threadIdx.x; k < P.nbins; k += blockDim.x) {
    int cnt = 0;
    double avdvv[QNB_MAX_STATES], sumv[QNB_MAX_STATES];
    for (int s = 0; s < QNB_MAX_STATES; s++) avdvv[s] = sumv[s] = 0.0;
    double avdvg = 0.0, avc1 = 0.0, avc2 = 0.0, avr = 0.0, sumg = 0.0;
    for (int f = P.nskip; f < W.nf; f++) {
      fep_frame(W, P, f, v, vg, c1, gap, r);
      if ((int)((gap - gapmin) / xint) != k) continue;
      const double veff = veff_of(W.lam, v, ns);
      for (int s = 0; s < ns; s++) avdvv[s] = avdvv[s] + (v[s] - veff);
      avdvg = avdvg + (vg - veff);
      avc1 = avc1 + c1;
      avc2 = avc2 + (W.noffd > 0 ? 1.0 - c1 : 0.0);
      avr = avr + r;
      cnt++;
    }
    if (cnt != 0) {
      avc1 = avc1 / (double)cnt;
      avc2 = avc2 / (double)cnt;
      avr = avr / (double)cnt;
      for (int s = 0; s < ns; s++) avdvv[s] = avdvv[s] / (double)cnt;
      avdvg = avdvg / (double)cnt;
    }
    for (int f = P.nskip; f < W.nf && cnt > 0; f++) {
      fep_frame(W, P, f, v, vg, c1, gap, r);
      if ((int)((gap - gapmin) / xint) != k) continue;
      const double veff = veff_of(W.lam, v, ns);
      for (int s = 0; s < ns; s++) sumv[s] = sumv[s] + exp(-(v[s] - veff - avdvv[s]) / P.kT);
      sumg = sumg + exp(-(vg - veff - avdvg) / P.kT);
    }
    double* o = rec + ((size_t)i * P.nbins + k) * FEP_REC;
    npts[(size_t)i * P.nbins + k] = cnt;
    if (cnt > 0 && cnt >= P.nmin) {
      for (int s = 0; s < ns; s++) o[4 + s] = dG[i] + (-P.kT * log(sumv[s] / (double)cnt) + avdvv[s]);
      o[0] = dG[i] + (-P.kT * log(sumg / (double)cnt) + avdvg);
      o[1] = avc1;
      o[2] = avc2;
      o[3] = avr;
    } else {
      for (int z = 0; z < FEP_REC; z++) o[z] = NAN;
    }
  }
}

// Part 3: per bin, the windows in order
__global__ __launch_bounds__(256) void k_fep_part3(int n, FepPar P, const double* rec, const int* npts, double* bins,
                                                   int* ptsum) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P.nbins; k += gridDim.x * blockDim.x) {
    double b1 = 0.0, b2 = 0.0, b3 = 0.0, b4 = 0.0;
    int pt = 0;
    for (int i = 0; i < n; i++) {
      const int c = npts[(size_t)i * P.nbins + k];
      if (c == 0 || c < P.nmin) continue;
      const double* o = rec + ((size_t)i * P.nbins + k) * FEP_REC;
      b2 = b2 + o[1] * (double)c;
      b3 = b3 + o[2] * (double)c;
      b4 = b4 + o[3] * (double)c;
      pt += c;
      b1 = b1 + o[0] * (double)c;
    }
    ptsum[k] = pt;
    const bool ok = pt > 0 && pt >= P.nmin;
    bins[4 * k + 0] = ok ? b1 / (double)pt : NAN;
    bins[4 * k + 1] = ok ? b2 / (double)pt : NAN;
    bins[4 * k + 2] = ok ? b3 / (double)pt : NAN;
    bins[4 * k + 3] = ok ? b4 / (double)pt : NAN;
  }
}

// Scratch layout: one device buffer of fep_scratch_doubles(n, nbins)
// doubles holds every intermediate and output, at the offsets fep_layout
// gives (the last region holds the int32 point counts).
void fep_layout(int n, int nbins, size_t off[8]) {
  off[0] = 0;                                  // dgf [n]
  off[1] = off[0] + n;                         // dgr [n]
  off[2] = off[1] + n;                         // dG [n]
  off[3] = off[2] + n;                         // wmin, wmax [2n]
  off[4] = off[3] + 2 * (size_t)n;             // range [2]
  off[5] = off[4] + 2;                         // rec [n][nbins][FEP_REC]
  off[6] = off[5] + (size_t)n * nbins * FEP_REC;  // bins [nbins][4]
  off[7] = off[6] + 4 * (size_t)nbins;         // ints follow: npts [n][nbins], ptsum [nbins]
}
size_t fep_scratch_doubles(int n, int nbins) {
  size_t off[8];
  fep_layout(n, nbins, off);
  return off[7] + ((size_t)n * nbins + nbins + 1) / 2 + 1;
}
size_t fep_rec_len() { return FEP_REC; }

// Part 1 and the gap range
void launch_fep_part1(hipStream_t st, const FepWin* wins, int n, const FepPar& P, double* scratch) {
  size_t off[8];
  fep_layout(n, P.nbins, off);
  double* sd = scratch;
  hipLaunchKernelGGL(k_fep_window, dim3(n), dim3(64), 0, st, wins, n, P, sd + off[0], sd + off[1], sd + off[3],
                     sd + off[3] + n);
  hipLaunchKernelGGL(k_fep_chain, dim3(1), dim3(64), 0, st, n, sd + off[0], sd + off[1], sd + off[3], sd + off[3] + n,
                     sd + off[2], sd + off[4]);
}

// Parts 2 and 3 (P.nbins > 0, a non-empty gap range)
void launch_fep_bins(hipStream_t st, const FepWin* wins, int n, const FepPar& P, double* scratch) {
  size_t off[8];
  fep_layout(n, P.nbins, off);
  double* sd = scratch;
  int* npts = reinterpret_cast<int*>(sd + off[7]);
  hipLaunchKernelGGL(k_fep_bins, dim3(n), dim3(256), 0, st, wins, P, sd + off[2], sd + off[4], sd + off[5], npts);
  hipLaunchKernelGGL(k_fep_part3, dim3((P.nbins + 255) / 256), dim3(256), 0, st, n, P, sd + off[5], npts, sd + off[6],
                     npts + (size_t)n * P.nbins);
}

// Runs the analysis for windows whose data already sit in device memory
// (hw: host copies of the window records) and copies the results out.
int fep_run(hipStream_t st, const std::vector<FepWin>& hw, const FepPar& P, qnb_fep_out* out) {
  const int n = (int)hw.size();
  DevAlloc mem;
  const FepWin* dw = nullptr;
  double* sd = nullptr;
  if (int rc = mem.upload(&dw, hw.data(), hw.size())) return rc;
  if (int rc = mem.alloc(&sd, fep_scratch_doubles(n, P.nbins))) return rc;
  size_t off[8];
  fep_layout(n, P.nbins, off);
  launch_fep_part1(st, dw, n, P, sd);
  HIPCHK(hipGetLastError());
  double range[2];
  HIPCHK(hipMemcpyAsync(range, sd + off[4], sizeof(range), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (P.nbins > 0) {
    if (!(range[1] > range[0])) return fail(QNB_EINVAL, "energy gap range is empty: no bins to fill");
    launch_fep_bins(st, dw, n, P, sd);
    HIPCHK(hipGetLastError());
  }
  auto get = [&](void* dst, const void* src, size_t bytes) -> int {
    if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    return 0;
  };
  int rc = 0;
  rc = rc ? rc : get(out->dgf, sd + off[0], n * sizeof(double));
  rc = rc ? rc : get(out->dgr, sd + off[1], n * sizeof(double));
  rc = rc ? rc : get(out->dG, sd + off[2], n * sizeof(double));
  rc = rc ? rc : get(out->gap_range, sd + off[4], 2 * sizeof(double));
  std::vector<double> rec, bins;
  std::vector<int32_t> ints;
  const size_t nb = (size_t)n * P.nbins;
  if (P.nbins > 0 && !rc) {
    rec.resize(nb * FEP_REC);
    bins.resize(4 * (size_t)P.nbins);
    ints.resize(nb + P.nbins);
    rc = rc ? rc : get(rec.data(), sd + off[5], rec.size() * sizeof(double));
    rc = rc ? rc : get(bins.data(), sd + off[6], bins.size() * sizeof(double));
    rc = rc ? rc : get(ints.data(), sd + off[7], ints.size() * sizeof(int32_t));
  }
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(st));
  if (P.nbins > 0) {
    for (size_t w = 0; w < nb; w++) {
      const double* o = rec.data() + w * FEP_REC;
      if (out->dGg) out->dGg[w] = o[0];
      if (out->c1) out->c1[w] = o[1];
      if (out->c2) out->c2[w] = o[2];
      if (out->r) out->r[w] = o[3];
      if (out->dGv)
        for (int s2 = 0; s2 < P.ns; s2++) out->dGv[w * P.ns + s2] = o[4 + s2];
      if (out->npts) out->npts[w] = ints[w];
    }
    for (int k = 0; k < P.nbins; k++) {
      if (out->bin_dGg) out->bin_dGg[k] = bins[4 * k];
      if (out->bin_c1) out->bin_c1[k] = bins[4 * k + 1];
      if (out->bin_c2) out->bin_c2[k] = bins[4 * k + 2];
      if (out->bin_r) out->bin_r[k] = bins[4 * k + 3];
      if (out->ptsum) out->ptsum[k] = ints[nb + k];
    }
  }
  return 0;
}

int fep_params(const qnb_fep_params* p, int ns, FepPar* P) {
  if (!p) return fail(QNB_EINVAL, "no parameters");
  if (!(p->kT > 0.0)) return fail(QNB_EINVAL, "kT must be positive");
  if (ns < 1 || ns > QNB_MAX_STATES) return fail(QNB_EINVAL, "nstates must be 1..%d", QNB_MAX_STATES);
  if (p->nskip < 0 || p->nbins < 0 || p->nmin < 0) return fail(QNB_EINVAL, "nskip, nbins and nmin must be >= 0");
  if (p->nbins > 0 && ns != 2)
    return fail(QNB_EINVAL, "the reaction free energy (nbins > 0) needs the 2-state EVB ground state");
  P->kT = p->kT;
  P->nskip = p->nskip;
  P->nbins = p->nbins;
  P->nmin = p->nmin;
  P->ns = ns;
  for (int s = 0; s < QNB_MAX_STATES; s++) {
    P->alpha[s] = s == 0 ? 0.0 : p->alpha[s];  // alfa(1) = zero (qfep.f90:447)
    P->coeff[s] = p->coeff[s];
  }
  P->A = p->A;
  P->mu = p->mu;
  P->eta = p->eta;
  P->r0 = p->r0;
  return 0;
}

}  // namespace qnb

using namespace qnb;

extern "C" int32_t qnb_free_energy(int32_t nwin, int32_t nstates, const int32_t* nframes, const double* const* V,
                                   const double* const* rkl, const double* lambdas, const qnb_fep_params* p,
                                   qnb_fep_out* out) {
  if (nwin < 1 || !nframes || !V || !lambdas || !out) return fail(QNB_EINVAL, "bad arguments");
  FepPar P;
  if (int rc = fep_params(p, nstates, &P)) return rc;
  DevAlloc mem;
  std::vector<FepWin> hw(nwin);
  for (int w = 0; w < nwin; w++) {
    if (!V[w] || nframes[w] <= P.nskip)
      return fail(QNB_EINVAL, "window %d has no frames after the %d skipped", w, P.nskip);
    FepWin& W = hw[w];
    W.hist = nullptr;
    W.nf = nframes[w];
    W.noffd = (rkl && rkl[w]) ? 1 : 0;
    const double* dv = nullptr;
    if (int rc = mem.upload(&dv, V[w], (size_t)nframes[w] * nstates)) return rc;
    W.V = dv;
    W.rkl = nullptr;
    if (W.noffd) {
      const double* dr = nullptr;
      if (int rc = mem.upload(&dr, rkl[w], (size_t)nframes[w])) return rc;
      W.rkl = dr;
    }
    for (int s = 0; s < QNB_MAX_STATES; s++) W.lam[s] = s < nstates ? lambdas[(size_t)w * nstates + s] : 0.0;
  }
  return fep_run(nullptr, hw, P, out);
}
